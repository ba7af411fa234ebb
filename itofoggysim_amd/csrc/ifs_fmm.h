// MLFMM plan and passes (ifs_fmm.cpp).  Internal.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct ifs_ctx;

namespace ifs {
namespace fmm {

struct Fmm;

// Builds the octree, the level hierarchy and the translation batches of one
// cluster (lib_mie_ms_ml_fmm_constructor -> lib_ml_fmm_constructor,
// FoggySim/src/lib/mie_theory/lib_mie_ms_ml_fmm_interface.f90:46-95,
// libmlfmm/src/lib/lib_ml_fmm.f90:77-155).
int build(ifs_ctx *c, Fmm **out);
void destroy(Fmm *f);
// acc += (far field via P2M, M2M, M2L, L2L, L2P) + (near field P2P) of xh
// (libmlfmm/src/lib/lib_ml_fmm.f90:231-262).
// prec: precision of the class products (see launch_class).
int run(ifs_ctx *c, Fmm *f, const double2 *xh, double2 *acc, int nrhs, int64_t ld, int prec = 0);
int tree_export(Fmm *f, int64_t *uindex_out, int32_t *levels_out);
// Hash of the device buffers and plan shapes a matvec's launches refer to.
uint64_t buffers_sig(Fmm *f);
int stats(Fmm *f, int64_t counts[6], int32_t levels_L[32]);

}  // namespace fmm
}  // namespace ifs
