// Translation engine and context (internal): shared by ifs_api.cpp, ifs_fmm.cpp
// and ifs_solver.cpp.
#pragma once
#include <hip/hip_runtime.h>

#include <complex>
#include <map>
#include <string>
#include <vector>

#include "../../include/ifs.h"
#include "ifs_internal.h"

namespace ifs {
namespace fmm {
struct Fmm;
}
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t b) {
        if (p && bytes >= b) return IFS_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (b == 0) return IFS_OK;
        if (hipMalloc(&p, b) != hipSuccess) {
            p = nullptr;
            return IFS_E_NOMEM;
        }
        bytes = b;
        return IFS_OK;
    }
    template <class T> T *as() const { return (T *)p; }
};

// Chunk of a translation batch that fits the workspace.
struct Chunk {
    int t_begin = 0, t_end = 0;      // task range (device task array)
    int s_begin = 0, s_end = 0;      // segment range
    int64_t n_d = 0, n_s = 0, n_ax = 0, n_y = 0, n_z = 0;
    int maxpairs = 0, maxK = 0, max_nmax = 0, max_Ls = 0, max_Lt = 0;
    int64_t max_rows_y = 0, max_rows_z = 0;
    int rowback_blocks = 1;
    double f_fwd = 0.0, f_ax = 0.0, f_back = 0.0;  // algorithmic flops of the chunk's products
    GqbChunk gqb;                    // quadrature tables (batches of regular translations)
};

// A translation batch: tasks from one source pool into one target pool.
struct Batch {
    std::vector<Task> tasks;     // sorted by target
    std::vector<Segment> segs;
    std::vector<Chunk> chunks;
    DevBuf d_tasks, d_segs;
    DevBuf own_d, own_s, own_ax, own_y, own_z;  // workspace of a cached batch
    bool cached = false;         // tables generated once and kept
    bool tables_ready = false;
    // regular (j_n) translations between spheres and boxes: coaxial tables by quadrature
    // (the recurrence in nu loses up to 1e-6 of the largest entry once the box order
    // and k|d| pass ~100); alpha takes the place of the sectorial columns at off_s
    bool gq = false;
    std::string name;
    const char *pn[6] = {"gen_wigner", "gen_sect", "gen_axial", "rot_fwd", "axial_apply", "rot_back"};
};

struct Profiler {
    bool on = false;
    struct Rec {
        const char *name;  // static string
        hipEvent_t a, b;
        double flops;
    };
    struct Tot {
        double ms = 0.0;
        int64_t launches = 0;
        double flops = 0.0;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    std::map<const char *, Tot> totals;  // keyed by the static name
    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        (void)hipEventCreate(&e);
        return e;
    }
    void flush() {
        for (auto &r : recs) {
            (void)hipEventSynchronize(r.b);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, r.a, r.b);
            Tot &t = totals[r.name];
            t.ms += ms;
            t.launches += 1;
            t.flops += r.flops;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        recs.clear();
    }
    ~Profiler() {
        for (auto &r : recs) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

}  // namespace ifs

struct ifs_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    ifs_params prm{};
    int ns = 0;
    int N = 0;
    std::vector<double> pos;        // 3*ns
    std::vector<double> radius;
    std::vector<std::complex<double>> mrel;
    std::vector<int32_t> Lj;        // requested per-sphere truncation
    std::vector<int32_t> Leff;      // min(L_j, N)
    std::vector<int32_t> active;    // row activity (reference row zeroing)
    std::vector<int64_t> off;       // compact normalised offsets
    std::vector<int32_t> group;     // cluster of each sphere (a multi-cluster context), else empty
    std::vector<int32_t> grp_first; // first sphere of each cluster, and ns (size clusters + 1)
    int64_t ncompact = 0;           // complex entries of the compact vector
    int64_t nflat = 0;              // complex entries of the flat vector

    ifs::DevBuf d_Leff, d_active, d_off, d_snm, d_an, d_bn, d_x, d_b;
    ifs::DevBuf d_pos;                   // sphere centres (3 ns doubles)
    void *ff_scratch = nullptr;          // this device's far-field scratch array (see ifs_optics.hip)
    ifs::DevBuf d_xh, d_acc;             // compact normalised input / accumulator
    ifs::DevBuf ws_d, ws_s, ws_ax, ws_y, ws_z;
    ifs::DevBuf ws_gq;                       // quadrature generator workspace
    std::map<int, ifs::DevBuf> gq_nodes;     // Q -> Gauss-Legendre nodes then weights (2Q doubles)
    double dsign = -1.0;            // d = dsign * (origin(source) - origin(target))
    size_t ws_budget = 0;           // translation workspace (chunked batches)
    size_t cache_budget = 0;        // tables kept across matvecs
    size_t ws_cached = 0;           // bytes held by cached batches

    ifs::Batch direct;              // sphere -> sphere (direct all pairs)
    ifs::fmm::Fmm *fmm = nullptr;

    // Krylov workspace
    ifs::DevBuf k_V, k_w, k_r, k_part, k_h, k_bhat, k_xhat;
    ifs::DevBuf k_stage;            // graph-mode staging of the Krylov vector (stable address)
    int k_restart = 0;

    ifs::Profiler prof;
    double flops_mv = 0.0, bytes_mv = 0.0;
    int64_t axial_tiles[2] = {0, 0};  // coaxial-shift launches: 64-wide tiles, 32-wide remainder tiles
    int force_rows[2] = {0, 0};       // rotation block heights set by ifs_set_class_tiles (0: autotuned)
    int autotune = 1;                 // time both rotation block heights on a batch's first product

    // HIP graphs of the matvec, one per argument set; a graph is replayed while the
    // signature of the buffers and plans it captured still holds
    struct Graph {
        const void *x;
        void *b;
        int normalised, nrhs, prec;
        int64_t ld, ldb;
        uint64_t sig;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
};


namespace ifs {
// Device address of the far-field scratch array on the current device (looked up
// once per device); ifs_create calls it with the context's device current.
int far_field_scratch(int device, void **out);
// Translations grouped in classes that share one translation vector (one octree
// offset), run as tiled products (ifs_gemm.hip).  At finalize the classes are
// ordered by (length, polar angle) and cut into chunks that fit the workspace;
// inside a chunk, consecutive classes with one polar angle form a rotation run
// (one Wigner-d table) and consecutive classes with one length an axial segment
// (one coaxial table).
struct ClassBatch {
    // input, filled by add_class
    std::vector<Task> gen;             // per class: table parameters (offsets 0)
    std::vector<int> ncols;            // per class
    std::vector<int64_t> src, dst;     // columns of all classes, class by class
    // plan
    struct Piece {
        GroupDesc g;
        int gen;                       // class whose Task generates the table
        int di;                        // index of g in descs (device copy)
    };
    struct ChunkPlan {
        int col_begin = 0, ncols = 0;
        std::vector<Piece> runs, segs;
        int scat_begin = 0, scat_end = 0;  // range in the scatter segment array
    };
    std::vector<ChunkPlan> chunks;
    std::vector<GroupDesc> descs;
    std::vector<int32_t> pidx;         // phase row per column
    std::vector<double2> phase;        // per class: exp(i m phi), m = -PL..PL
    std::vector<ScatterSeg> scat;
    std::vector<int32_t> scat_cols;
    int PL = 0;
    int plan_nrhs = 0;                 // right-hand sides the plan was cut for
    int n_dslots = 1;                  // Wigner tables kept at once (chunk runs share slots r % n_dslots)
    // Table keys: one Wigner table per |cos(theta)| (a run of cos(theta) < 0 mirrors
    // it), one coaxial table per (length, kind).  When the context's cache budget
    // holds every key's table, the tables persist across matvecs (generated once).
    std::vector<int> dkey, axkey;      // per class
    std::vector<Task> dkeys;           // per Wigner key: generator task (cth = |cth|)
    std::vector<int> axkey_cls;        // per coaxial key: a class that generates it
    std::vector<GqPlan> gq;            // per coaxial key: quadrature plan
    std::vector<char> d_ready, ax_ready;
    bool d_persist = false, ax_persist = false;
    DevBuf own_d, own_ax, d_dkeys;
    int rot_rows[2][2] = {{0, 0}, {0, 0}};  // [precision][forward, back]: rotation block height, 0 = not tuned
    int64_t src_stride = 0, dst_stride = 0;  // pool stride between right-hand sides (complex)
    DevBuf d_desc, d_gen, d_src, d_pidx, d_phase, d_scat, d_scat_cols;
    int64_t n_d = 0, n_s = 0, n_ax = 0, n_y = 0, n_z = 0;  // workspace (doubles for d, y, z; complex for s, ax)
    std::string name;
};
// b_r = (I + T A) x_r for r < nrhs: x at stride ld, b at stride ldb (-1: ld); normalised
// != 0 means x and b are in the normalised basis.
// prec: 0 FP64, 1 FP32 class products (M2M, M2L, L2L); everything else stays FP64.
int matvec_impl(ifs_ctx *c, const double2 *x, double2 *b, int normalised, int nrhs, int64_t ld, int64_t ldb = -1,
                int prec = 0);
// matvec_impl through a HIP graph: the first call with an argument set runs eagerly
// (it sizes every buffer and builds the kept tables), the second captures the launches
// on the context's stream and replays them, later calls replay.
int matvec_graph(ifs_ctx *c, const double2 *x, double2 *b, int normalised, int nrhs, int64_t ld, int64_t ldb = -1,
                 int prec = 0);
void graphs_clear(ifs_ctx *c);
int class_batch_finalize(ifs_ctx *c, ClassBatch &B);
// Device Gauss-Legendre rule of Q nodes (cached per context): mu, then w.
int gq_rule(ifs_ctx *c, int Q, const double **mu, const double **w);
int class_batch_plan(ifs_ctx *c, ClassBatch &B, int nrhs);
// After every class batch is finalized: decides which tables persist, then plans each.
int class_batches_prepare(ifs_ctx *c, const std::vector<ClassBatch *> &bs);
int class_batch_run(ifs_ctx *c, ClassBatch &B, const double2 *src_pool, double2 *dst_pool, int nrhs, int prec = 0);
void class_batch_cost(const ClassBatch &B, double &flops, double &bytes);

int batch_finalize(ifs_ctx *c, Batch &B, size_t budget_bytes, bool allow_cache);
int batch_run(ifs_ctx *c, Batch &B, const double2 *src_pool, double2 *dst_pool, int nrhs = 1,
              int64_t src_stride = 0, int64_t dst_stride = 0);
void batch_cost(const Batch &B, double &flops, double &bytes);
Task make_task(const double *dv, double k, int kind, int Ls, int Lt, int64_t src, int64_t dst);

// Profiling scope: records HIP events around the launches it encloses.
struct Scope {
    ifs_ctx *c;
    const char *name;
    hipEvent_t a{}, b{};
    double flops;
    Scope(ifs_ctx *c_, const char *n, double algorithmic_flops = 0.0);
    ~Scope();
};
}  // namespace ifs
