/*
 * ifs.h -- C ABI of the MI355X multi-sphere scattering matvec (ItoFoggySim hot path).
 *
 * Conventions
 *   - Every call returns IFS_OK (0) or a negative IFS_E* status; nothing is printed.
 *   - Complex vectors are interleaved doubles (re, im), i.e. the memory layout of
 *     Fortran complex(c_double_complex) and C double _Complex.
 *   - Vector layout is the reference's flat layout: per sphere j a block of
 *     N(N+2) a_nm followed by N(N+2) b_nm, (n,m) at n*n-1+(m+n), N = n_global
 *     (FoggySim/src/lib/mie_theory/lib_mie_ms_solver_interface_helper_functions.f90:49-78).
 *   - A context owns device memory and one HIP stream; the caller owns host buffers.
 *     There is no global state: contexts are independent.  Calls on one context are
 *     not thread-safe; calls on different contexts are.
 */
#ifndef IFS_H
#define IFS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IFS_OK 0
#define IFS_E_ARG -1        /* invalid argument / size mismatch */
#define IFS_E_NOMEM -2      /* device allocation failed */
#define IFS_E_HIP -3        /* HIP runtime error */
#define IFS_E_NODEVICE -4   /* no GPU visible */
#define IFS_E_NOCONV -5     /* solver reached max iterations without converging */
#define IFS_E_BREAKDOWN -6  /* solver breakdown (zero residual direction) */

typedef struct ifs_ctx ifs_ctx;

/* One cluster of spheres.  Replaces the sphere part of the reference's global
 * simulation_data (lib_mie_type.f90:43-99, lib_mie_ms_data_container.f90:25). */
typedef struct {
    int32_t n_spheres;
    const double *centres; /* [3*n] sphere centres d_0_j, same length unit as 1/k */
    const double *radii;   /* [n] radii */
    const double *m_rel;   /* [2*n] complex relative refractive index (re, im) per sphere */
    const int32_t *lmax;   /* [n] per-sphere truncation L_j; NULL: L_j = ceil(x+4.05 x^(1/3)+2), x = k r */
} ifs_cluster;

#define IFS_ZSEL_BESSEL_J 1
#define IFS_ZSEL_HANKEL1 3

/* Parameters.  Replaces simulation_data%spherical_harmonics and the solver
 * parameter type (lib_mie_type.f90:84-99, lib_mie_ms_solver_interface.f90:91-129). */
typedef struct {
    double k;                /* wavenumber in the medium, 2 pi n_medium / lambda_0 */
    int32_t n_global;        /* N: global truncation of the flat vector layout */
    int32_t z_selector;      /* radial kind of sphere-to-sphere translations (3: h1, 1: j) */
    int32_t ref_row_zeroing; /* 1: reproduce the reference's zeroed rows for spheres with L_j < N */
    int32_t use_fmm;         /* 0: direct all-pairs matvec, 1: MLFMM */
    int32_t leaf_size;       /* MLFMM: max spheres per leaf box (reference tree_s_opt) */
    double fmm_d0;           /* MLFMM: box truncation L = kD + 1.8 d0^(2/3) (kD)^(1/3); D = box edge */
    int32_t fmm_lmax_cap;    /* MLFMM: upper bound on box truncation (0: none) */
    int32_t reserved[4];
} ifs_params;

/* Builds device tables, the Mie T-matrix and (MLFMM) the octree.
 * Replaces lib_mie_ms_solver_constructor + lib_mie_ms_ml_fmm_constructor
 * (lib_mie_ms_solver_interface.f90:91-129; lib_mie_ms_ml_fmm_interface.f90:46-95). */
int ifs_create(ifs_ctx **ctx, const ifs_cluster *cluster, const ifs_params *params);
int ifs_destroy(ifs_ctx *ctx);

/* Length of the flat vector (complex entries): 2 * n_spheres * N(N+2). */
int64_t ifs_vector_length(const ifs_ctx *ctx);

/* b = (I + T A) x on host buffers (copies in and out over PCIe).
 * Replaces the GMRES callback calc_vector_b => lib_mie_ms_solver_calculate_vector_b
 * (lib_solver_GMRES.f90:52-57,72-77; lib_mie_ms_solver_interface.f90:739-783). */
int ifs_matvec(ifs_ctx *ctx, const double *x, double *b, int64_t n);

/* Same, on device buffers already resident in HBM, enqueued on the context stream. */
int ifs_matvec_device(ifs_ctx *ctx, const void *x_dev, void *b_dev, int64_t n);

/* Launch stream: NULL restores the context's own stream.  The stream must be a
 * default-priority HIP stream. */
int ifs_set_stream(ifs_ctx *ctx, void *hip_stream);
void *ifs_get_stream(ifs_ctx *ctx);

/* Mie coefficients a_n^j, b_n^j, n = 1..N, as [n_spheres * N] interleaved complex.
 * Replaces lib_mie_type_func_get_sphere_parameter (lib_mie_type_functions.f90:46-80). */
int ifs_get_mie(ifs_ctx *ctx, double *an, double *bn);

/* Right-hand side for one x-polarised plane wave travelling along +z:
 * b = T p (p_nm, q_nm of lib_mie_illumination.f90:636-763 at alpha = beta = 0). */
int ifs_plane_wave_rhs(ifs_ctx *ctx, double *rhs, int64_t n);

typedef struct {
    int32_t restart;      /* Krylov subspace size (reference default 10; BASELINE uses 50) */
    int32_t max_iter;     /* maximum matvecs */
    double tol;           /* relative residual ||b - M x|| / ||b|| */
    int32_t use_x0;       /* 1: x holds the initial guess on entry */
    int32_t reserved[3];
} ifs_solve_opts;

typedef struct {
    int32_t iterations;   /* matvecs performed */
    int32_t converged;
    double rel_residual;  /* true relative residual at exit, recomputed in FP64 */
    double seconds;
} ifs_solve_info;

/* Solves (I + T A) x = rhs with restarted GMRES on the device (FP64 Arnoldi,
 * Givens and true-residual recomputation).  Replaces lib_mie_ms_solver_gmres_run
 * (lib_solver_GMRES.f90:116-316). Host buffers. */
int ifs_solve(ifs_ctx *ctx, const double *rhs, double *x, int64_t n, const ifs_solve_opts *opts,
              ifs_solve_info *info);
/* Same on device buffers. */
int ifs_solve_device(ifs_ctx *ctx, const void *rhs_dev, void *x_dev, int64_t n, const ifs_solve_opts *opts,
                     ifs_solve_info *info);

/* Octree export for bit-exact checks (lib_tree.f90:147-270, lib_ml_fmm_helper_functions.f90:366-446).
 *   uindex_out[n_spheres]: Morton index of each sphere's leaf box at level l_max
 *   levels_out[2]: l_min, l_max.  MLFMM contexts only. */
int ifs_tree_export(ifs_ctx *ctx, int64_t *uindex_out, int32_t *levels_out);

/* Host-only octree functions (no device work), bit-exact restatements of libtree:
 *   scaling to [0,1)^3                         lib_tree.f90:320-374 (x_min_D: x_min[3], D)
 *   Morton universal index at a level          lib_tree_helper_functions.f90:202-364
 *   26 neighbours at distance k, -1 if outside lib_tree_helper_functions.f90:914-997
 *   box centre (scaled coordinates)            lib_tree_helper_functions.f90:774-799
 *   interaction list I4 (ascending; returns its length)   lib_tree.f90:994-1143
 *   l_max for at most s elements per box       lib_tree.f90:1188-1242 */
int ifs_tree_scale(const double *pts, int32_t n, double *scaled, double *x_min_D);
int ifs_tree_universal_index(const double *scaled, int32_t n, int32_t level, int64_t *out);
int ifs_tree_neighbours(int64_t k, int64_t box, int32_t level, int64_t out[26]);
int ifs_tree_centre(int64_t box, int32_t level, double out[3]);
int ifs_tree_domain_i4(int64_t box, int32_t level, int64_t *out, int32_t cap);
int ifs_tree_level_max(const int64_t *sorted_uindex, int64_t n, int64_t s);

/* Interaction counts of the MLFMM plan per kind (P2M, M2M, M2L, L2L, L2P, P2P)
 * and the box truncation per level (levels_L[l] for l = 0..31, 0 if unused). */
int ifs_fmm_stats(ifs_ctx *ctx, int64_t counts[6], int32_t levels_L[32]);

/* Per-kernel device time of the last profiled calls: enable = 1 starts
 * recording HIP events around every launch on the context stream. */
int ifs_profile_enable(ifs_ctx *ctx, int32_t enable);
/* Returns the number of kernel kinds written (<= cap); names are static strings;
 * flops is the algorithmic flop count of the recorded launches (0 where not defined). */
int ifs_profile_read(ifs_ctx *ctx, const char **names, double *total_ms, int64_t *launches, double *flops,
                     int32_t cap);
/* Algorithmic flop count of one matvec (rotation form (128/3) L^3 per translation,
 * plus T-apply and identity) and compulsory bytes, from the actual plan. */
int ifs_matvec_cost(ifs_ctx *ctx, double *flops, double *bytes);

#ifdef __cplusplus
}
#endif
#endif
