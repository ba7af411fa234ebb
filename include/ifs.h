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

/* Translation vector of the sphere/box couplings:
 *   IFS_D_PHYSICAL  d = origin(target) - origin(source): the sign for which Xu's
 *                   coefficients as the reference evaluates them conserve energy
 *                   (lossless clusters: far-field integral = extinction);
 *   IFS_D_REFERENCE d = origin(source) - origin(target), as the reference's direct
 *                   path passes it (lib_mie_ms_solver_interface.f90:941-943). */
#define IFS_D_PHYSICAL 0
#define IFS_D_REFERENCE 1

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
    int32_t workspace_mb;    /* translation workspace limit in MiB (0: min(free/4, 48 GiB)); small
                                values split the translation classes into more workspace chunks */
    int32_t d_convention;    /* IFS_D_PHYSICAL (default) or IFS_D_REFERENCE */
    int32_t cache_mb;        /* device memory for translation tables kept across matvecs, in MiB
                                (0: 65% of free memory at creation; < 0: none, tables are
                                regenerated in every matvec) */
    int32_t precision;       /* ifs_matvec* arithmetic: IFS_PREC_FP64 (0) or IFS_PREC_MIXED (1) */
} ifs_params;

/* Precision of a matvec or of a solve's Krylov products:
 *   IFS_PREC_FP64   FP64 throughout;
 *   IFS_PREC_MIXED  the box translations (M2M, M2L, L2L: the MLFMM's dense products) run
 *                   as FP32 matrix-core products with FP32 accumulation on the FP64 tables;
 *                   sphere-level translations, T-apply, vectors and solver state stay FP64. */
#define IFS_PREC_FP64 0
#define IFS_PREC_MIXED 1

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

/* b_r = (I + T A) x_r for nrhs right-hand sides in one pass: every translation table is
 * generated once and applied to all of them (the translations become nrhs columns each
 * of the class products).  Device buffers holding the vectors at stride ld >= n
 * (complex); no reference counterpart (the reference applies one vector per call). */
int ifs_matvec_batch_device(ifs_ctx *ctx, int32_t nrhs, const void *x_dev, void *b_dev, int64_t n, int64_t ld);
/* Same on host buffers [nrhs][n]. */
int ifs_matvec_batch(ifs_ctx *ctx, int32_t nrhs, const double *x, double *b, int64_t n);
/* ifs_matvec_device through a HIP graph of its launches: the first call with a given
 * (x_dev, b_dev) pair runs eagerly (sizing every buffer and building the kept tables),
 * the second captures the launches on the context's stream, later calls replay them
 * (one hipGraphLaunch).  A graph is dropped and captured again when a buffer or plan
 * it refers to has changed; profiling (ifs_profile_enable) runs eagerly.  No
 * counterpart in the reference (its matvec is a Fortran callback). */
int ifs_matvec_graph_device(ifs_ctx *ctx, const void *x_dev, void *b_dev, int64_t n);

/* Arithmetic of the context's ifs_matvec* calls from now on (IFS_PREC_FP64 or
 * IFS_PREC_MIXED; ifs_params.precision at creation).  The tables are shared. */
int ifs_set_precision(ifs_ctx *ctx, int32_t precision);

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

/* Incident plane wave of unit amplitude travelling along (alpha, beta) (polar and
 * azimuthal angle of the propagation direction; at alpha = 0 the field is polarised
 * along (cos beta, sin beta, 0)): p_mn, q_mn at every sphere centre in the flat layout,
 *   p_mn = e^{i k khat.d_j} tau_mn(alpha) e^{-i m beta} / (n(n+1)), q_mn with pi_mn
 * (lib_mie_illumination.f90:636-763, get_p_q_j_j_plane_wave_core).  Host buffers. */
int ifs_plane_wave(ifs_ctx *ctx, double alpha, double beta, double *p, int64_t n);
/* rhs = T p for that wave; rows beyond a sphere's L_j (and of inactive spheres) are 0.
 * Generalises ifs_plane_wave_rhs (alpha = beta = 0). */
int ifs_plane_wave_rhs_dir(ifs_ctx *ctx, double alpha, double beta, double *rhs, int64_t n);
/* The two polarisations of the reference's illumination: te = 0 is the wave above
 * (TM), te = 1 the same wave with p and q exchanged (TE, lib_mie_illumination.f90:
 * 430-470 and 347-352).  Coefficients p, and rhs = T p. */
int ifs_plane_wave_mode(ifs_ctx *ctx, double alpha, double beta, int32_t te, double *p, int64_t n);
int ifs_plane_wave_rhs_mode(ifs_ctx *ctx, double alpha, double beta, int32_t te, double *rhs, int64_t n);

/* Cluster cross sections from the solved coefficients x and the incident p of the same
 * wave (flat layout): out = {C_sca, C_ext, C_abs} in the squared length unit of 1/k.
 * Per-sphere sums of lib_mie_ss_helper_functions.f90:1380-1460 (Xu eq. 57) for the
 * extinction; the absorption of each sphere from its exciting field a^j_mn / a_n^j;
 * C_sca = C_ext - C_abs. */
int ifs_cross_sections(ifs_ctx *ctx, const double *x, const double *p, int64_t n, double out[3]);
int ifs_cross_sections_device(ifs_ctx *ctx, const void *x_dev, const void *p_dev, int64_t n, double out[3]);

/* Far field of the scattered wave, E_s ~ e^{ikr}/(kr) (F_theta theta_hat + F_phi phi_hat),
 * at ndir directions (theta[i], phi[i]) from the solved coefficients x (flat layout):
 * out[4i..4i+3] = Re F_theta, Im F_theta, Re F_phi, Im F_phi.  No reference counterpart
 * for clusters (single sphere: the Barber-Hill S2 program, light_scattering.f90:54-204). */
int ifs_far_field(ifs_ctx *ctx, const double *x, int64_t n, int32_t ndir, const double *theta, const double *phi,
                  double *out);
int ifs_far_field_device(ifs_ctx *ctx, const void *x_dev, int64_t n, int32_t ndir, const void *theta_dev,
                         const void *phi_dev, void *out_dev);
/* Far field as ifs_far_field_device, allocating nothing: the per-direction tables go to
 * a fixed device array of the library, IFS_FF_SCRATCH_ENTRIES complex (32 MiB) on every GPU
 * the library runs on, held until the process ends; the sphere centres are kept in the
 * context.  ndir * N(N+2) must not exceed IFS_FF_SCRATCH_ENTRIES (else IFS_E_ARG and nothing
 * runs: split the directions).  Enqueued on the context stream, no synchronisation.  All
 * calls on one GPU share the array, so calls that overlap on one GPU race.  No counterpart
 * in the reference: for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). */
#define IFS_FF_SCRATCH_ENTRIES (1 << 21)
int ifs_far_field_device_noalloc(ifs_ctx *ctx, const void *x_dev, int64_t n, int32_t ndir, const void *theta_dev,
                                 const void *phi_dev, void *out_dev);
/* Amplitude matrix S1..S4 (Bohren & Huffman 3.12, incidence along +z) from the far
 * fields Fx, Fy of the x- (beta = 0) and y-polarised (beta = pi/2) waves:
 * S[8i..8i+7] = S1, S2, S3, S4 (re, im). */
int ifs_amplitude_matrix(int32_t ndir, const double *phi, const double *Fx, const double *Fy, double *S);
/* Mueller matrix (Bohren & Huffman 3.16): M[16i + 4r + c] = S_{r+1,c+1}. */
int ifs_mueller(int32_t ndir, const double *S, double *M);

typedef struct {
    int32_t restart;      /* Krylov subspace size (reference default 10; the 2000-sphere benchmark uses 50) */
    int32_t max_iter;     /* maximum matvecs */
    double tol;           /* relative residual in the normalised basis (ifs_solve_info.rel_residual) */
    int32_t use_x0;       /* 1: x holds the initial guess on entry */
    int32_t verbose;      /* 1: progress lines on stderr (every 5 matvecs and every restart) */
    int32_t precision;    /* IFS_PREC_FP64: every product FP64.  IFS_PREC_MIXED: the Arnoldi products
                             in mixed precision (ifs_params.precision), the true residual of every
                             restart cycle in FP64, so the cycles refine the iterate in FP64 until the
                             FP64 residual meets tol (iterative refinement) */
    int32_t flags;        /* IFS_SOLVE_GRAPH: the products replay a HIP graph of the matvec (captured
                             on the context's stream at the second product; one staging vector
                             carries the Krylov vector into it).  0: eager launches */
} ifs_solve_opts;

#define IFS_SOLVE_GRAPH 1

typedef struct {
    int32_t iterations;        /* matvecs performed */
    int32_t converged;
    double rel_residual;       /* true relative residual at exit, recomputed in FP64, in the normalised
                                  basis x_hat = s_nm x, s_nm = sqrt(n(n+1)(2n+1)(n-m)!/(n+m)!):
                                  ||S (b - M x)|| / ||S b||, the norm the tolerance refers to */
    double seconds;
    double rel_residual_flat;  /* the same residual in the caller's flat basis, ||b - M x|| / ||b||
                                  (the norm of the reference's GMRES driver); entries there span
                                  1e+-80 at N = 50, so this norm is dominated by the rows m = n */
} ifs_solve_info;

/* Solves (I + T A) x = rhs with restarted GMRES on the device (FP64 Arnoldi,
 * Givens and true-residual recomputation).  Replaces lib_mie_ms_solver_gmres_run
 * (lib_solver_GMRES.f90:116-316). Host buffers. */
int ifs_solve(ifs_ctx *ctx, const double *rhs, double *x, int64_t n, const ifs_solve_opts *opts,
              ifs_solve_info *info);
/* Same on device buffers. */
int ifs_solve_device(ifs_ctx *ctx, const void *rhs_dev, void *x_dev, int64_t n, const ifs_solve_opts *opts,
                     ifs_solve_info *info);
/* nrhs systems (I + T A) x_r = rhs_r by restarted GMRES in lock step: one batched
 * matvec per Arnoldi step for all right-hand sides, per-RHS orthogonalisation and
 * convergence (info[nrhs]).  Device buffers at stride ld; host variant [nrhs][n]. */
int ifs_solve_batch_device(ifs_ctx *ctx, int32_t nrhs, const void *rhs_dev, void *x_dev, int64_t n, int64_t ld,
                           const ifs_solve_opts *opts, ifs_solve_info *info);
int ifs_solve_batch(ifs_ctx *ctx, int32_t nrhs, const double *rhs, double *x, int64_t n, const ifs_solve_opts *opts,
                    ifs_solve_info *info);

/* nrhs systems (I + T A) x_r = rhs_r by restarted block GMRES: one Krylov space shared
 * by all right-hand sides (restart = block steps per cycle, m nrhs basis vectors; one
 * batched product per step), so each right-hand side is corrected from every direction
 * the others explored.  info[r].iterations counts the products per right-hand side.
 * No reference counterpart (the reference solves one right-hand side per run). */
int ifs_solve_block_device(ifs_ctx *ctx, int32_t nrhs, const void *rhs_dev, void *x_dev, int64_t n, int64_t ld,
                           const ifs_solve_opts *opts, ifs_solve_info *info);
int ifs_solve_block(ifs_ctx *ctx, int32_t nrhs, const double *rhs, double *x, int64_t n, const ifs_solve_opts *opts,
                    ifs_solve_info *info);

/* Many independent clusters in one context: the spheres of clusters[0..nclusters)
 * in order, each cluster interacting only with itself (a block-diagonal operator), so
 * that one matvec launch covers every cluster's pairs.  Direct plans only
 * (params->use_fmm = 0, IFS_E_ARG otherwise); one wavenumber and one n_global for
 * all; per-sphere orders (lmax) for every cluster or for none.  The flat vector is the
 * clusters' flat vectors one after another.  No counterpart in the reference, which
 * solves one cluster per run (lib_mie_multi_sphere.f90:111-235). */
int ifs_create_multi(ifs_ctx **out, int32_t nclusters, const ifs_cluster *clusters, const ifs_params *params);
/* Clusters of a context (1 for ifs_create), and where each starts:
 * first_sphere[nclusters + 1], first_entry[nclusters + 1] (flat complex entries);
 * either pointer may be NULL. */
int32_t ifs_cluster_count(const ifs_ctx *ctx);
int ifs_cluster_offsets(const ifs_ctx *ctx, int64_t *first_sphere, int64_t *first_entry);
/* Every cluster of the context solved on its own Krylov space by restarted GMRES in
 * lock step (one product per Arnoldi step for all clusters; per-cluster
 * orthogonalisation, convergence and info[nclusters]).  Device and host buffers. */
int ifs_solve_multi_device(ifs_ctx *ctx, const void *rhs_dev, void *x_dev, int64_t n, const ifs_solve_opts *opts,
                           ifs_solve_info *info);
int ifs_solve_multi(ifs_ctx *ctx, const double *rhs, double *x, int64_t n, const ifs_solve_opts *opts,
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

/* Octree build on the GPU, bit-exact with the host functions above (same scaling,
 * Morton index, s-rule): the universal indices of the n points at the threshold
 * level 21, sorted ascending (uindex_sorted[n]), the permutation that sorts them
 * (perm[n], stable), levels = {l_min, l_max}.  Replaces lib_tree_constructor's
 * sort and level search (lib_tree.f90:147-270, 1162-1242).  Host buffers. */
int ifs_tree_build_gpu(const double *pts, int32_t n, int32_t s, int64_t *uindex_sorted, int32_t *perm,
                       int32_t levels[2]);
/* I4 (k = 1) of every box of an ascending list of distinct boxes at one level,
 * restricted to the boxes of the list, ascending: out[nb * 189] (-1 padded),
 * count[nb] (lib_tree.f90:994-1143).  GPU; host buffers. */
int ifs_tree_i4_gpu(const int64_t *boxes, int64_t nb, int32_t level, int64_t *out, int32_t *count);

/* Serving: continuous batching of scattering requests on one GPU.  A request is a
 * cluster with its parameters and ndir incidence directions (alpha[i], beta[i]); its
 * result is {C_sca, C_ext, C_abs} per direction, solved to the relative residual tol
 * with lock-step GMRES(restart) in the request's precision (ifs_params.precision).
 * `slots` worker threads each run one request at a time on a context with its own
 * stream, so admitted requests overlap on the device.  At submit the request's device
 * memory is planned on the host (vectors and Krylov basis, MLFMM pools from an octree
 * built with the ifs_tree_* functions, translation workspace, kept tables), and its
 * context is created with that workspace and table cache (ifs_params.workspace_mb and
 * cache_mb are overridden), so it stays inside its footprint; a queued request is
 * admitted when a slot is free and its footprint fits what the running ones leave of
 * hbm_budget_gb (or nothing else runs).  Inputs are copied at submit.  No reference
 * counterpart (the reference solves one cluster per process). */
typedef struct ifs_server ifs_server;
typedef struct {
    int32_t status;          /* IFS_OK or the request's error */
    int32_t iterations;      /* largest matvec count over the directions */
    double latency_s;        /* submit -> done */
    double queue_s;          /* submit -> admitted */
    int64_t footprint_bytes; /* estimate used for admission */
} ifs_request_info;
int ifs_server_create(ifs_server **srv, int32_t slots, double hbm_budget_gb);
int ifs_server_submit(ifs_server *srv, const ifs_cluster *cluster, const ifs_params *params, int32_t ndir,
                      const double *alpha, const double *beta, double tol, int32_t restart, int64_t *request_id);
/* Blocks until the request is done; csec[3 * ndir]; returns the request's status. */
int ifs_server_wait(ifs_server *srv, int64_t request_id, double *csec, ifs_request_info *info);
/* out = {completed, p50 latency s, p99 latency s, completed per second since create,
 *        peak concurrent requests, peak admitted footprint bytes}. */
int ifs_server_stats(ifs_server *srv, double out[6]);
/* Finishes the queued requests, then stops the workers. */
int ifs_server_destroy(ifs_server *srv);
/* The admission plan of a request, on the host only (no device work): out = bytes of
 * vectors, MLFMM pools, translation workspace, kept tables, fixed buffers, their total,
 * then the workspace_mb and cache_mb (-1: none kept) its context would be created with,
 * for ndir directions, GMRES(restart), a budget of hbm_budget_gb shared by slots. */
int ifs_server_plan(const ifs_cluster *cluster, const ifs_params *params, int32_t ndir, int32_t restart,
                    double hbm_budget_gb, int32_t slots, double out[8]);
/* 1: each request's GMRES steps and its completion are reported on stderr; 0 (default): silent. */
int ifs_server_set_verbose(ifs_server *srv, int32_t verbose);

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
/* Algorithmic flop count of one matvec and compulsory bytes, from the actual plan:
 * per translation the products this library performs (folded rotations
 * 8 sum_n [(kk+1)(n+1) + kk n] each way, kk = min(n, K); decoupled coaxial shift
 * 16 sum_k (Lt-q+1)(Ls-q+1), q = max(1,|k|); about 21 L^3 in all), plus the T-apply
 * and the identity. */
int ifs_matvec_cost(ifs_ctx *ctx, double *flops, double *bytes);

/* Test support, no counterpart in the reference: order k of the coaxial translation
 * table (kind 1: regular j_n; 3: outgoing -> regular h1_n) of length t = k|d| along
 * the z axis, truncations Lt (target) and Ls (source), orders |m| <= K, generated on
 * the device by method 0 (the degree recurrence of the sphere-level tables) or 1 (the
 * Gauss-Legendre quadrature of the box tables), in the normalised basis.  out receives
 * 2 (Lt-q+1)(Ls-q+1) complex, q = max(1,|k|): the block P = A + B, then Q = A - B,
 * each row-major over n = q..Lt (rows) and nu = q..Ls (columns).  Synchronous; it
 * allocates and frees its own device memory. */
int ifs_coaxial_block(int32_t kind, double t, int32_t Lt, int32_t Ls, int32_t K, int32_t method, int32_t k,
                      double *out);

/* Instrumentation: coaxial-shift launches of the box translations since creation,
 * counts[0] with 64-wide translation tiles and counts[1] with the 32-wide tile that a
 * segment remainder of at most 32 translations (columns x right-hand sides) takes. */
int ifs_class_tile_stats(ifs_ctx *ctx, int64_t counts[2]);
/* Block heights of the box translations' rotation products: 64 or 32 rows (the same
 * results), or 0 to let the autotuner choose.  With autotune = 1 (the default) a
 * batch's first product times both heights on every launch and keeps the faster per
 * product; choices are remembered per shape for the rest of the process.  Drops the
 * context's captured graphs. */
int ifs_set_class_tiles(ifs_ctx *ctx, int32_t fwd_rows, int32_t back_rows, int32_t autotune);

#ifdef __cplusplus
}
#endif
#endif
