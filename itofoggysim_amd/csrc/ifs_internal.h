// Internal declarations shared by the host orchestration (ifs_api.cpp) and the
// HIP kernels (ifs_kernels.hip).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ifs {

// Coefficient storage
// -------------------
// Every expansion (sphere or box) of order L is stored in the normalised basis
//     c_hat(n,m) = s(n,m) c(n,m),   s(n,m) = sqrt(n(n+1)(2n+1) (n-m)!/(n+m)!)
// as 2*L(L+2) complex doubles: the a-part (n = 1..L, m = -n..n at n*n-1+m+n)
// followed by the b-part.  In this basis a translation is unitary for regular
// waves and factors into rotation (Wigner d), coaxial shift and back-rotation.

// One translation task: source expansion -> target expansion.
//   out(target) += R(theta,phi)^-1 * Axial(t, kind) * R(theta,phi) * src
// with R = d(theta) * diag(exp(i m phi)).
struct Task {
    int64_t src;     // complex offset of the source expansion (a-part) in the source pool
    int64_t dst;     // complex offset of the target expansion (a-part) in the target pool
    int64_t off_d;   // workspace offset (doubles) of the Wigner-d table
    int64_t off_ax;  // workspace offset (complex) of the coaxial table
    int64_t off_y;   // workspace offset (complex) of the rotated source
    int64_t off_z;   // workspace offset (complex) of the shifted expansion
    int64_t off_s;   // workspace offset (complex) of the sectorial scalar columns
    double t;        // k |d|
    double cth;      // cos(theta) of the translation direction
    double phi;      // azimuth of the translation direction
    int32_t Ls, Lt;  // source / target truncation
    int32_t K;       // min(Ls, Lt): largest |m| carried through the coaxial shift
    int32_t kind;    // 1: regular (j_n), 3: outgoing -> regular (h1_n)
};

// Work descriptor of the class-batched products (ifs_gemm.hip).  Translations
// are columns; a rotation run is a column range whose translation vectors share
// the polar angle (one Wigner-d table), an axial segment a column range whose
// vectors share the length and radial kind (one coaxial table).  Columns live in
// a chunk workspace: Y and Z are row-major with 4*ncy doubles per row, W (the
// back-rotated columns) is column-major, 2*Lt(Lt+2) complex per column.
//
// With several right-hand sides every translation is one column per RHS: a run
// or segment covers ncols * nrhs workspace columns starting at ycol0 * nrhs (RHS r,
// translation t at column ycol0 * nrhs + r * ncols + t), and W holds RHS r of the
// chunk's translation c at column r * ncy + c.  Source and target pools hold the
// RHS at strides src_stride / dst_stride (complex).
struct GroupDesc {
    int64_t off_d;     // Wigner-d table (doubles) in the d workspace
    int64_t off_ax;    // coaxial table (complex) in the coaxial workspace
    int32_t Ls, Lt, K;
    int32_t ncols;     // translations of this run / segment
    int32_t col_begin; // first column in the batch column arrays (source offset, phase row)
    int32_t ycol0;     // first translation in the chunk workspace
    int32_t ncy;       // translations of the chunk
    int32_t PL;        // phase rows hold exp(i m phi), m = -PL..PL
    int32_t nrhs;      // right-hand sides
    int32_t mirror;    // run of the supplementary polar angle pi - b: its tables are those of b,
                       // with d^n_{km}(pi - b) = (-1)^{n+k} d^n_{k,-m}(b) applied in the products
    int64_t src_stride, dst_stride;
};

// Deterministic reduction of back-rotated columns into one target expansion.
struct ScatterSeg {
    int64_t dst;       // complex offset of the target expansion (a-part)
    int32_t beg, end;  // range in the chunk's sorted column list
};

// A run of consecutive tasks that share one target expansion.
struct Segment {
    int64_t dst;     // complex offset of the target expansion
    int32_t Lt;
    int32_t t_begin, t_end;  // task index range [begin, end)
};

// Sizes (in doubles for d, in complex entries otherwise).
// Folded Wigner tables: per degree n (1..L), with kk = min(n,K), the matrices
//   E[k][m] = d^n_{km} + (-1)^m d^n_{k,-m},  O[k][m] = d^n_{km} - (-1)^m d^n_{k,-m}
// for k = 0..kk, m = 0..n (so E[k][0] = 2 d^n_{k0}, O[k][0] = O[0][m] = 0), row-major,
// E then O.  d_{-k,-m} = (-1)^{k-m} d_{km} makes them carry all of d^n.
__host__ __device__ inline int64_t fold_size(int L, int K) {
    // sum_{n=1}^{L} 2 (min(n,K)+1)(n+1)
    int nk = L < K ? L : K;
    int64_t a = nk + 1;
    int64_t s = 2 * (a * (a + 1) * (2 * a + 1) / 6 - 1);  // n = 1..nk: 2 (n+1)^2
    if (L > K) s += (int64_t)(K + 1) * ((int64_t)(L + 1) * (L + 2) - (int64_t)(K + 1) * (K + 2));
    return s;
}
__host__ __device__ inline int64_t fold_off(int n, int K) { return fold_size(n - 1, K); }

__host__ __device__ inline int64_t rows_size(int L, int K) {
    // sum_{n=1}^{L} (2 min(n,K)+1)
    int nk = L < K ? L : K;
    int64_t s = (int64_t)(nk + 1) * (nk + 1) - 1;
    if (L > K) s += (int64_t)(2 * K + 1) * (L - K);
    return s;
}
__host__ __device__ inline int64_t rows_off(int n, int K) { return rows_size(n - 1, K); }
// Rows (n, k >= 0) of degrees 1..L: sum_{n=1}^{L} (min(n,K)+1).
__host__ __device__ inline int64_t frow_size(int L, int K) {
    int nk = L < K ? L : K;
    int64_t s = (int64_t)nk * (nk + 3) / 2;
    if (L > K) s += (int64_t)(K + 1) * (L - K);
    return s;
}

__host__ __device__ inline int64_t ax_block(int k, int Lt, int Ls) {
    int q = k < 0 ? -k : k;
    if (q < 1) q = 1;
    return (int64_t)(Lt - q + 1) * (Ls - q + 1);
}
// sum_{q=a}^{b} (Lt+1-q)(Ls+1-q)
__host__ __device__ inline int64_t ax_sum(int a, int b, int Lt, int Ls) {
    if (b < a) return 0;
    int64_t P = Lt + 1, Q = Ls + 1, A = a, B = b;
    int64_t cnt = B - A + 1;
    int64_t s1 = (A + B) * cnt / 2;
    int64_t s2 = B * (B + 1) * (2 * B + 1) / 6 - (A - 1) * A * (2 * A - 1) / 6;
    return cnt * P * Q - (P + Q) * s1 + s2;
}
// Offset (complex) of the A block of coaxial order k; the B block follows it.
// Blocks are ordered k = -K..K, block k holds rows n = q..Lt, cols nu = q..Ls, q = max(1,|k|).
__host__ __device__ inline int64_t ax_off(int k, int Lt, int Ls, int K) {
    if (k <= 0) {
        int j = -k;  // blocks kk = -K..-(j+1) have q = K..j+1
        return 2 * ax_sum(j + 1, K, Lt, Ls);
    }
    return 2 * ax_sum(1, K, Lt, Ls) + 2 * ax_block(0, Lt, Ls) + 2 * ax_sum(1, k - 1, Lt, Ls);
}
__host__ __device__ inline int64_t ax_size(int Lt, int Ls, int K) { return ax_off(K, Lt, Ls, K) + 2 * ax_block(K, Lt, Ls); }
__host__ __device__ inline int64_t sect_size(int K, int nmax) { return (int64_t)(K + 1) * (nmax + 2); }

// Coaxial tables by Gauss-Legendre quadrature (ifs_coax_gq.hip): one table of
// (kind, t, Lt, Ls, K) in the layout of launch_gen_axial, generated in batches of
// mb orders m within ws_bytes of workspace.
constexpr int GQ_MAXB = 96;
struct GqPlan {
    int kind = 1;
    double t = 0.0;
    int Lt = 0, Ls = 0, K = 0;
    int P = 0;                         // outgoing kind: Legendre degree of T
    int nb = 1;                        // bands of T: band b holds p in [lo[b], lo[b+1])
    int lo[GQ_MAXB] = {0};
    int Q = 0, Nr = 0, mb = 1;         // nodes, Legendre rows, orders per batch
    size_t ws_bytes = 0;
};
GqPlan gq_plan(int kind, double t, int Lt, int Ls, int K, size_t ws_cap_bytes);
// mu, w: the Q Gauss-Legendre nodes and weights on the device
void launch_gen_axial_gq(const GqPlan &g, const double *mu, const double *w, void *ws, double2 *wax, hipStream_t s);

// Regular coaxial tables of generic tasks (sphere <-> box) by quadrature, batched:
// alpha of task T at T.off_s ((K+1) (Lt+2) (Ls+1) complex), tables at T.off_ax.
struct GqbChunk {
    int Q = 0;         // shared Gauss-Legendre nodes
    int Nr = 0;        // Legendre rows of the shared table: max over tasks of max(Lt + 1, Ls)
    int maxK = 0;
    int max_big = 0;   // max over tasks of max(Lt + 1, Ls)
    int max_small = 0; // max over tasks of min(Lt + 1, Ls)
    int64_t max_ent = 0;  // max over tasks of Lt * Ls
};
int gqb_nodes(double t, int Lt, int Ls);
inline int64_t gqb_alpha_size(int Lt, int Ls, int K) { return (int64_t)(K + 1) * (Lt + 2) * (Ls + 1); }
// V: (maxK + 1) (Nr + 1) Q doubles of workspace
// Sphere <-> box coaxial shift applied through the quadrature (no table): the chunk's
// Legendre table V at the nodes (launch_gqb_legendre), then per right-hand side
// launch_gqb_apply (Y -> Z of every task).  Dynamic LDS: gqb_apply_lds(Q, Nr) bytes.
size_t gqb_apply_lds(int Q, int Nr);
void launch_gqb_legendre(const GqbChunk &g, const double *mu, double *V, hipStream_t s);
void launch_gqb_apply(const Task *tasks, int ntasks, const GqbChunk &g, const double *mu, const double *w,
                      const double *V, const double2 *wy, double2 *wz, hipStream_t s);
void launch_gen_axial_gqb(const Task *tasks, int ntasks, const GqbChunk &g, const double *mu, const double *w,
                          double *V, double2 *walpha, double2 *wax, hipStream_t s);

// Kernel launchers (ifs_kernels.hip).  All enqueue on `s`.
void launch_gen_wigner_fold(const Task *tasks, int ntasks, int maxpairs, double *wd, hipStream_t s);
// Same tables stored in FP32 (the recurrence runs in FP64).
void launch_gen_wigner_fold_f32(const Task *tasks, int ntasks, int maxpairs, float *wd, hipStream_t s);
void launch_gen_sect(const Task *tasks, int ntasks, int max_nmax, double2 *wsect, hipStream_t s);
void launch_gen_axial(const Task *tasks, int ntasks, int maxK, int max_nmax, const double2 *wsect, double2 *wax,
                      hipStream_t s);
// max_rows: largest frow_size(Ls, K) of the tasks
void launch_rot_fwd(const Task *tasks, int ntasks, int64_t max_rows, int max_Ls, const double2 *src_pool,
                    const double *wd, double2 *wy, hipStream_t s);
// Forward rotation for high source orders: block per (task, degree), lanes along m.
void launch_rot_fwd_wide(const Task *tasks, int ntasks, int max_Ls, const double2 *src_pool, const double *wd,
                         double2 *wy, hipStream_t s);
void launch_axial_apply(const Task *tasks, int ntasks, int64_t max_rows, const double2 *wax, const double2 *wy,
                        double2 *wz, hipStream_t s);
// Coaxial shift Y -> Z with the coefficients generated on the fly (needs the
// sectorial columns of launch_gen_sect; target orders < 1024 when Ls <= Lt).
// Coaxial-shift tiles of a segment of nt translation columns: full 64-wide tiles, and one
// 32-wide tile for a remainder of at most 32 (a larger remainder takes a 64-wide tile).
inline void axial_tile_split(int nt, int &nfull, bool &tail) {
    const int rem = nt % 64;
    nfull = rem > 0 && rem <= 32 ? nt / 64 : (nt + 63) / 64;
    tail = nfull * 64 < nt;
}
void launch_axial_fused(const Task *tasks, int ntasks, int maxK, int max_nmax, int max_Lt, const double2 *wsect,
                        const double2 *wy, double2 *wz, hipStream_t s);
void launch_rot_back(const Task *tasks, const Segment *segs, int nsegs, int yblocks, const double *wd,
                     const double2 *wz, double2 *dst_pool, hipStream_t s);

// Precision of the class products (prec): 0 FP64 throughout; 1 FP32 products with FP32
// accumulation (pools and coaxial tables stay FP64 and are converted where staged into
// LDS; Y and Z hold FP32, the back-rotated columns W FP64).  ftab: the Wigner tables at
// wd are FP32 (prec 1 only; tables regenerated per matvec are then generated in FP32).
// stage 0: forward rotation of a run (source gather with exp(i m phi)) -> Y
// stage 1: coaxial shift of a segment Y -> Z
// stage 2: back rotation of a run (exp(-i m phi)) Z -> W (aliases Y)
// rows: block height of the rotation products, 64 or 32 (the same results; the faster
// one depends on the orders and the column count, see the autotuner in class_batch_run)
void launch_class(const GroupDesc &h, const GroupDesc *d_desc, const int64_t *src_off, const int32_t *pidx,
                  const double2 *phase, const double2 *src_pool, const double *wd, const double2 *wax, double *wy,
                  double *wz, int stage, int prec, int ftab, int rows, hipStream_t s);
// dst_pool[r dst_stride + seg.dst + e] += sum_{i in [beg,end)} W[r ncy + cols[i]][e],
// e < nbox = 2 Lt(Lt+2), r < nrhs
void launch_scatter(const ScatterSeg *segs, int nsegs, const int32_t *cols, const double2 *W, int64_t nbox,
                    double2 *dst_pool, int ncy, int nrhs, int64_t dst_stride, hipStream_t s);

void launch_mie(int ns, const double *xsize, const double2 *mrel, int N, double2 *an, double2 *bn, hipStream_t s);
void launch_pack_in(int ns, int N, const int32_t *Leff, const int64_t *off, const double *snm, const double2 *x,
                    double2 *xh, int nrhs, int64_t ldx, int64_t ldxh, hipStream_t s);
void launch_finalize(int ns, int N, const int32_t *Leff, const int32_t *active, const int64_t *off,
                     const double *snm, const double2 *an, const double2 *bn, const double2 *x, const double2 *acc,
                     double2 *b, int nrhs, int64_t ldx, int64_t ldb, int64_t ldacc, hipStream_t s);
void launch_zero(double2 *p, int64_t n, hipStream_t s);
void launch_scale_flat(int ns, int N, const double *snm, const double2 *x, double2 *y, int inverse, hipStream_t s);

// BLAS-1 style kernels for the Krylov solver (FP64 complex).
void launch_dots(const double2 *V, int64_t ldv, int nv, const double2 *w, int64_t n, double2 *out, double2 *partial,
                 int nblocks, hipStream_t s);
void launch_gemv_sub(const double2 *V, int64_t ldv, int nv, const double2 *h, double2 *w, int64_t n, hipStream_t s);
void launch_axpby(double2 a, const double2 *x, double2 b, double2 *y, int64_t n, hipStream_t s);
void launch_scale_copy(double2 a, const double2 *x, double2 *y, int64_t n, hipStream_t s);
void launch_gemv_add(const double2 *V, int64_t ldv, int nv, const double2 *c, double2 *x, int64_t n, hipStream_t s);

}  // namespace ifs
