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

// A translation class: columns [col_begin, col_begin + ncols) of the class
// source/target offset arrays share one translation vector and truncations.
struct ClassDesc {
    int64_t off_d;    // Wigner-d table (doubles) in the d workspace
    int64_t off_ax;   // coaxial table (complex) in the coaxial workspace
    double phi;       // azimuth of the translation vector
    int32_t Ls, Lt, K;
    int32_t ncols;
    int32_t col_begin;
    int32_t pad;
};

// A run of consecutive tasks that share one target expansion.
struct Segment {
    int64_t dst;     // complex offset of the target expansion
    int32_t Lt;
    int32_t t_begin, t_end;  // task index range [begin, end)
};

// Sizes (in doubles for d, in complex entries otherwise).
__host__ __device__ inline int64_t wig_size(int Lmax, int K) {
    // sum_{n=1}^{Lmax} (2 min(n,K)+1)(2n+1)
    int64_t s = 0;
    int nk = Lmax < K ? Lmax : K;
    s = (int64_t)(nk + 1) * (2 * nk + 1) * (2 * nk + 3) / 3 - 1;  // n = 1..nk: (2n+1)^2
    if (Lmax > K) s += (int64_t)(2 * K + 1) * ((int64_t)(Lmax + 1) * (Lmax + 1) - (int64_t)(K + 1) * (K + 1));
    return s;
}
__host__ __device__ inline int64_t wig_off(int n, int K) { return wig_size(n - 1, K); }

__host__ __device__ inline int64_t rows_size(int L, int K) {
    // sum_{n=1}^{L} (2 min(n,K)+1)
    int nk = L < K ? L : K;
    int64_t s = (int64_t)(nk + 1) * (nk + 1) - 1;
    if (L > K) s += (int64_t)(2 * K + 1) * (L - K);
    return s;
}
__host__ __device__ inline int64_t rows_off(int n, int K) { return rows_size(n - 1, K); }

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

// Kernel launchers (ifs_kernels.hip).  All enqueue on `s`.
void launch_gen_wigner(const Task *tasks, int ntasks, int maxpairs, double *wd, hipStream_t s);
void launch_gen_sect(const Task *tasks, int ntasks, int max_nmax, double2 *wsect, hipStream_t s);
void launch_gen_axial(const Task *tasks, int ntasks, int maxK, int max_nmax, const double2 *wsect, double2 *wax,
                      hipStream_t s);
void launch_rot_fwd(const Task *tasks, int ntasks, int64_t max_rows, int max_Ls, const double2 *src_pool,
                    const double *wd, double2 *wy, hipStream_t s);
void launch_axial_apply(const Task *tasks, int ntasks, int64_t max_rows, const double2 *wax, const double2 *wy,
                        double2 *wz, hipStream_t s);
void launch_rot_back(const Task *tasks, const Segment *segs, int nsegs, int yblocks, const double *wd,
                     const double2 *wz, double2 *dst_pool, hipStream_t s);

void launch_class(const ClassDesc &h, const ClassDesc *d_cls, const int64_t *src_off, const int64_t *dst_off,
                  const double2 *src_pool, double2 *dst_pool, const double *wd, const double2 *wax, double *wy,
                  double *wz, int stage, hipStream_t s);

void launch_mie(int ns, const double *xsize, const double2 *mrel, int N, double2 *an, double2 *bn, hipStream_t s);
void launch_pack_in(int ns, int N, const int32_t *Leff, const int64_t *off, const double *snm, const double2 *x,
                    double2 *xh, int nrhs, int64_t ldx, int64_t ldxh, hipStream_t s);
void launch_finalize(int ns, int N, const int32_t *Leff, const int32_t *active, const int64_t *off,
                     const double *snm, const double2 *an, const double2 *bn, const double2 *x, const double2 *acc,
                     double2 *b, int nrhs, int64_t ldx, int64_t ldacc, hipStream_t s);
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
