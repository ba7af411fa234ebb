// Class-batched translations as tiled FP64 GEMMs (gfx950).
//
// A translation class is a set of translations that share one translation
// vector (the same octree offset at one level, e.g. one M2L offset or one child
// octant of M2M/L2L).  Its Wigner-d table and coaxial table are generated once
// and every translation of the class becomes a column of three dense products:
//
//   forward rotation, degree nu:  Y^nu (rows k)  = D^nu (k x m) * X^nu (m x cols)
//   coaxial shift, order k:       Z_k  (rows n)  = [A_k B_k] * [Y_k^a ; Y_k^b]   (complex, a/b mixed)
//   back rotation, degree n:      W^n  (rows m)  = D^n^T (m x k) * Z^n (k x cols)
//
// Columns are the translations of the class, 4 reals each (a.re, a.im, b.re, b.im).
// The gather of the sources with the phase e^{i m phi} is fused into the loads of
// the forward product, the phase e^{-i m phi} and the scatter-add into the targets
// into the epilogue of the back product.  Targets of one class are distinct, so
// the scatter needs no atomics.
#include <hip/hip_runtime.h>

#include "ifs_internal.h"

namespace ifs {

namespace {

constexpr int BM = 64;   // rows per tile
constexpr int BN = 64;   // real columns per tile (16 translations)
constexpr int BK = 16;   // inner step
constexpr int TM = 4;    // rows per thread
constexpr int TN = 4;    // columns per thread (one translation: a.re a.im b.re b.im)

__device__ __forceinline__ int fidx(int n, int m) { return n * n - 1 + m + n; }

}  // namespace

// Forward rotation of one class: grid (col tiles, row tiles, degree-1).
__global__ __launch_bounds__(256) void k_class_rot_fwd(const ClassDesc *__restrict__ cls,
                                                       const int64_t *__restrict__ src_off,
                                                       const double2 *__restrict__ src_pool,
                                                       const double *__restrict__ wd, double *__restrict__ wy) {
    const ClassDesc C = *cls;
    const int nu = blockIdx.z + 1;
    if (nu > C.Ls) return;
    const int K = C.K;
    const int kk = nu < K ? nu : K;
    const int M = 2 * kk + 1;        // rows k = -kk..kk
    const int KI = 2 * nu + 1;       // inner m = -nu..nu
    const int NC = 4 * C.ncols;
    const int row0 = blockIdx.y * BM;
    const int col0 = blockIdx.x * BN;
    if (row0 >= M) return;
    __shared__ double As[BK][BM + 1];
    __shared__ double Bs[BK][BN + 1];
    const double *A = wd + C.off_d + wig_off(nu, K);  // row-major (M x KI)
    const int tid = threadIdx.x;
    const int tr = tid / (BN / TN);  // 0..15
    const int tc = tid % (BN / TN);  // 0..15
    double acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = 0.0;
    const int Lsrc = C.Ls;
    const int64_t nsrc = (int64_t)Lsrc * (Lsrc + 2);
    for (int k0 = 0; k0 < KI; k0 += BK) {
        // A tile: BM x BK (1024 values, 4 per thread)
        for (int e = tid; e < BM * BK; e += 256) {
            int r = e / BK, q = e % BK;
            int gr = row0 + r, gq = k0 + q;
            As[q][r] = (gr < M && gq < KI) ? A[(int64_t)gr * KI + gq] : 0.0;
        }
        // B tile: BK x BN, fused gather with phase: one thread per (m, translation)
        for (int e = tid; e < BK * (BN / 4); e += 256) {
            int q = e / (BN / 4), t = e % (BN / 4);
            int gq = k0 + q;
            int j = (col0 / 4) + t;
            double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
            if (gq < KI && j < C.ncols) {
                int m = gq - nu;
                const double2 *x = src_pool + src_off[C.col_begin + j] + fidx(nu, m);
                double2 xa = x[0], xb = x[nsrc];
                double s, c;
                sincos((double)m * C.phi, &s, &c);
                v0 = c * xa.x - s * xa.y;
                v1 = c * xa.y + s * xa.x;
                v2 = c * xb.x - s * xb.y;
                v3 = c * xb.y + s * xb.x;
            }
            Bs[q][4 * t + 0] = v0;
            Bs[q][4 * t + 1] = v1;
            Bs[q][4 * t + 2] = v2;
            Bs[q][4 * t + 3] = v3;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < BK; ++q) {
            double a[TM], b[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) a[i] = As[q][tr * TM + i];
#pragma unroll
            for (int j = 0; j < TN; ++j) b[j] = Bs[q][tc * TN + j];
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
    double *Y = wy + rows_off(nu, K) * (int64_t)NC;
#pragma unroll
    for (int i = 0; i < TM; ++i) {
        int gr = row0 + tr * TM + i;
        if (gr >= M) continue;
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            int gc = col0 + tc * TN + j;
            if (gc < NC) Y[(int64_t)gr * NC + gc] = acc[i][j];
        }
    }
}

// Coaxial shift of one class: grid (translation tiles, row tiles, 2K+1 orders).
// Z_k^a(n) = sum_nu A_k(n,nu) Y_k^a(nu) + B_k(n,nu) Y_k^b(nu);  Z_k^b = B Y^a + A Y^b
constexpr int AM = 32;   // rows n per tile
constexpr int AN = 32;   // translations per tile
constexpr int AK = 8;

__global__ __launch_bounds__(256) void k_class_axial(const ClassDesc *__restrict__ cls,
                                                     const double2 *__restrict__ wax, const double *__restrict__ wy,
                                                     double *__restrict__ wz) {
    const ClassDesc C = *cls;
    const int K = C.K;
    const int k = (int)blockIdx.z - K;
    const int ak = k < 0 ? -k : k;
    const int q = ak < 1 ? 1 : ak;
    const int Lt = C.Lt, Ls = C.Ls;
    const int M = Lt - q + 1;
    const int KI = Ls - q + 1;
    const int row0 = blockIdx.y * AM;
    const int col0 = blockIdx.x * AN;
    if (row0 >= M) return;
    const int NC = 4 * C.ncols;
    const double2 *Ab = wax + C.off_ax + ax_off(k, Lt, Ls, K);
    const double2 *Bb = Ab + ax_block(k, Lt, Ls);
    __shared__ double2 As[AK][AM];
    __shared__ double2 Bs2[AK][AM];
    __shared__ double2 Ya[AK][AN + 1];
    __shared__ double2 Yb[AK][AN + 1];
    const int tid = threadIdx.x;
    const int tr = tid / 16;  // 0..15 -> 2 rows each
    const int tc = tid % 16;  // 0..15 -> 2 translations each
    double2 za[2][2], zb[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            za[i][j] = make_double2(0.0, 0.0);
            zb[i][j] = make_double2(0.0, 0.0);
        }
    for (int k0 = 0; k0 < KI; k0 += AK) {
        for (int e = tid; e < AM * AK; e += 256) {
            int r = e / AK, c = e % AK;
            int gr = row0 + r, gc = k0 + c;
            bool ok = gr < M && gc < KI;
            As[c][r] = ok ? Ab[(int64_t)gr * KI + gc] : make_double2(0.0, 0.0);
            Bs2[c][r] = ok ? Bb[(int64_t)gr * KI + gc] : make_double2(0.0, 0.0);
        }
        for (int e = tid; e < AK * AN; e += 256) {
            int c = e / AN, t = e % AN;
            int nu = q + k0 + c;
            int j = col0 + t;
            double2 va = make_double2(0.0, 0.0), vb = make_double2(0.0, 0.0);
            if (nu <= Ls && j < C.ncols) {
                int kk = nu < K ? nu : K;
                const double *row = wy + (rows_off(nu, K) + (k + kk)) * (int64_t)NC + 4 * j;
                va = make_double2(row[0], row[1]);
                vb = make_double2(row[2], row[3]);
            }
            Ya[c][t] = va;
            Yb[c][t] = vb;
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < AK; ++c) {
            double2 a[2], b[2], ya[2], yb[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                a[i] = As[c][tr * 2 + i];
                b[i] = Bs2[c][tr * 2 + i];
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                ya[j] = Ya[c][tc * 2 + j];
                yb[j] = Yb[c][tc * 2 + j];
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    // za += a*ya + b*yb ; zb += b*ya + a*yb
                    za[i][j].x = fma(a[i].x, ya[j].x, za[i][j].x);
                    za[i][j].x = fma(-a[i].y, ya[j].y, za[i][j].x);
                    za[i][j].y = fma(a[i].x, ya[j].y, za[i][j].y);
                    za[i][j].y = fma(a[i].y, ya[j].x, za[i][j].y);
                    za[i][j].x = fma(b[i].x, yb[j].x, za[i][j].x);
                    za[i][j].x = fma(-b[i].y, yb[j].y, za[i][j].x);
                    za[i][j].y = fma(b[i].x, yb[j].y, za[i][j].y);
                    za[i][j].y = fma(b[i].y, yb[j].x, za[i][j].y);
                    zb[i][j].x = fma(b[i].x, ya[j].x, zb[i][j].x);
                    zb[i][j].x = fma(-b[i].y, ya[j].y, zb[i][j].x);
                    zb[i][j].y = fma(b[i].x, ya[j].y, zb[i][j].y);
                    zb[i][j].y = fma(b[i].y, ya[j].x, zb[i][j].y);
                    zb[i][j].x = fma(a[i].x, yb[j].x, zb[i][j].x);
                    zb[i][j].x = fma(-a[i].y, yb[j].y, zb[i][j].x);
                    zb[i][j].y = fma(a[i].x, yb[j].y, zb[i][j].y);
                    zb[i][j].y = fma(a[i].y, yb[j].x, zb[i][j].y);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        int n = q + row0 + tr * 2 + i;
        if (n > Lt) continue;
        int kk = n < K ? n : K;
        double *row = wz + (rows_off(n, K) + (k + kk)) * (int64_t)NC;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            int t = col0 + tc * 2 + j;
            if (t >= C.ncols) continue;
            row[4 * t + 0] = za[i][j].x;
            row[4 * t + 1] = za[i][j].y;
            row[4 * t + 2] = zb[i][j].x;
            row[4 * t + 3] = zb[i][j].y;
        }
    }
}

// Back rotation of one class with fused phase and scatter-add:
// grid (col tiles, row tiles, degree-1); rows are m = -n..n, inner k.
__global__ __launch_bounds__(256) void k_class_rot_back(const ClassDesc *__restrict__ cls,
                                                        const int64_t *__restrict__ dst_off,
                                                        const double *__restrict__ wd, const double *__restrict__ wz,
                                                        double2 *__restrict__ dst_pool) {
    const ClassDesc C = *cls;
    const int n = blockIdx.z + 1;
    if (n > C.Lt) return;
    const int K = C.K;
    const int kk = n < K ? n : K;
    const int M = 2 * n + 1;    // rows m
    const int KI = 2 * kk + 1;  // inner k
    const int NC = 4 * C.ncols;
    const int row0 = blockIdx.y * BM;
    const int col0 = blockIdx.x * BN;
    if (row0 >= M) return;
    __shared__ double As[BK][BM + 1];
    __shared__ double Bs[BK][BN + 1];
    const double *D = wd + C.off_d + wig_off(n, K);  // (KI x M) row-major: D[k][m]
    const double *Z = wz + rows_off(n, K) * (int64_t)NC;
    const int tid = threadIdx.x;
    const int tr = tid / (BN / TN);
    const int tc = tid % (BN / TN);
    double acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = 0.0;
    for (int k0 = 0; k0 < KI; k0 += BK) {
        for (int e = tid; e < BM * BK; e += 256) {
            int q = e / BM, r = e % BM;  // coalesced along m
            int gr = row0 + r, gq = k0 + q;
            As[q][r] = (gr < M && gq < KI) ? D[(int64_t)gq * M + gr] : 0.0;
        }
        for (int e = tid; e < BK * BN; e += 256) {
            int q = e / BN, cc = e % BN;
            int gq = k0 + q, gc = col0 + cc;
            Bs[q][cc] = (gq < KI && gc < NC) ? Z[(int64_t)gq * NC + gc] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < BK; ++q) {
            double a[TM], b[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) a[i] = As[q][tr * TM + i];
#pragma unroll
            for (int j = 0; j < TN; ++j) b[j] = Bs[q][tc * TN + j];
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
    // epilogue: thread owns rows m and exactly one translation (TN = 4 reals)
    const int j = (col0 + tc * TN) / 4;
    if (j >= C.ncols) return;
    double2 *dst = dst_pool + dst_off[C.col_begin + j];
    const int64_t ndst = (int64_t)C.Lt * (C.Lt + 2);
#pragma unroll
    for (int i = 0; i < TM; ++i) {
        int gr = row0 + tr * TM + i;
        if (gr >= M) continue;
        int m = gr - n;
        double s, c;
        sincos(-(double)m * C.phi, &s, &c);
        double2 wa = make_double2(c * acc[i][0] - s * acc[i][1], c * acc[i][1] + s * acc[i][0]);
        double2 wb = make_double2(c * acc[i][2] - s * acc[i][3], c * acc[i][3] + s * acc[i][2]);
        int64_t o = fidx(n, m);
        double2 da = dst[o], db = dst[o + ndst];
        dst[o] = make_double2(da.x + wa.x, da.y + wa.y);
        dst[o + ndst] = make_double2(db.x + wb.x, db.y + wb.y);
    }
}

void launch_class(const ClassDesc &h, const ClassDesc *d_cls, const int64_t *src_off, const int64_t *dst_off,
                  const double2 *src_pool, double2 *dst_pool, const double *wd, const double2 *wax, double *wy,
                  double *wz, int stage, hipStream_t s) {
    const int NC = 4 * h.ncols;
    const int K = h.K;
    if (stage == 0) {
        int mt = (2 * (h.Ls < K ? h.Ls : K) + 1 + BM - 1) / BM;
        dim3 grid((NC + BN - 1) / BN, mt, h.Ls);
        hipLaunchKernelGGL(k_class_rot_fwd, grid, dim3(256), 0, s, d_cls, src_off, src_pool, wd, wy);
    } else if (stage == 1) {
        int mt = (h.Lt + AM - 1) / AM;
        dim3 grid((h.ncols + AN - 1) / AN, mt, 2 * K + 1);
        hipLaunchKernelGGL(k_class_axial, grid, dim3(256), 0, s, d_cls, wax, wy, wz);
    } else {
        int mt = (2 * h.Lt + 1 + BM - 1) / BM;
        dim3 grid((NC + BN - 1) / BN, mt, h.Lt);
        hipLaunchKernelGGL(k_class_rot_back, grid, dim3(256), 0, s, d_cls, dst_off, wd, wz, dst_pool);
    }
}

}  // namespace ifs
