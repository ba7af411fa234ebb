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

// Real products (rotations): 128 x 128 tiles, 256 threads, 8 x 8 outputs per thread.
// Thread (ty, tx) owns rows ty + 16 i (i < 8) and the 2 translations tx + 16 j
// (j < 2), i.e. the 8 real columns 4 (tx + 16 j) + p.
constexpr int BM = 128;
constexpr int BN = 128;  // 32 translations
constexpr int BK = 8;

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
    const int M = 2 * kk + 1;   // rows k = -kk..kk
    const int KI = 2 * nu + 1;  // inner m = -nu..nu
    const int NC = 4 * C.ncols;
    const int row0 = blockIdx.y * BM;
    const int col0 = blockIdx.x * BN;
    if (row0 >= M) return;
    __shared__ double As[BK][BM];
    __shared__ double Bs[BK][BN];
    const double *A = wd + C.off_d + wig_off(nu, K);  // row-major (M x KI)
    const int tid = threadIdx.x;
    const int ty = tid >> 4, tx = tid & 15;
    double acc[8][8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = 0.0;
    const int64_t nsrc = (int64_t)C.Ls * (C.Ls + 2);
    // gather assignment: thread -> (inner row q, translation t) for the B tile
    const int gq_l = tid >> 5, gt = tid & 31;
    const int gj = (col0 >> 2) + gt;
    const double2 *xs = (gj < C.ncols) ? src_pool + src_off[C.col_begin + gj] : nullptr;
    for (int k0 = 0; k0 < KI; k0 += BK) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int idx = tid + 256 * e;       // 0..1023 over 128 rows x 8 inner
            int r = idx >> 3, q = idx & 7;
            int gr = row0 + r, gqq = k0 + q;
            As[q][r] = (gr < M && gqq < KI) ? A[(int64_t)gr * KI + gqq] : 0.0;
        }
        {
            int gqq = k0 + gq_l;
            double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
            if (gqq < KI && xs) {
                int m = gqq - nu;
                const double2 *x = xs + fidx(nu, m);
                double2 xa = x[0], xb = x[nsrc];
                double sn, cs;
                sincos((double)m * C.phi, &sn, &cs);
                v0 = cs * xa.x - sn * xa.y;
                v1 = cs * xa.y + sn * xa.x;
                v2 = cs * xb.x - sn * xb.y;
                v3 = cs * xb.y + sn * xb.x;
            }
            Bs[gq_l][4 * gt + 0] = v0;
            Bs[gq_l][4 * gt + 1] = v1;
            Bs[gq_l][4 * gt + 2] = v2;
            Bs[gq_l][4 * gt + 3] = v3;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < BK; ++q) {
            double a[8], b[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) a[i] = As[q][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int p = 0; p < 4; ++p) b[4 * j + p] = Bs[q][4 * (tx + 16 * j) + p];
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
    double *Y = wy + rows_off(nu, K) * (int64_t)NC;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        int gr = row0 + ty + 16 * i;
        if (gr >= M) continue;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            int t = (col0 >> 2) + tx + 16 * j;
            if (t >= C.ncols) continue;
            double *y = Y + (int64_t)gr * NC + 4 * t;
            y[0] = acc[i][4 * j + 0];
            y[1] = acc[i][4 * j + 1];
            y[2] = acc[i][4 * j + 2];
            y[3] = acc[i][4 * j + 3];
        }
    }
}

// Coaxial shift of one class: grid (translation tiles, row tiles, 2K+1 orders).
// Z_k^a(n) = sum_nu A_k(n,nu) Y_k^a(nu) + B_k(n,nu) Y_k^b(nu);  Z_k^b = B Y^a + A Y^b
// 64 rows x 64 translations per tile; thread (ty, tx) owns rows ty + 16 i and
// translations tx + 16 j (i, j < 4).
constexpr int AM = 64;
constexpr int AN = 64;
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
    __shared__ double2 Ya[AK][AN];
    __shared__ double2 Yb[AK][AN];
    const int tid = threadIdx.x;
    const int ty = tid >> 4, tx = tid & 15;
    double2 za[4][4], zb[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            za[i][j] = make_double2(0.0, 0.0);
            zb[i][j] = make_double2(0.0, 0.0);
        }
    for (int k0 = 0; k0 < KI; k0 += AK) {
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            int idx = tid + 256 * e;  // 0..511 over 64 rows x 8 inner
            int r = idx >> 3, c = idx & 7;
            int gr = row0 + r, gc = k0 + c;
            bool ok = gr < M && gc < KI;
            As[c][r] = ok ? Ab[(int64_t)gr * KI + gc] : make_double2(0.0, 0.0);
            Bs2[c][r] = ok ? Bb[(int64_t)gr * KI + gc] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            int idx = tid + 256 * e;  // 0..511 over 8 inner x 64 translations
            int c = idx >> 6, t = idx & 63;
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
            double2 a[4], b[4], ya[4], yb[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                a[i] = As[c][ty + 16 * i];
                b[i] = Bs2[c][ty + 16 * i];
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                ya[j] = Ya[c][tx + 16 * j];
                yb[j] = Yb[c][tx + 16 * j];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
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
    for (int i = 0; i < 4; ++i) {
        int n = q + row0 + ty + 16 * i;
        if (n > Lt) continue;
        int kk = n < K ? n : K;
        double *row = wz + (rows_off(n, K) + (k + kk)) * (int64_t)NC;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int t = col0 + tx + 16 * j;
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
    __shared__ double As[BK][BM];
    __shared__ double Bs[BK][BN];
    const double *D = wd + C.off_d + wig_off(n, K);  // (KI x M) row-major: D[k][m]
    const double *Z = wz + rows_off(n, K) * (int64_t)NC;
    const int tid = threadIdx.x;
    const int ty = tid >> 4, tx = tid & 15;
    double acc[8][8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = 0.0;
    for (int k0 = 0; k0 < KI; k0 += BK) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int idx = tid + 256 * e;  // 8 inner x 128 rows, coalesced along m
            int q = idx >> 7, r = idx & 127;
            int gr = row0 + r, gq = k0 + q;
            As[q][r] = (gr < M && gq < KI) ? D[(int64_t)gq * M + gr] : 0.0;
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int idx = tid + 256 * e;  // 8 inner x 128 columns
            int q = idx >> 7, cc = idx & 127;
            int gq = k0 + q, gc = col0 + cc;
            Bs[q][cc] = (gq < KI && gc < NC) ? Z[(int64_t)gq * NC + gc] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < BK; ++q) {
            double a[8], b[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) a[i] = As[q][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int p = 0; p < 4; ++p) b[4 * j + p] = Bs[q][4 * (tx + 16 * j) + p];
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
    const int64_t ndst = (int64_t)C.Lt * (C.Lt + 2);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        int t = (col0 >> 2) + tx + 16 * j;
        if (t >= C.ncols) continue;
        double2 *dst = dst_pool + dst_off[C.col_begin + t];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            int gr = row0 + ty + 16 * i;
            if (gr >= M) continue;
            int m = gr - n;
            double sn, cs;
            sincos(-(double)m * C.phi, &sn, &cs);
            double2 wa = make_double2(cs * acc[i][4 * j] - sn * acc[i][4 * j + 1],
                                      cs * acc[i][4 * j + 1] + sn * acc[i][4 * j]);
            double2 wb = make_double2(cs * acc[i][4 * j + 2] - sn * acc[i][4 * j + 3],
                                      cs * acc[i][4 * j + 3] + sn * acc[i][4 * j + 2]);
            int64_t o = fidx(n, m);
            double2 da = dst[o], db = dst[o + ndst];
            dst[o] = make_double2(da.x + wa.x, da.y + wa.y);
            dst[o + ndst] = make_double2(db.x + wb.x, db.y + wb.y);
        }
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
