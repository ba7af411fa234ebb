// Class-batched translations as tiled FP64 matrix-core (MFMA) products (gfx950).
//
// A translation class is a set of translations that share one translation
// vector (the same octree offset at one level, e.g. one M2L offset or one child
// octant of M2M/L2L).  Its Wigner-d table and coaxial table are generated once
// and every translation of the class becomes a column of three dense products:
//
//   forward rotation, degree nu:  Y^nu (rows k)  = D^nu (k x m) * X^nu (m x cols)
//   coaxial shift, order k:       Z_k^+ = P_k Y_k^+,  Z_k^- = Q_k Y_k^-            (complex)
//   back rotation, degree n:      W^n  (rows m)  = D^n^T (m x k) * Z^n (k x cols)
//
// with x_+- = x_a +- x_b and P = A + B, Q = A - B (the [A B; B A] coaxial operator
// decoupled).  Columns are the translations of the class, 4 reals each
// (+.re, +.im, -.re, -.im).
// The gather of the sources with the phase e^{i m phi} is fused into the loads of
// the forward product, the phase e^{-i m phi} and the scatter-add into the targets
// into the epilogue of the back product.  Targets of one class are distinct, so
// the scatter needs no atomics.
#include <hip/hip_runtime.h>

#include "ifs_internal.h"

namespace ifs {

namespace {

typedef double dv4 __attribute__((ext_vector_type(4)));

// FP64 matrix-core tile: D(16x16) += A(16x4) B(4x16) with lane l holding
// A[l%16][l/16], B[l/16][l%16] and D[4v+l/16][l%16], v = 0..3.
__device__ __forceinline__ dv4 mfma(double a, double b, dv4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int fidx(int n, int m) { return n * n - 1 + m + n; }

// Rotation products: block tile 64 rows x 128 real columns (32 translations),
// 4 waves of 32 x 64, inner step 16 (four 16x16x4 steps).
constexpr int RM = 64;
constexpr int RN = 128;
constexpr int RK = 16;
constexpr int RPAD = 2;

}  // namespace

// Phase table e^{i m phi}, m = -L..L, into LDS (ph[m + L]).
__device__ __forceinline__ void phase_table(double2 *ph, int L, double phi, double sign) {
    for (int m = threadIdx.x; m <= 2 * L; m += blockDim.x) {
        double sn, cs;
        sincos(sign * (double)(m - L) * phi, &sn, &cs);
        ph[m] = make_double2(cs, sn);
    }
}

constexpr int PH_MAX = 2 * 1024 + 1;  // phase table entries (L <= 1024)

// Forward rotation of one class: grid (col tiles, row tiles, degree-1).
//   Y^nu[k][4j+p] = sum_m d^nu_{km} (e^{i m phi} x_j(nu,m))_p
// The next inner step is prefetched into registers while the matrix cores work.
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
    const int row0 = blockIdx.y * RM;
    const int col0 = blockIdx.x * RN;
    if (row0 >= M) return;
    __shared__ double As[RK][RM + RPAD];
    __shared__ double Bs[RK][RN + RPAD];
    __shared__ double2 ph[PH_MAX];
    phase_table(ph, nu, C.phi, 1.0);
    const double *A = wd + C.off_d + wig_off(nu, K);  // row-major (M x KI)
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 64;
    dv4 acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = dv4{0.0, 0.0, 0.0, 0.0};
    const int64_t nsrc = (int64_t)C.Ls * (C.Ls + 2);
    // gather: thread -> (translation t, inner q): idx = tid + 256 e, t = idx / 16, q = idx % 16
    const double2 *xs[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        int t = (tid + 256 * e) >> 4;
        int j = (col0 >> 2) + t;
        xs[e] = (j < C.ncols) ? src_pool + src_off[C.col_begin + j] + fidx(nu, 0) : nullptr;
    }
    __syncthreads();  // phase table ready
    double ra[4];
    double2 rxa[2], rxb[2];
    auto load = [&](int k0) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int idx = tid + 256 * e;  // 64 rows x 16 inner
            int r = idx >> 4, q = idx & 15;
            int gr = row0 + r, gq = k0 + q;
            ra[e] = (gr < M && gq < KI) ? A[(int64_t)gr * KI + gq] : 0.0;
        }
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            int gq = k0 + ((tid + 256 * e) & 15);
            if (gq < KI && xs[e]) {
                rxa[e] = xs[e][gq - nu];
                rxb[e] = xs[e][gq - nu + nsrc];
            } else {
                rxa[e] = make_double2(0.0, 0.0);
                rxb[e] = make_double2(0.0, 0.0);
            }
        }
    };
    load(0);
    for (int k0 = 0; k0 < KI; k0 += RK) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int idx = tid + 256 * e;
            As[idx & 15][idx >> 4] = ra[e];
        }
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            int idx = tid + 256 * e;
            int t = idx >> 4, q = idx & 15;
            int gq = k0 + q;
            double2 p = (gq < KI) ? ph[gq] : make_double2(1.0, 0.0);  // m + nu = gq
            double2 xa = make_double2(rxa[e].x + rxb[e].x, rxa[e].y + rxb[e].y);  // x_+
            double2 xb = make_double2(rxa[e].x - rxb[e].x, rxa[e].y - rxb[e].y);  // x_-
            Bs[q][4 * t + 0] = p.x * xa.x - p.y * xa.y;
            Bs[q][4 * t + 1] = p.x * xa.y + p.y * xa.x;
            Bs[q][4 * t + 2] = p.x * xb.x - p.y * xb.y;
            Bs[q][4 * t + 3] = p.x * xb.y + p.y * xb.x;
        }
        __syncthreads();
        if (k0 + RK < KI) load(k0 + RK);
#pragma unroll
        for (int kc = 0; kc < RK; kc += 4) {
            double a[2], b[4];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = As[kc + lk][wr + 16 * i + li];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kc + lk][wc + 16 * j + li];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = mfma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
    double *Y = wy + rows_off(nu, K) * (int64_t)NC;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int gc = col0 + wc + 16 * j + li;
            if (gc >= NC) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int gr = row0 + wr + 16 * i + 4 * v + lk;
                if (gr < M) Y[(int64_t)gr * NC + gc] = acc[i][j][v];
            }
        }
}

// Coaxial shift of one class: grid (translation tiles, row tiles, 2K+1 orders).
//   Z^+ = P Y^+,  Z^- = Q Y^-  (complex), computed as real products of
//   [[Pr, -Pi], [Pi, Pr]] with [Yr; Yi].  One 16x16x4 tile = 8 rows n x 2 parts
//   (tile row r: n = r % 8, part = r / 8), inner = 2 nu x 2 parts, 16 translations.
//   Block: 64 rows n x 64 translations; wave w owns rows 16 w .. 16 w + 15.
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
    const double2 *Pb = wax + C.off_ax + ax_off(k, Lt, Ls, K);
    const double2 *Qb = Pb + ax_block(k, Lt, Ls);
    __shared__ double2 Ps[AK][AM];
    __shared__ double2 Qs[AK][AM];
    __shared__ double Ys[AK][4 * AN + 2];
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int nl = li & 7, pout = li >> 3;   // A-fragment row: (n within 8, output part)
    const int nul = lk >> 1, pin = lk & 1;   // inner: (nu within 2, input part)
    const bool use_im = pout != pin;
    const double sgn = (pout == 0 && pin == 1) ? -1.0 : 1.0;
    dv4 accP[2][4], accQ[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            accP[i][j] = dv4{0.0, 0.0, 0.0, 0.0};
            accQ[i][j] = dv4{0.0, 0.0, 0.0, 0.0};
        }
    double2 pa[2], pb[2];
    double py[8];
    auto load = [&](int k0) {
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            int idx = tid + 256 * e;  // 64 rows x 8 inner
            int r = idx >> 3, c = idx & 7;
            int gr = row0 + r, gc = k0 + c;
            bool ok = gr < M && gc < KI;
            pa[e] = ok ? Pb[(int64_t)gr * KI + gc] : make_double2(0.0, 0.0);
            pb[e] = ok ? Qb[(int64_t)gr * KI + gc] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            int idx = tid + 256 * e;  // 8 inner x 256 reals (64 translations x 4)
            int c = idx >> 8, rr = idx & 255;
            int nu = q + k0 + c;
            int j = col0 + (rr >> 2);
            double v = 0.0;
            if (nu <= Ls && j < C.ncols) {
                int kq = nu < K ? nu : K;
                v = wy[(rows_off(nu, K) + (k + kq)) * (int64_t)NC + 4 * (int64_t)col0 + rr];
            }
            py[e] = v;
        }
    };
    load(0);
    for (int k0 = 0; k0 < KI; k0 += AK) {
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            int idx = tid + 256 * e;
            Ps[idx & 7][idx >> 3] = pa[e];
            Qs[idx & 7][idx >> 3] = pb[e];
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            int idx = tid + 256 * e;
            Ys[idx >> 8][idx & 255] = py[e];
        }
        __syncthreads();
        if (k0 + AK < KI) load(k0 + AK);
#pragma unroll
        for (int cc = 0; cc < AK; cc += 2) {
            double aP[2], aQ[2], bP[4], bQ[4];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                int r = w * 16 + 8 * i + nl;
                double2 p = Ps[cc + nul][r], qv = Qs[cc + nul][r];
                aP[i] = sgn * (use_im ? p.y : p.x);
                aQ[i] = sgn * (use_im ? qv.y : qv.x);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                int col = 16 * j + li;
                bP[j] = Ys[cc + nul][4 * col + pin];
                bQ[j] = Ys[cc + nul][4 * col + 2 + pin];
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    accP[i][j] = mfma(aP[i], bP[j], accP[i][j]);
                    accQ[i][j] = mfma(aQ[i], bQ[j], accQ[i][j]);
                }
        }
        __syncthreads();
    }
    // D row 4 v + lk of tile i: n = 8 i + ((4 v + lk) & 7), part = (4 v + lk) >> 3, i.e.
    // v = 0, 2 -> (re, im) of n = 8 i + lk;  v = 1, 3 -> (re, im) of n = 8 i + 4 + lk
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            int n = q + row0 + w * 16 + 8 * i + 4 * h + lk;
            if (n > Lt) continue;
            int kq = n < K ? n : K;
            double *row = wz + (rows_off(n, K) + (k + kq)) * (int64_t)NC;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                int t = col0 + 16 * j + li;
                if (t >= C.ncols) continue;
                row[4 * t + 0] = accP[i][j][h];
                row[4 * t + 1] = accP[i][j][h + 2];
                row[4 * t + 2] = accQ[i][j][h];
                row[4 * t + 3] = accQ[i][j][h + 2];
            }
        }
}

// Back rotation of one class with fused phase and scatter-add:
// grid (col tiles, row tiles, degree-1); rows m = -n..n, inner k.
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
    const int row0 = blockIdx.y * RM;
    const int col0 = blockIdx.x * RN;
    if (row0 >= M) return;
    __shared__ double As[RK][RM + RPAD];
    __shared__ double Bs[RK][RN + RPAD];
    __shared__ double2 ph[RM];
    for (int r = threadIdx.x; r < RM; r += blockDim.x) {
        double sn, cs;
        sincos(-(double)(row0 + r - n) * C.phi, &sn, &cs);
        ph[r] = make_double2(cs, sn);
    }
    const double *D = wd + C.off_d + wig_off(n, K);  // (KI x M) row-major: D[k][m]
    const double *Z = wz + rows_off(n, K) * (int64_t)NC;
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 64;
    dv4 acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = dv4{0.0, 0.0, 0.0, 0.0};
    double ra[4], rb[8];
    auto load = [&](int k0) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int idx = tid + 256 * e;  // 16 inner x 64 rows, coalesced along m
            int qq = idx >> 6, r = idx & 63;
            int gr = row0 + r, gq = k0 + qq;
            ra[e] = (gr < M && gq < KI) ? D[(int64_t)gq * M + gr] : 0.0;
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            int idx = tid + 256 * e;  // 16 inner x 128 columns
            int qq = idx >> 7, cc = idx & 127;
            int gq = k0 + qq, gc = col0 + cc;
            rb[e] = (gq < KI && gc < NC) ? Z[(int64_t)gq * NC + gc] : 0.0;
        }
    };
    load(0);
    for (int k0 = 0; k0 < KI; k0 += RK) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int idx = tid + 256 * e;
            As[idx >> 6][idx & 63] = ra[e];
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            int idx = tid + 256 * e;
            Bs[idx >> 7][idx & 127] = rb[e];
        }
        __syncthreads();
        if (k0 + RK < KI) load(k0 + RK);
#pragma unroll
        for (int kc = 0; kc < RK; kc += 4) {
            double a[2], b[4];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = As[kc + lk][wr + 16 * i + li];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kc + lk][wc + 16 * j + li];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = mfma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
    // epilogue: the 4 parts of a translation sit in lanes li..li+3 (li % 4 == 0)
    const int64_t ndst = (int64_t)C.Lt * (C.Lt + 2);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                double c0 = acc[i][j][v];
                double c1 = __shfl_down(c0, 1, 64);
                double c2 = __shfl_down(c0, 2, 64);
                double c3 = __shfl_down(c0, 3, 64);
                int gc = col0 + wc + 16 * j + li;
                int lr = wr + 16 * i + 4 * v + lk;
                int gr = row0 + lr;
                if ((li & 3) != 0 || gc >= NC || gr >= M) continue;
                int t = gc >> 2;
                int m = gr - n;
                double2 p = ph[lr];
                double2 *dst = dst_pool + dst_off[C.col_begin + t];
                int64_t o = fidx(n, m);
                // (c0, c1) = back-rotated z_+, (c2, c3) = z_-: a = (z_+ + z_-)/2, b = (z_+ - z_-)/2
                double ar = 0.5 * (c0 + c2), ai = 0.5 * (c1 + c3);
                double br = 0.5 * (c0 - c2), bi = 0.5 * (c1 - c3);
                double2 da = dst[o], db = dst[o + ndst];
                dst[o] = make_double2(da.x + p.x * ar - p.y * ai, da.y + p.x * ai + p.y * ar);
                dst[o + ndst] = make_double2(db.x + p.x * br - p.y * bi, db.y + p.x * bi + p.y * br);
            }
}

void launch_class(const ClassDesc &h, const ClassDesc *d_cls, const int64_t *src_off, const int64_t *dst_off,
                  const double2 *src_pool, double2 *dst_pool, const double *wd, const double2 *wax, double *wy,
                  double *wz, int stage, hipStream_t s) {
    const int NC = 4 * h.ncols;
    const int K = h.K;
    if (stage == 0) {
        int mt = (2 * (h.Ls < K ? h.Ls : K) + 1 + RM - 1) / RM;
        dim3 grid((NC + RN - 1) / RN, mt, h.Ls);
        hipLaunchKernelGGL(k_class_rot_fwd, grid, dim3(256), 0, s, d_cls, src_off, src_pool, wd, wy);
    } else if (stage == 1) {
        int mt = (h.Lt + AM - 1) / AM;
        dim3 grid((h.ncols + AN - 1) / AN, mt, 2 * K + 1);
        hipLaunchKernelGGL(k_class_axial, grid, dim3(256), 0, s, d_cls, wax, wy, wz);
    } else {
        int mt = (2 * h.Lt + 1 + RM - 1) / RM;
        dim3 grid((NC + RN - 1) / RN, mt, h.Lt);
        hipLaunchKernelGGL(k_class_rot_back, grid, dim3(256), 0, s, d_cls, dst_off, wd, wz, dst_pool);
    }
}

}  // namespace ifs
