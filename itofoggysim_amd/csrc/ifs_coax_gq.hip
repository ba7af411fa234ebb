// Coaxial translation tables by Gauss-Legendre quadrature (stable at high order).
//
// The scalar coaxial coefficients of order m are integrals over mu = cos(theta):
//   alpha^m(n, nu) = i^{n - nu} * int_{-1}^{1} Pb_n^m(mu) Pb_nu^m(mu) T(mu) dmu
// with Pb_n^m the orthonormal associated Legendre functions and
//   T(mu) = exp(i t mu)                                   (regular, j_n)
//   T(mu) = sum_{p=0}^{P} i^p (2p+1) h1_p(t) P_p(mu)      (outgoing -> regular, P = Lt + 1 + Ls)
// (for j_p the second sum is exp(i t mu) itself: the plane-wave expansion).  The
// integrand is a polynomial of degree n + nu + P (or an entire function whose
// Legendre coefficients (2p+1) j_p(t) vanish to rounding beyond p ~ t + 8 t^(1/3) + 30),
// so Q Gauss nodes with 2Q > n + nu + P integrate it to rounding.  Unlike the
// three-term recurrences in nu and m, whose parasitic solutions grow without bound
// once the orders exceed ~100 near the turning point n ~ t, every term here is
// bounded.
//
// For the outgoing kind, |h1_p(t)| grows steeply for p > t, and a single T would
// put the rounding of its largest terms (~1e10 at p = 2L) into every entry.  T is
// therefore split in bands of p, a new band starting wherever |h1_p| exceeds 1e3
// times its value at the band's first order; band b contributes only to the
// entries with n + nu >= its first order (below that its integral vanishes
// exactly), so each entry sees rounding of at most 1e3 times the size of its own
// largest terms.
//
// Pipeline per table: T bands at the nodes (k_gq_T), Pb_n^m at the nodes for a
// batch of orders m (k_gq_leg), alpha for the batch (k_gq_alpha, a banded
// V^T diag(T) V product), then the vector coefficients and the P = A + B,
// Q = A - B blocks in the layout of the recurrence generator (k_gq_vector).
#include <hip/hip_runtime.h>
#include <math.h>

#include <vector>

#include "ifs_internal.h"

namespace ifs {

namespace {

__device__ __forceinline__ double2 cmulg(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// i^k * v
__device__ __forceinline__ double2 ipowg(int k, double2 v) {
    switch (k & 3) {
        case 0: return v;
        case 1: return make_double2(-v.y, v.x);
        case 2: return make_double2(-v.x, -v.y);
        default: return make_double2(v.y, -v.x);
    }
}
__device__ __forceinline__ double coef_ag(int n, int m) {
    int am = m < 0 ? -m : m;
    if (n < am) return 0.0;
    return sqrt((double)(n + 1 + am) * (double)(n + 1 - am) / ((double)(2 * n + 1) * (double)(2 * n + 3)));
}

// h1_p(t) = j_p(t) + i y_p(t), p = 0..P, one thread: j by Miller's downward
// recurrence, y upward.
__device__ void hankel_into(double x, int P, double2 *out) {
    const double big = (double)P > x ? (double)P : x;
    const int nstart = (int)big + 30 + (int)sqrt(40.0 * big);
    double fp1 = 0.0, f = 1e-200;
    for (int n = nstart; n >= 1; --n) {
        double fm1 = (2.0 * n + 1.0) / x * f - fp1;
        fp1 = f;
        f = fm1;
        if (n - 1 <= P) out[n - 1].x = f;
        if (fabs(f) > 1e200) {
            f *= 1e-200;
            fp1 *= 1e-200;
            if (n - 1 <= P)
                for (int q = n - 1; q <= P; ++q) out[q].x *= 1e-200;
        }
    }
    const double j0 = sin(x) / x;
    const double j1 = sin(x) / (x * x) - cos(x) / x;
    const double scale = (P >= 1 && fabs(j1) > fabs(j0)) ? j1 / out[1].x : j0 / out[0].x;
    for (int n = 0; n <= P; ++n) out[n].x *= scale;
    double ym1 = -cos(x) / x;
    out[0].y = ym1;
    if (P >= 1) {
        double y = -cos(x) / (x * x) - sin(x) / x;
        out[1].y = y;
        for (int n = 1; n < P; ++n) {
            double yp = (2.0 * n + 1.0) / x * y - ym1;
            ym1 = y;
            y = yp;
            out[n + 1].y = y;
        }
    }
}

}  // namespace

// T bands at the nodes, weights folded in: Tb[b * Q + q] = w_q T_b(mu_q).
//   kind 1: one band, exp(i t mu);  kind 3: band b = terms p in [lo[b], lo[b+1]).
__global__ void k_gq_T(const GqPlan g, const double *__restrict__ mu, const double *__restrict__ w,
                       double2 *__restrict__ Tb) {
    extern __shared__ double2 hsh[];
    const int Q = g.Q;
    if (g.kind == 3) {
        if (threadIdx.x == 0) hankel_into(g.t, g.P, hsh);
        __syncthreads();
    }
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= Q) return;
    const double x = mu[q], wq = w[q];
    if (g.kind != 3) {
        double s, c;
        sincos(g.t * x, &s, &c);
        Tb[q] = make_double2(wq * c, wq * s);
        return;
    }
    double2 acc = make_double2(0.0, 0.0);
    int band = 0;
    double pm1 = 0.0, pc = 1.0;  // P_{p-1}, P_p
    for (int p = 0; p <= g.P; ++p) {
        if (p >= 1) {
            double pn = (p == 1) ? x : ((2.0 * p - 1.0) * x * pc - (p - 1.0) * pm1) / p;
            pm1 = pc;
            pc = pn;
        }
        while (band + 1 < g.nb && p >= g.lo[band + 1]) {
            Tb[(int64_t)band * Q + q] = make_double2(wq * acc.x, wq * acc.y);
            acc = make_double2(0.0, 0.0);
            ++band;
        }
        double2 cp = ipowg(p, hsh[p]);
        const double f = (2.0 * p + 1.0) * pc;
        acc.x += f * cp.x;
        acc.y += f * cp.y;
    }
    Tb[(int64_t)band * Q + q] = make_double2(wq * acc.x, wq * acc.y);
    for (int b = band + 1; b < g.nb; ++b) Tb[(int64_t)b * Q + q] = make_double2(0.0, 0.0);
}

// V[(j * (Nr + 1) + n) * Q + q] = Pb_n^{m0 + j}(mu_q), n = 0..Nr (0 for n < m).
__global__ void k_gq_leg(int m0, int nm, int Nr, int Q, const double *__restrict__ mu, double *__restrict__ V) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (q >= Q || j >= nm) return;
    const int m = m0 + j;
    double *out = V + (int64_t)j * (Nr + 1) * Q + q;
    for (int n = 0; n < m && n <= Nr; ++n) out[(int64_t)n * Q] = 0.0;
    if (m > Nr) return;
    const double x = mu[q];
    const double s2 = fmax(1.0 - x * x, 0.0);
    // Pb_m^m = sqrt((2m+1)/2 (2m)!/(2^m m!)^2) (1 - x^2)^(m/2), in logs (underflow -> 0)
    const double lc = 0.5 * (log((2.0 * m + 1.0) * 0.5) + lgamma(2.0 * m + 1.0) -
                             2.0 * (m * 0.6931471805599453 + lgamma(m + 1.0)));
    double pmm = (m == 0) ? exp(lc) : (s2 > 0.0 ? exp(lc + 0.5 * m * log(s2)) : 0.0);
    out[(int64_t)m * Q] = pmm;
    if (m + 1 > Nr) return;
    double p1 = sqrt(2.0 * m + 3.0) * x * pmm;
    out[(int64_t)(m + 1) * Q] = p1;
    double p2 = pmm;
    const double mm = (double)m * m;
    for (int n = m + 2; n <= Nr; ++n) {
        const double nn = (double)n;
        const double a = sqrt((4.0 * nn * nn - 1.0) / (nn * nn - mm));
        const double b = sqrt(((nn - 1.0) * (nn - 1.0) - mm) * (2.0 * nn + 1.0) / ((nn * nn - mm) * (2.0 * nn - 3.0)));
        const double pn = a * x * p1 - b * p2;
        out[(int64_t)n * Q] = pn;
        p2 = p1;
        p1 = pn;
    }
}

// alpha[(j * (Lt + 2) + n) * (Ls + 1) + nu] = i^{n-nu} sum_b [n + nu >= lo_b] sum_q V_n V_nu Tb_b,
// n = 0..Lt+1, nu = 0..Ls.  Block: 64 x 64 entries, thread: 4 x 4, inner step 16 nodes.
constexpr int GT = 64;
constexpr int GK = 16;
__global__ __launch_bounds__(256) void k_gq_alpha(const GqPlan g, int nm, const double *__restrict__ V,
                                                  const double2 *__restrict__ Tb, double2 *__restrict__ alpha) {
    const int Lt = g.Lt, Ls = g.Ls, Nr = g.Nr, Q = g.Q;
    const int j = blockIdx.z;
    const int r0 = blockIdx.y * GT, c0 = blockIdx.x * GT;
    const int NR = Lt + 2, NCc = Ls + 1;
    if (r0 >= NR || c0 >= NCc) return;
    __shared__ double Vr[GK][GT + 1];
    __shared__ double Vc[GK][GT + 1];
    __shared__ double2 Ts[GK];
    const double *Vm = V + (int64_t)j * (Nr + 1) * Q;
    const int tid = threadIdx.x;
    const int tr = (tid >> 4) * 4, tc = (tid & 15) * 4;
    double ar[4][4], ai[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) ar[a][b] = ai[a][b] = 0.0;
    const int rmax = min(NR - 1, r0 + GT - 1), cmax = min(NCc - 1, c0 + GT - 1);
    for (int band = 0; band < g.nb; ++band) {
        const int lo = g.lo[band];
        if (lo > rmax + cmax) break;
        const double2 *T = Tb + (int64_t)band * Q;
        double br[4][4], bi[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) br[a][b] = bi[a][b] = 0.0;
        for (int k0 = 0; k0 < Q; k0 += GK) {
            // stage 16 nodes x 64 rows and 64 columns (coalesced along q)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                int idx = tid + 256 * e;  // 64 x 16
                int rr = idx >> 4, qq = idx & 15;
                int gq = k0 + qq;
                int gr = r0 + rr, gc = c0 + rr;
                Vr[qq][rr] = (gq < Q && gr <= Nr && gr < NR) ? Vm[(int64_t)gr * Q + gq] : 0.0;
                Vc[qq][rr] = (gq < Q && gc <= Nr && gc < NCc) ? Vm[(int64_t)gc * Q + gq] : 0.0;
            }
            if (tid < GK) Ts[tid] = (k0 + tid < Q) ? T[k0 + tid] : make_double2(0.0, 0.0);
            __syncthreads();
#pragma unroll 4
            for (int qq = 0; qq < GK; ++qq) {
                const double2 tq = Ts[qq];
                double vr[4], vc[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    vr[a] = Vr[qq][tr + a] * tq.x;
                    vc[a] = Vc[qq][tc + a];
                }
                double vi[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) vi[a] = Vr[qq][tr + a] * tq.y;
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        br[a][b] = fma(vr[a], vc[b], br[a][b]);
                        bi[a][b] = fma(vi[a], vc[b], bi[a][b]);
                    }
            }
            __syncthreads();
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int n = r0 + tr + a, nu = c0 + tc + b;
                if (n + nu >= lo) {
                    ar[a][b] += br[a][b];
                    ai[a][b] += bi[a][b];
                }
            }
    }
    double2 *out = alpha + (int64_t)j * NR * NCc;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int n = r0 + tr + a, nu = c0 + tc + b;
            if (n < NR && nu < NCc) out[(int64_t)n * NCc + nu] = ipowg(((n - nu) % 4 + 4) & 3, make_double2(ar[a][b], ai[a][b]));
        }
}

// Vector coaxial coefficients (normalised basis) from alpha, orders m0..m0+nm-1:
//   A_m(n,nu) = i^{nu-n} [sqrt(n(n+1)) al(n,nu) + t a_n sqrt(n/(n+1)) al(n+1,nu)
//                         + t a_{n-1} sqrt((n+1)/n) al(n-1,nu)] / sqrt(nu(nu+1))
//   B_m(n,nu) = i^{nu-n+1} t m al(n,nu) / sqrt(n(n+1) nu(nu+1))
// stored as P = A + B, Q = A - B (order -m: A unchanged, B negated), in the block
// layout of the recurrence generator.
__global__ void k_gq_vector(int m0, int nm, int Lt, int Ls, int K, double t, const double2 *__restrict__ alpha,
                            double2 *__restrict__ wax) {
    const int j = blockIdx.y;
    const int m = m0 + j;
    if (j >= nm || m > K) return;
    const int q = m < 1 ? 1 : m;
    const int nrow = Lt - q + 1, ncol = Ls - q + 1;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (nrow <= 0 || ncol <= 0 || e >= (int64_t)nrow * ncol) return;
    const int n = q + (int)(e / ncol), nu = q + (int)(e % ncol);
    const int NR = Lt + 2, NCc = Ls + 1;
    const double2 *al = alpha + (int64_t)j * NR * NCc;
    const double2 c0 = al[(int64_t)n * NCc + nu];
    const double2 cp = al[(int64_t)(n + 1) * NCc + nu];
    const double2 cm = (n - 1 >= m) ? al[(int64_t)(n - 1) * NCc + nu] : make_double2(0.0, 0.0);
    const double nn = (double)n;
    const double inv_nu = 1.0 / sqrt((double)nu * (nu + 1.0));
    const double f0 = sqrt(nn * (nn + 1.0)), fp = t * coef_ag(n, m) * sqrt(nn / (nn + 1.0)),
                 fm = t * coef_ag(n - 1, m) * sqrt((nn + 1.0) / nn);
    double2 s = make_double2(f0 * c0.x + fp * cp.x + fm * cm.x, f0 * c0.y + fp * cp.y + fm * cm.y);
    const double2 av = ipowg(((nu - n) % 4 + 4) & 3, make_double2(inv_nu * s.x, inv_nu * s.y));
    const double fb = t * (double)m * inv_nu / sqrt(nn * (nn + 1.0));
    const double2 bv = ipowg(((nu - n + 1) % 4 + 4) & 3, make_double2(fb * c0.x, fb * c0.y));
    const double2 pv = make_double2(av.x + bv.x, av.y + bv.y), qv = make_double2(av.x - bv.x, av.y - bv.y);
    double2 *Ap = wax + ax_off(m, Lt, Ls, K);
    double2 *Bp = Ap + ax_block(m, Lt, Ls);
    const int64_t o = (int64_t)(n - q) * ncol + (nu - q);
    Ap[o] = pv;
    Bp[o] = qv;
    if (m > 0) {
        double2 *An = wax + ax_off(-m, Lt, Ls, K);
        double2 *Bn = An + ax_block(-m, Lt, Ls);
        An[o] = qv;
        Bn[o] = pv;
    }
}

// Regular (j_n) coaxial tables of the generic translations (sphere <-> box, one
// table per task) by the same quadrature, batched over the tasks of a chunk.  The
// tasks share one node set and one table V of Pb_n^m at the nodes (k_gq_leg); per
// task only the plane-wave factor w_q exp(i t mu_q) differs:
//   I(n, nu) = sum_q V_n(mu_q) V_nu(mu_q) w_q exp(i t mu_q),   alpha = i^{n-nu} I.
// I is symmetric in (n, nu), so a tile runs its 256 rows along the larger of the two
// index ranges (n <= Lt + 1 for sphere -> box, nu <= Ls for box -> sphere) and its 16
// columns along the smaller one.  Thread: 4 rows x 4 columns; inner step 16 nodes.
// alpha of task T, order m: walpha[T.off_s + (m (Lt + 2) + n) (Ls + 1) + nu].
constexpr int GB_R = 256;
constexpr int GB_C = 16;
constexpr int GB_K = 16;
__global__ __launch_bounds__(256) void k_gqb_alpha(const Task *__restrict__ tasks, int Q, int Nr, int rtiles,
                                                   int ctiles, const double *__restrict__ mu,
                                                   const double *__restrict__ w, const double *__restrict__ V,
                                                   double2 *__restrict__ walpha) {
    const int ntile = rtiles * ctiles;
    const int ti = blockIdx.x / ntile, tile = blockIdx.x - ti * ntile;
    const int m = blockIdx.y;
    const Task T = tasks[ti];
    if (m > T.K) return;
    const int Lt = T.Lt, Ls = T.Ls;
    const bool rows_n = Lt + 1 >= Ls;
    const int nbig = rows_n ? Lt + 1 : Ls;
    const int nsml = rows_n ? Ls : Lt + 1;
    const int r0 = m + (tile / ctiles) * GB_R, c0 = m + (tile % ctiles) * GB_C;
    if (r0 > nbig || c0 > nsml) return;
    __shared__ double Vr[GB_K][GB_R];
    __shared__ double Vc[GB_K][GB_C];
    __shared__ double2 Ts[GB_K];
    const double *Vm = V + (int64_t)m * (Nr + 1) * Q;
    const int tid = threadIdx.x;
    const int rg = (tid >> 2) * 4, cg = (tid & 3) * 4;
    double br[4][4], bi[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) br[a][b] = bi[a][b] = 0.0;
    for (int k0 = 0; k0 < Q; k0 += GB_K) {
#pragma unroll
        for (int e = 0; e < GB_R * GB_K / 256; ++e) {
            const int idx = tid + 256 * e;  // 256 rows x 16 nodes, nodes fastest
            const int rr = idx >> 4, qq = idx & 15;
            const int gq = k0 + qq, gr = r0 + rr;
            Vr[qq][rr] = (gq < Q && gr <= nbig) ? Vm[(int64_t)gr * Q + gq] : 0.0;
        }
        {
            const int cc = tid >> 4, qq = tid & 15;
            const int gq = k0 + qq, gc = c0 + cc;
            Vc[qq][cc] = (gq < Q && gc <= nsml) ? Vm[(int64_t)gc * Q + gq] : 0.0;
        }
        if (tid < GB_K) {
            const int gq = k0 + tid;
            if (gq < Q) {
                double s, c;
                sincos(T.t * mu[gq], &s, &c);
                Ts[tid] = make_double2(w[gq] * c, w[gq] * s);
            } else {
                Ts[tid] = make_double2(0.0, 0.0);
            }
        }
        __syncthreads();
#pragma unroll 4
        for (int qq = 0; qq < GB_K; ++qq) {
            const double2 tq = Ts[qq];
            double vr[4], vi[4], vc[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const double v = Vr[qq][rg + a];
                vr[a] = v * tq.x;
                vi[a] = v * tq.y;
                vc[a] = Vc[qq][cg + a];
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    br[a][b] = fma(vr[a], vc[b], br[a][b]);
                    bi[a][b] = fma(vi[a], vc[b], bi[a][b]);
                }
        }
        __syncthreads();
    }
    double2 *al = walpha + T.off_s + (int64_t)m * (Lt + 2) * (Ls + 1);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int ri = r0 + rg + a, ci = c0 + cg + b;
            if (ri > nbig || ci > nsml) continue;
            const int n = rows_n ? ri : ci, nu = rows_n ? ci : ri;
            al[(int64_t)n * (Ls + 1) + nu] = ipowg(((n - nu) % 4 + 4) & 3, make_double2(br[a][b], bi[a][b]));
        }
}

// Vector coefficients of the batched tables (as k_gq_vector, per task and order m).
__global__ void k_gqb_vector(const Task *__restrict__ tasks, int eblk, const double2 *__restrict__ walpha,
                             double2 *__restrict__ wax) {
    const int ti = blockIdx.x / eblk, eb = blockIdx.x - ti * eblk;
    const int m = blockIdx.y;
    const Task T = tasks[ti];
    if (m > T.K) return;
    const int Lt = T.Lt, Ls = T.Ls, K = T.K;
    const int q = m < 1 ? 1 : m;
    const int nrow = Lt - q + 1, ncol = Ls - q + 1;
    const int64_t e = (int64_t)eb * blockDim.x + threadIdx.x;
    if (nrow <= 0 || ncol <= 0 || e >= (int64_t)nrow * ncol) return;
    const int n = q + (int)(e / ncol), nu = q + (int)(e % ncol);
    const int NCc = Ls + 1;
    const double2 *al = walpha + T.off_s + (int64_t)m * (Lt + 2) * NCc;
    const double2 c0 = al[(int64_t)n * NCc + nu];
    const double2 cp = al[(int64_t)(n + 1) * NCc + nu];
    const double2 cm = (n - 1 >= m) ? al[(int64_t)(n - 1) * NCc + nu] : make_double2(0.0, 0.0);
    const double t = T.t;
    const double nn = (double)n;
    const double inv_nu = 1.0 / sqrt((double)nu * (nu + 1.0));
    const double f0 = sqrt(nn * (nn + 1.0)), fp = t * coef_ag(n, m) * sqrt(nn / (nn + 1.0)),
                 fm = t * coef_ag(n - 1, m) * sqrt((nn + 1.0) / nn);
    const double2 s = make_double2(f0 * c0.x + fp * cp.x + fm * cm.x, f0 * c0.y + fp * cp.y + fm * cm.y);
    const double2 av = ipowg(((nu - n) % 4 + 4) & 3, make_double2(inv_nu * s.x, inv_nu * s.y));
    const double fb = t * (double)m * inv_nu / sqrt(nn * (nn + 1.0));
    const double2 bv = ipowg(((nu - n + 1) % 4 + 4) & 3, make_double2(fb * c0.x, fb * c0.y));
    const double2 pv = make_double2(av.x + bv.x, av.y + bv.y), qv = make_double2(av.x - bv.x, av.y - bv.y);
    double2 *Ap = wax + T.off_ax + ax_off(m, Lt, Ls, K);
    double2 *Bp = Ap + ax_block(m, Lt, Ls);
    const int64_t o = (int64_t)(n - q) * ncol + (nu - q);
    Ap[o] = pv;
    Bp[o] = qv;
    if (m > 0) {
        double2 *An = wax + T.off_ax + ax_off(-m, Lt, Ls, K);
        double2 *Bn = An + ax_block(-m, Lt, Ls);
        An[o] = qv;
        Bn[o] = pv;
    }
}

// The sphere <-> box coaxial shift applied through the quadrature, without forming
// its table.  With I(n, nu) = sum_q V_n(mu_q) V_nu(mu_q) T_q (T_q = w_q exp(i t mu_q))
// and the vector relation of k_gqb_vector, a vector y of source orders shifts as
//   u(n') = sum_q V_n'(mu_q) T_q g(q),   g(q) = sum_nu V_nu(mu_q) y(nu) / sqrt(nu (nu + 1)),
//   (P y)(n) = f0 u(n) + i fp u(n+1) - i fm u(n-1) + i fb u(n),   (Q y)(n): -i fb u(n),
// f0 = sqrt(n(n+1)), fp = t a_n sqrt(n/(n+1)), fm = t a_{n-1} sqrt((n+1)/n) (n-1 >= m),
// fb = t m / sqrt(n(n+1)).  That is (Lt + Ls) Q multiply-adds per vector and order
// instead of the table's Lt Ls Q.  Block per (task, m, sign of k): the two vectors
// y_+(., k), y_-(., k) of order k = +-m; z_+(n, k) = P_m y_+ and z_-(n, k) = Q_m y_- for
// k = +m, the roles of P and Q exchanged for k = -m.  V: the chunk's table of Pb_n^m at
// the nodes (k_gq_leg), [m][n][q].  LDS: 2 (Nr + 2) + 2 Q complex.
__global__ __launch_bounds__(256) void k_gqb_apply(const Task *__restrict__ tasks, int Q, int Nr,
                                                   const double *__restrict__ mu, const double *__restrict__ w,
                                                   const double *__restrict__ V, const double2 *__restrict__ wy,
                                                   double2 *__restrict__ wz) {
    extern __shared__ double2 shq[];
    const Task T = tasks[blockIdx.x];
    const int m = blockIdx.y;
    const int sgn = blockIdx.z == 0 ? 1 : -1;
    if (m > T.K || (m == 0 && sgn < 0)) return;
    const int Lt = T.Lt, Ls = T.Ls, K = T.K;
    const int k = sgn * m;
    const int q0 = m < 1 ? 1 : m;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    double2 *yv = shq;                 // [2][Nr + 2]: y / sqrt(nu(nu+1)), then u
    double2 *gq = shq + 2 * (Nr + 2);  // [2][Q]
    const int64_t ny = rows_size(Ls, K), nz = rows_size(Lt, K);
    const double2 *ya = wy + T.off_y, *yb = ya + ny;
    for (int nu = q0 + tid; nu <= Ls; nu += blockDim.x) {
        const int kk = nu < K ? nu : K;
        const int64_t iy = rows_off(nu, K) + k + kk;
        const double inv = 1.0 / sqrt((double)nu * (nu + 1.0));
        const double2 a = ya[iy], b = yb[iy];
        yv[nu] = make_double2(inv * a.x, inv * a.y);
        yv[(Nr + 2) + nu] = make_double2(inv * b.x, inv * b.y);
    }
    __syncthreads();
    const double *Vm = V + (int64_t)m * (Nr + 1) * Q;
    // g(q) = T_q sum_nu V_nu(q) yv(nu): threads along q (coalesced table rows)
    for (int q = tid; q < Q; q += blockDim.x) {
        double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
        for (int nu = q0; nu <= Ls; ++nu) {
            const double v = Vm[(int64_t)nu * Q + q];
            const double2 a = yv[nu], b = yv[(Nr + 2) + nu];
            ar = fma(v, a.x, ar);
            ai = fma(v, a.y, ai);
            br = fma(v, b.x, br);
            bi = fma(v, b.y, bi);
        }
        double sn, cs;
        sincos(T.t * mu[q], &sn, &cs);
        const double tr = w[q] * cs, ti = w[q] * sn;
        gq[q] = make_double2(tr * ar - ti * ai, tr * ai + ti * ar);
        gq[Q + q] = make_double2(tr * br - ti * bi, tr * bi + ti * br);
    }
    __syncthreads();
    // u(n') = sum_q V_n'(q) g(q), n' = m .. Lt + 1: a wave per row, lanes along q
    for (int n = m + wv; n <= Lt + 1; n += (int)(blockDim.x >> 6)) {
        double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
        const double *Vn = Vm + (int64_t)n * Q;
        for (int q = lane; q < Q; q += 64) {
            const double v = Vn[q];
            const double2 a = gq[q], b = gq[Q + q];
            ar = fma(v, a.x, ar);
            ai = fma(v, a.y, ai);
            br = fma(v, b.x, br);
            bi = fma(v, b.y, bi);
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            ar += __shfl_xor(ar, o, 64);
            ai += __shfl_xor(ai, o, 64);
            br += __shfl_xor(br, o, 64);
            bi += __shfl_xor(bi, o, 64);
        }
        if (lane == 0) {
            yv[n] = make_double2(ar, ai);
            yv[(Nr + 2) + n] = make_double2(br, bi);
        }
    }
    __syncthreads();
    // z_+(n, k), z_-(n, k): P_m / Q_m on y_+ / y_- for k = +m, exchanged for k = -m
    double2 *z = wz + T.off_z;
    const double t = T.t;
    for (int n = q0 + tid; n <= Lt; n += blockDim.x) {
        const double nn = (double)n;
        const double f0 = sqrt(nn * (nn + 1.0)), fp = t * coef_ag(n, m) * sqrt(nn / (nn + 1.0));
        const double fm = n - 1 >= m ? t * coef_ag(n - 1, m) * sqrt((nn + 1.0) / nn) : 0.0;
        const double fb = t * (double)m / sqrt(nn * (nn + 1.0));
        const int kk = n < K ? n : K;
        const int64_t e = rows_off(n, K) + k + kk;
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            const double2 *u = yv + v * (Nr + 2);
            const double2 u0 = u[n], up = u[n + 1];
            const double2 um = n - 1 >= m ? u[n - 1] : make_double2(0.0, 0.0);
            // A part: f0 u(n) + i fp u(n+1) - i fm u(n-1); B part: i fb u(n)
            const double2 av = make_double2(f0 * u0.x - fp * up.y + fm * um.y, f0 * u0.y + fp * up.x - fm * um.x);
            const double2 bv = make_double2(-fb * u0.y, fb * u0.x);
            // P = A + B on y_+ for k = +m and on y_- for k = -m; Q = A - B otherwise
            const bool plus = (v == 0) == (sgn > 0);
            z[e + v * nz] = plus ? make_double2(av.x + bv.x, av.y + bv.y) : make_double2(av.x - bv.x, av.y - bv.y);
        }
    }
}

size_t gqb_apply_lds(int Q, int Nr) { return sizeof(double2) * (2 * (size_t)(Nr + 2) + 2 * (size_t)Q); }

void launch_gqb_legendre(const GqbChunk &g, const double *mu, double *V, hipStream_t s) {
    hipLaunchKernelGGL(k_gq_leg, dim3((g.Q + 255) / 256, g.maxK + 1), dim3(256), 0, s, 0, g.maxK + 1, g.Nr, g.Q, mu,
                       V);
}

void launch_gqb_apply(const Task *tasks, int ntasks, const GqbChunk &g, const double *mu, const double *w,
                      const double *V, const double2 *wy, double2 *wz, hipStream_t s) {
    if (ntasks <= 0) return;
    hipLaunchKernelGGL(k_gqb_apply, dim3((unsigned)ntasks, g.maxK + 1, 2), dim3(256), gqb_apply_lds(g.Q, g.Nr), s,
                       tasks, g.Q, g.Nr, mu, w, V, wy, wz);
}

int gqb_nodes(double t, int Lt, int Ls) {
    const int Pdeg = (int)ceil(t + 8.0 * cbrt(t) + 30.0);
    return (Lt + 1 + Ls + Pdeg) / 2 + 4;
}

void launch_gen_axial_gqb(const Task *tasks, int ntasks, const GqbChunk &g, const double *mu, const double *w,
                          double *V, double2 *walpha, double2 *wax, hipStream_t s) {
    if (ntasks <= 0) return;
    hipLaunchKernelGGL(k_gq_leg, dim3((g.Q + 255) / 256, g.maxK + 1), dim3(256), 0, s, 0, g.maxK + 1, g.Nr, g.Q, mu,
                       V);
    const int rtiles = (g.max_big + 1 + GB_R - 1) / GB_R, ctiles = (g.max_small + 1 + GB_C - 1) / GB_C;
    hipLaunchKernelGGL(k_gqb_alpha, dim3((unsigned)ntasks * rtiles * ctiles, g.maxK + 1), dim3(256), 0, s, tasks,
                       g.Q, g.Nr, rtiles, ctiles, mu, w, V, walpha);
    const int eblk = (int)((g.max_ent + 255) / 256);
    hipLaunchKernelGGL(k_gqb_vector, dim3((unsigned)ntasks * eblk, g.maxK + 1), dim3(256), 0, s, tasks, eblk, walpha,
                       wax);
}

// |h1_p(t)|, p = 0..P, on the host (Miller's recurrence for j_p, upward for y_p).
static void hankel_abs_host(double x, int P, std::vector<double> &out) {
    std::vector<double> j(P + 1, 0.0), y(P + 1, 0.0);
    const double big = (double)P > x ? (double)P : x;
    const int nstart = (int)big + 30 + (int)sqrt(40.0 * big);
    double fp1 = 0.0, f = 1e-200;
    for (int n = nstart; n >= 1; --n) {
        double fm1 = (2.0 * n + 1.0) / x * f - fp1;
        fp1 = f;
        f = fm1;
        if (n - 1 <= P) j[n - 1] = f;
        if (fabs(f) > 1e200) {
            f *= 1e-200;
            fp1 *= 1e-200;
            if (n - 1 <= P)
                for (int q = n - 1; q <= P; ++q) j[q] *= 1e-200;
        }
    }
    const double j0 = sin(x) / x, j1 = sin(x) / (x * x) - cos(x) / x;
    const double scale = (P >= 1 && fabs(j1) > fabs(j0)) ? j1 / j[1] : j0 / j[0];
    y[0] = -cos(x) / x;
    if (P >= 1) y[1] = -cos(x) / (x * x) - sin(x) / x;
    for (int n = 1; n < P; ++n) y[n + 1] = (2.0 * n + 1.0) / x * y[n] - y[n - 1];
    out.resize(P + 1);
    for (int n = 0; n <= P; ++n) out[n] = hypot(j[n] * scale, y[n]);
}

// Plan of one table: node count, band split, workspace.
GqPlan gq_plan(int kind, double t, int Lt, int Ls, int K, size_t ws_cap_bytes) {
    GqPlan g;
    g.kind = kind;
    g.t = t;
    g.Lt = Lt;
    g.Ls = Ls;
    g.K = K;
    int Pdeg;
    g.nb = 1;
    g.lo[0] = 0;
    if (kind == 3) {
        g.P = Lt + 1 + Ls;
        Pdeg = g.P;
        std::vector<double> h;
        hankel_abs_host(t, g.P, h);
        double ref = h[0];
        for (int p = 1; p <= g.P; ++p) {
            if (h[p] > 1e3 * ref && g.nb < GQ_MAXB) {
                g.lo[g.nb++] = p;
                ref = h[p];
            }
        }
    } else {
        g.P = 0;
        Pdeg = (int)ceil(t + 8.0 * cbrt(t) + 30.0);
    }
    g.Q = (Lt + 1 + Ls + Pdeg) / 2 + 4;
    g.Nr = (Lt + 1 > Ls ? Lt + 1 : Ls);
    const size_t per_m = (size_t)(g.Nr + 1) * g.Q * 8 + (size_t)(Lt + 2) * (Ls + 1) * 16;
    int mb = (int)(ws_cap_bytes / (per_m ? per_m : 1));
    if (mb < 1) mb = 1;
    if (mb > K + 1) mb = K + 1;
    g.mb = mb;
    g.ws_bytes = per_m * mb + (size_t)g.nb * g.Q * 16;
    return g;
}

void launch_gen_axial_gq(const GqPlan &g, const double *mu, const double *w, void *ws, double2 *wax, hipStream_t s) {
    double2 *Tb = (double2 *)ws;
    double *V = (double *)(Tb + (size_t)g.nb * g.Q);
    double2 *alpha = (double2 *)(V + (size_t)g.mb * (g.Nr + 1) * g.Q);
    {
        size_t shm = g.kind == 3 ? sizeof(double2) * (size_t)(g.P + 1) : 0;
        hipLaunchKernelGGL(k_gq_T, dim3((g.Q + 255) / 256), dim3(256), shm, s, g, mu, w, Tb);
    }
    for (int m0 = 0; m0 <= g.K; m0 += g.mb) {
        const int nm = (g.K + 1 - m0) < g.mb ? (g.K + 1 - m0) : g.mb;
        hipLaunchKernelGGL(k_gq_leg, dim3((g.Q + 255) / 256, nm), dim3(256), 0, s, m0, nm, g.Nr, g.Q, mu, V);
        dim3 ga((g.Ls + 1 + GT - 1) / GT, (g.Lt + 2 + GT - 1) / GT, nm);
        hipLaunchKernelGGL(k_gq_alpha, ga, dim3(256), 0, s, g, nm, V, Tb, alpha);
        const int64_t ent = (int64_t)g.Lt * g.Ls;
        hipLaunchKernelGGL(k_gq_vector, dim3((unsigned)((ent + 255) / 256), nm), dim3(256), 0, s, m0, nm, g.Lt, g.Ls,
                           g.K, g.t, alpha, wax);
    }
}

}  // namespace ifs
