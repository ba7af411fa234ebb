// HIP kernels (gfx950) for the multi-sphere scattering matvec.
//
// Translation pipeline per task (source expansion -> target expansion), all FP64:
//   gen_wigner : Wigner d^n_{km}(theta) tables by the three-term recurrence in n
//                (folded E/O tables for the rotations, see ifs_internal.h)
//   gen_sect   : sectorial scalar coaxial coefficients (recurrence in m)
//   gen_axial  : vector coaxial coefficients A_k(n,nu), B_k(n,nu) (recurrence in nu)
//   rot_fwd    : y_+- = d(theta) diag(e^{i m phi}) (x_a +- x_b)   (forward rotation)
//   axial_apply: z_+ = P_k y_+, z_- = Q_k y_- per coaxial order k, P = A + B, Q = A - B
//                (the [A B; B A] coaxial operator decoupled)
//   rot_back   : target_a/b += diag(e^{-i m phi}) d(theta)^T (z_+ +- z_-)/2, summed over
//                the tasks of one target in a fixed order (deterministic, no atomics)
// The algorithm replaces the Gaunt/3j evaluation of Xu's A/B of
// libmath/src/math/lib_math_vector_spherical_harmonics.f90:1354-1585 by the
// rotation / coaxial-shift factorisation; both produce the same operator.
#include <hip/hip_runtime.h>
#include <math.h>

#include "ifs_internal.h"
#include "ifs_wigner.h"

namespace ifs {

// ------------------------------------------------------------------------
// complex helpers
// ------------------------------------------------------------------------
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cscale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
    double d = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
// i^p * a
__device__ __forceinline__ double2 ipow_mul(int p, double2 a) {
    switch (p & 3) {
    case 0: return a;
    case 1: return make_double2(-a.y, a.x);
    case 2: return make_double2(-a.x, -a.y);
    default: return make_double2(a.y, -a.x);
    }
}
__device__ __forceinline__ int fidx(int n, int m) { return n * n - 1 + m + n; }

// a^m_n = sqrt((n+1+|m|)(n+1-|m|)/((2n+1)(2n+3))), 0 for n < |m|
__device__ __forceinline__ double coef_a(int n, int m) {
    int am = m < 0 ? -m : m;
    if (n < am) return 0.0;
    return sqrt((double)(n + 1 + am) * (double)(n + 1 - am) / ((double)(2 * n + 1) * (double)(2 * n + 3)));
}
// b^m_n = sign(m) sqrt((n-m-1)(n-m)/((2n-1)(2n+1))) for |m| <= n, 0 otherwise (sign(0) = +)
__device__ __forceinline__ double coef_b(int n, int m) {
    int am = m < 0 ? -m : m;
    if (am > n) return 0.0;
    double p = (double)(n - m - 1) * (double)(n - m);
    if (p <= 0.0) return 0.0;
    double v = sqrt(p / ((double)(2 * n - 1) * (double)(2 * n + 1)));
    return m >= 0 ? v : -v;
}

// ------------------------------------------------------------------------
// Wigner d tables
// ------------------------------------------------------------------------
// Folded tables (layout at fold_off): threads over (k, m), k = 0..K, m = 0..L; each
// runs the recurrences of d_{k,m} and d_{k,-m} side by side from n0 = max(k, m).
template <typename TW>
__global__ void k_gen_wigner_fold(const Task *__restrict__ tasks, TW *__restrict__ wd) {
    const Task T = tasks[blockIdx.y];
    const int K = T.K;
    const int L = T.Ls > T.Lt ? T.Ls : T.Lt;
    const int width = L + 1;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (K + 1) * width) return;
    const int k = p / width;
    const int m = p % width;
    const int n0 = k > m ? k : m;
    const double cb = T.cth;
    const double ch = sqrt(fmax(0.0, 0.5 * (1.0 + cb)));
    const double sh = sqrt(fmax(0.0, 0.5 * (1.0 - cb)));
    const double sgm = (m & 1) ? -1.0 : 1.0;
    // Degree n's E block starts at fold_off(n, K) and holds (min(n,K)+1) rows of n+1;
    // the O block follows it.  The entry (k, m) of degree n+1 lies
    // 2 (min(n,K)+1)(n+1) + k past that of degree n, so a pointer is carried from
    // degree to degree (32-bit increments: box orders below 32000).
    const int nfirst = n0 == 0 ? 1 : n0;
    TW *pe = wd + T.off_d + fold_off(nfirst, K) + (int64_t)k * (nfirst + 1) + m;
    const bool m0 = m == 0, k0 = k == 0;
    auto put = [&](int n, double dp, double dm) {
        const int kk = n < K ? n : K;
        const double e = m0 ? 2.0 * dp : dp + sgm * dm;
        const double od = (m0 || k0) ? 0.0 : dp - sgm * dm;
        const int ob = (kk + 1) * (n + 1);
        pe[0] = (TW)e;
        pe[ob] = (TW)od;
        pe += 2 * ob + k;
    };
    double p1, p2, q1, q2;  // chains d_{k,m} and d_{k,-m}: values at n-1 and n-2
    int nstart;
    if (n0 == 0) {
        p2 = q2 = 1.0;  // d^0_{00}
        p1 = q1 = cb;   // d^1_{00}
        if (L >= 1) put(1, p1, q1);
        nstart = 2;
    } else {
        p2 = q2 = 0.0;
        p1 = wigner_seed(k, m, ch, sh);
        q1 = (m == 0) ? p1 : wigner_seed(k, -m, ch, sh);
        put(n0, p1, q1);
        nstart = n0 + 1;
    }
    // The three-term recurrence of wigner_step for both chains: they share
    // s_n = sqrt((n^2 - k^2)(n^2 - m^2)), which also serves as the d^{n-2} coefficient
    // of the next degree.  One reciprocal square root per degree gives both the divisor
    // and s_n: with x = (n^2 - k^2)(n^2 - m^2) > 0 (n > max(k, m)) and
    // r = 1 / sqrt((n-1)^2 x),  1 / ((n-1) s_n) = r  and  s_n = (n-1) x r.
    const double k2 = (double)k * k, m2 = (double)m * m, km = (double)k * m;
    double sprev = sqrt(fmax(0.0, ((nstart - 1.0) * (nstart - 1.0) - k2) * ((nstart - 1.0) * (nstart - 1.0) - m2)));
    for (int n = nstart; n <= L; ++n) {
        const double nn = (double)n;
        const double x = (nn * nn - k2) * (nn * nn - m2);
        const double nm1x = (nn - 1.0) * x;
        const double inv = rsqrt((nn - 1.0) * nm1x);
        const double scur = nm1x * inv;
        const double a = (2.0 * nn - 1.0), b = nn * (nn - 1.0) * cb, c2 = nn * sprev;
        const double pn = (a * (b - km) * p1 - c2 * p2) * inv;
        const double qn = (m == 0) ? pn : (a * (b + km) * q1 - c2 * q2) * inv;
        put(n, pn, qn);
        p2 = p1;
        p1 = pn;
        q2 = q1;
        q1 = qn;
        sprev = scur;
    }
}

void launch_gen_wigner_fold(const Task *tasks, int ntasks, int maxpairs, double *wd, hipStream_t s) {
    if (ntasks <= 0) return;
    dim3 grid((maxpairs + 255) / 256, ntasks);
    hipLaunchKernelGGL(k_gen_wigner_fold<double>, grid, dim3(256), 0, s, tasks, wd);
}

void launch_gen_wigner_fold_f32(const Task *tasks, int ntasks, int maxpairs, float *wd, hipStream_t s) {
    if (ntasks <= 0) return;
    dim3 grid((maxpairs + 255) / 256, ntasks);
    hipLaunchKernelGGL(k_gen_wigner_fold<float>, grid, dim3(256), 0, s, tasks, wd);
}

// ------------------------------------------------------------------------
// Radial functions (one thread)
// ------------------------------------------------------------------------
// j_n(x), n = 0..nmax, by Miller's downward recurrence written straight into
// out[n].x, normalised with j_0 or j_1; y_n upward into out[n].y (kind 3).
__device__ void radial_into(double x, int nmax, int kind, double2 *out) {
    if (x == 0.0) {
        for (int n = 0; n <= nmax; ++n) out[n] = make_double2(n == 0 ? 1.0 : 0.0, 0.0);
        return;
    }
    const double big = (double)nmax > x ? (double)nmax : x;
    const int nstart = (int)big + 30 + (int)sqrt(40.0 * big);
    double fp1 = 0.0, f = 1e-200;  // f_{n+1}, f_n
    for (int n = nstart; n >= 1; --n) {
        double fm1 = (2.0 * n + 1.0) / x * f - fp1;
        fp1 = f;
        f = fm1;
        if (n - 1 <= nmax) out[n - 1].x = f;
        if (fabs(f) > 1e200) {
            f *= 1e-200;
            fp1 *= 1e-200;
            if (n - 1 <= nmax)
                for (int q = n - 1; q <= nmax; ++q) out[q].x *= 1e-200;
        }
    }
    const double j0 = sin(x) / x;
    const double j1 = sin(x) / (x * x) - cos(x) / x;
    const double scale = (nmax >= 1 && fabs(j1) > fabs(j0)) ? j1 / out[1].x : j0 / out[0].x;
    for (int n = 0; n <= nmax; ++n) out[n].x *= scale;
    if (kind == 3) {
        double ym1 = -cos(x) / x;
        out[0].y = ym1;
        if (nmax >= 1) {
            double y = -cos(x) / (x * x) - sin(x) / x;
            out[1].y = y;
            for (int n = 1; n < nmax; ++n) {
                double yp = (2.0 * n + 1.0) / x * y - ym1;
                ym1 = y;
                y = yp;
                out[n + 1].y = y;
            }
        }
    } else {
        for (int n = 0; n <= nmax; ++n) out[n].y = 0.0;
    }
}

// ------------------------------------------------------------------------
// Sectorial scalar coaxial columns alpha^m(n, nu = m), m = 0..K
//   alpha^0(n,0)     = (-1)^n sqrt(2n+1) f_n(t)
//   alpha^{m+1}(n,m+1) = [b^{-m-1}_n alpha^m(n-1,m) - b^m_{n+1} alpha^m(n+1,m)] / b^{-m-1}_{m+1}
// Column m is valid for n in [m, nmax - m]; stored with stride nmax+2.
// ------------------------------------------------------------------------
__global__ void k_gen_sect(const Task *__restrict__ tasks, int max_nmax, double2 *__restrict__ wsect) {
    extern __shared__ double2 sh[];
    const Task T = tasks[blockIdx.x];
    const int nmax = T.Lt + T.Ls + 2;
    double2 *cur = sh;
    double2 *nxt = sh + (max_nmax + 2);
    if (threadIdx.x == 0) radial_into(T.t, nmax, T.kind, cur);
    __syncthreads();
    for (int n = threadIdx.x; n <= nmax; n += blockDim.x) {
        double s = sqrt(2.0 * n + 1.0) * ((n & 1) ? -1.0 : 1.0);
        cur[n] = cscale(s, cur[n]);
    }
    __syncthreads();
    double2 *out = wsect + T.off_s;
    const int stride = nmax + 2;
    for (int m = 0; m <= T.K; ++m) {
        const int top = nmax - m;
        for (int n = m + threadIdx.x; n <= top; n += blockDim.x) out[(int64_t)m * stride + n] = cur[n];
        if (m == T.K) break;
        const double den = coef_b(m + 1, -m - 1);
        for (int n = m + 1 + threadIdx.x; n <= top - 1; n += blockDim.x) {
            double2 lo = cscale(coef_b(n, -m - 1), cur[n - 1]);
            double2 hi = cscale(coef_b(n + 1, m), cur[n + 1]);
            nxt[n] = cscale(1.0 / den, csub(lo, hi));
        }
        __syncthreads();
        double2 *tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
}

void launch_gen_sect(const Task *tasks, int ntasks, int max_nmax, double2 *wsect, hipStream_t s) {
    if (ntasks <= 0) return;
    size_t shm = sizeof(double2) * 2 * (size_t)(max_nmax + 2);
    hipLaunchKernelGGL(k_gen_sect, dim3(ntasks), dim3(256), shm, s, tasks, max_nmax, wsect);
}

// ------------------------------------------------------------------------
// Vector coaxial tables for one (task, m >= 0): recurrence in nu
//   alpha(n,nu+1) = [a_{nu-1} alpha(n,nu-1) - a_n alpha(n+1,nu) + a_{n-1} alpha(n-1,nu)] / a_nu
// and the vector relation (normalised basis)
//   A_m(n,nu) = i^{nu-n} [sqrt(n(n+1)) al(n,nu) + t a_n sqrt(n/(n+1)) al(n+1,nu)
//                         + t a_{n-1} sqrt((n+1)/n) al(n-1,nu)] / sqrt(nu(nu+1))
//   B_m(n,nu) = i^{nu-n+1} t m al(n,nu) / sqrt(n(n+1) nu(nu+1))
// Order -m shares A and has B negated.  The table stores P = A + B and Q = A - B:
// with y_+- = y_a +- y_b the shift decouples into z_+ = P y_+ and z_- = Q y_-.
// ------------------------------------------------------------------------
__global__ void k_gen_axial(const Task *__restrict__ tasks, int max_nmax, const double2 *__restrict__ wsect,
                            double2 *__restrict__ wax) {
    extern __shared__ double2 sh[];
    const Task T = tasks[blockIdx.y];
    const int m = blockIdx.x;
    if (m > T.K) return;
    const int nmax = T.Lt + T.Ls + 2;
    const int Lt = T.Lt, Ls = T.Ls, K = T.K;
    const double t = T.t;
    double2 *P = sh;
    double2 *C = sh + (max_nmax + 2);
    double2 *N = sh + 2 * (max_nmax + 2);
    const double2 *src = wsect + T.off_s + (int64_t)m * (nmax + 2);
    int top = nmax - m;
    for (int n = threadIdx.x; n <= nmax + 1; n += blockDim.x) {
        P[n] = make_double2(0.0, 0.0);
        C[n] = (n >= m && n <= top) ? src[n] : make_double2(0.0, 0.0);
        N[n] = make_double2(0.0, 0.0);
    }
    __syncthreads();
    const int q = m < 1 ? 1 : m;
    double2 *Ap = wax + T.off_ax + ax_off(m, Lt, Ls, K);
    double2 *Bp = Ap + ax_block(m, Lt, Ls);
    double2 *An = wax + T.off_ax + ax_off(-m, Lt, Ls, K);
    double2 *Bn = An + ax_block(-m, Lt, Ls);
    const int ncol = Ls - q + 1;
    // A/B of target order n, source order nu from alpha(n-1|n|n+1, nu)
    auto emit = [&](int n, int nu, double2 c0, double2 cp, double2 cm) {
        const double nn = (double)n;
        const double inv_nu = 1.0 / sqrt((double)nu * (nu + 1));
        double2 s = cscale(sqrt(nn * (nn + 1.0)), c0);
        s = cadd(s, cscale(t * coef_a(n, m) * sqrt(nn / (nn + 1.0)), cp));
        s = cadd(s, cscale(t * coef_a(n - 1, m) * sqrt((nn + 1.0) / nn), cm));
        double2 av = ipow_mul((nu - n) & 3, cscale(inv_nu, s));
        double2 bv = ipow_mul((nu - n + 1) & 3, cscale(t * (double)m * inv_nu / sqrt(nn * (nn + 1.0)), c0));
        int64_t o = (int64_t)(n - q) * ncol + (nu - q);
        double2 pv = cadd(av, bv), qv = csub(av, bv);
        Ap[o] = pv;  // P_m = A_m + B_m
        Bp[o] = qv;  // Q_m = A_m - B_m
        if (m > 0) {  // order -m: A unchanged, B negated
            An[o] = qv;
            Bn[o] = pv;
        }
    };
    if (Ls > Lt) {
        // The recurrence advances the column index and its error grows with the
        // number of steps, so it runs over the smaller order: columns c = m..Lt+1
        // (target orders) with the source orders as rows, and
        //   alpha(n, nu) = (-1)^{n+nu} alpha(nu, n).
        for (int c = m;; ++c) {
            if (c <= Lt) {
                const double ac = coef_a(c, m);
                const double acm1 = coef_a(c - 1, m);
                for (int r = m + threadIdx.x; r <= top - 1; r += blockDim.x) {
                    double2 v = cscale(acm1, P[r]);
                    v = csub(v, cscale(coef_a(r, m), C[r + 1]));
                    if (r - 1 >= m) v = cadd(v, cscale(coef_a(r - 1, m), C[r - 1]));
                    N[r] = cscale(1.0 / ac, v);
                }
                for (int r = top + threadIdx.x; r <= nmax + 1; r += blockDim.x) N[r] = make_double2(0.0, 0.0);
            }
            __syncthreads();
            if (c >= q && c <= Lt) {
                const int n = c;
                for (int nu = q + threadIdx.x; nu <= Ls; nu += blockDim.x) {
                    const double sg = ((n + nu) & 1) ? -1.0 : 1.0;  // (-1)^{n+nu}
                    double2 c0 = cscale(sg, C[nu]);
                    double2 cp = cscale(-sg, N[nu]);
                    double2 cm = (n - 1 >= m) ? cscale(-sg, P[nu]) : make_double2(0.0, 0.0);
                    emit(n, nu, c0, cp, cm);
                }
            }
            if (c >= Lt) break;
            __syncthreads();
            double2 *tmp = P;
            P = C;
            C = N;
            N = tmp;
            top -= 1;
        }
        return;
    }
    for (int nu = m; nu <= Ls; ++nu) {
        if (nu >= 1) {
            for (int n = q + threadIdx.x; n <= Lt; n += blockDim.x) {
                double2 c0 = C[n];
                double2 cp = C[n + 1];
                double2 cm = (n - 1 >= m) ? C[n - 1] : make_double2(0.0, 0.0);
                emit(n, nu, c0, cp, cm);
            }
        }
        if (nu == Ls) break;
        // advance nu -> nu+1 on rows [m, top-1]
        const double anu = coef_a(nu, m);
        const double anum1 = coef_a(nu - 1, m);
        for (int n = m + threadIdx.x; n <= top - 1; n += blockDim.x) {
            double2 v = cscale(anum1, P[n]);
            v = csub(v, cscale(coef_a(n, m), C[n + 1]));
            if (n - 1 >= m) v = cadd(v, cscale(coef_a(n - 1, m), C[n - 1]));
            N[n] = cscale(1.0 / anu, v);
        }
        __syncthreads();
        // rotate P <- C <- N
        double2 *tmp = P;
        P = C;
        C = N;
        N = tmp;
        top -= 1;
        for (int n = top + 1 + threadIdx.x; n <= nmax + 1; n += blockDim.x) C[n] = make_double2(0.0, 0.0);
        __syncthreads();
    }
}

// ------------------------------------------------------------------------
// Coaxial shift with its coefficients generated on the fly (one right-hand side):
// the recurrence of k_gen_axial, and instead of storing P_{+-m}, Q_{+-m} each
// coefficient is applied at once,
//   z_+(n, m) += P_m y_+(nu, m),   z_-(n, m) += Q_m y_-(nu, m),
//   z_+(n,-m) += Q_m y_+(nu,-m),   z_-(n,-m) += P_m y_-(nu,-m)   (P_{-m} = Q_m, Q_{-m} = P_m),
// so the tables never reach HBM.  Block per (order m >= 0, task) of 64 or 256
// threads; rows of a thread: FR at most (target orders up to FR * threads - 1).
// ------------------------------------------------------------------------
// Every factor that depends on one order only (a^m_r, 1/a^m_r, 1/sqrt(nu(nu+1)) and the
// four target-order factors of the vector relation) is tabulated in LDS once per block,
// and the source rows y(nu, +-m) are staged there too: the serial recurrence then runs
// on LDS operands and multiply-adds, with no square root, division or global load in
// its steps.  LDS: 168 (max_nmax + 2) + 256 bytes.  FR: target rows per thread (1 when
// every target order fits one row per thread, which halves the registers).
template <int FR>
__global__ void k_axial_fused(const Task *__restrict__ tasks, int max_nmax, const double2 *__restrict__ wsect,
                              const double2 *__restrict__ wy, double2 *__restrict__ wz) {
    extern __shared__ double2 sh[];
    const Task T = tasks[blockIdx.y];
    const int m = blockIdx.x;
    if (m > T.K) return;
    const int nmax = T.Lt + T.Ls + 2;
    const int Lt = T.Lt, Ls = T.Ls, K = T.K;
    const double t = T.t;
    const int M2 = max_nmax + 2;
    double2 *P = sh;
    double2 *C = sh + M2;
    double2 *N = sh + 2 * M2;
    double2 *red = sh + 3 * M2;  // 4 waves x 4 partial sums
    double2 *Y0 = red + 16, *Y1 = Y0 + M2, *Y2 = Y1 + M2, *Y3 = Y2 + M2;  // y_+(nu,m) y_-(nu,m) y_+(nu,-m) y_-(nu,-m)
    double *KA = (double *)(Y3 + M2);  // a^m_r
    double *IKA = KA + M2;             // 1 / a^m_r
    double *RN = IKA + M2;             // 1 / sqrt(nu(nu+1))
    double *K0 = RN + M2;              // sqrt(n(n+1))
    double *KP = K0 + M2;              // t a^m_n sqrt(n/(n+1))
    double *KM = KP + M2;              // t a^m_{n-1} sqrt((n+1)/n)
    double *KB = KM + M2;              // t m / sqrt(n(n+1))
    const double2 *src = wsect + T.off_s + (int64_t)m * (nmax + 2);
    const int q = m < 1 ? 1 : m;
    const int64_t ny = rows_size(Ls, K), nz = rows_size(Lt, K);
    const double2 *yp = wy + T.off_y;  // y_+ rows (nu, k); y_- at + ny
    double2 *zp = wz + T.off_z;        // z_+ rows (n, k); z_- at + nz
    auto yrow = [&](int nu, int k) { return rows_off(nu, K) + k + (nu < K ? nu : K); };
    int top = nmax - m;
    const double2 zero = make_double2(0.0, 0.0);
    for (int n = threadIdx.x; n <= nmax + 1; n += blockDim.x) {
        P[n] = zero;
        C[n] = (n >= m && n <= top) ? src[n] : zero;
        N[n] = zero;
        const double a = coef_a(n, m);
        KA[n] = a;
        IKA[n] = a != 0.0 ? 1.0 / a : 0.0;
        if (n >= q && n <= Ls) {
            RN[n] = 1.0 / sqrt((double)n * (n + 1));
            const int64_t iy = yrow(n, m);
            Y0[n] = yp[iy];
            Y1[n] = yp[iy + ny];
            if (m > 0) {
                const int64_t iyn = yrow(n, -m);
                Y2[n] = yp[iyn];
                Y3[n] = yp[iyn + ny];
            } else {
                Y2[n] = zero;
                Y3[n] = zero;
            }
        }
        if (n >= q && n <= Lt) {
            const double nn = (double)n;
            K0[n] = sqrt(nn * (nn + 1.0));
            KP[n] = t * a * sqrt(nn / (nn + 1.0));
            KM[n] = t * coef_a(n - 1, m) * sqrt((nn + 1.0) / nn);
            KB[n] = t * (double)m / sqrt(nn * (nn + 1.0));
        }
    }
    __syncthreads();
    // P = A + B, Q = A - B of target order n, source order nu
    auto coef = [&](int n, double inv_nu, int nu, double2 c0, double2 cp, double2 cm, double2 &pv, double2 &qv) {
        double2 s = cscale(K0[n], c0);
        s = cadd(s, cscale(KP[n], cp));
        s = cadd(s, cscale(KM[n], cm));
        double2 av = ipow_mul((nu - n) & 3, cscale(inv_nu, s));
        double2 bv = ipow_mul((nu - n + 1) & 3, cscale(KB[n] * inv_nu, c0));
        pv = cadd(av, bv);
        qv = csub(av, bv);
    };
    if (Ls > Lt) {
        // recurrence over the target order (see k_gen_axial); the sum over nu is a
        // block reduction per target row
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        for (int c = m;; ++c) {
            if (c <= Lt) {
                const double iac = IKA[c];
                const double acm1 = c >= 1 ? KA[c - 1] : 0.0;
                for (int r = m + threadIdx.x; r <= top - 1; r += blockDim.x) {
                    double2 v = cscale(acm1, P[r]);
                    v = csub(v, cscale(KA[r], C[r + 1]));
                    if (r - 1 >= m) v = cadd(v, cscale(KA[r - 1], C[r - 1]));
                    N[r] = cscale(iac, v);
                }
                for (int r = top + threadIdx.x; r <= nmax + 1; r += blockDim.x) N[r] = zero;
            }
            __syncthreads();
            if (c >= q && c <= Lt) {
                const int n = c;
                double2 s4[4] = {zero, zero, zero, zero};
                for (int nu = q + threadIdx.x; nu <= Ls; nu += blockDim.x) {
                    const double sg = ((n + nu) & 1) ? -1.0 : 1.0;
                    double2 c0 = cscale(sg, C[nu]);
                    double2 cp = cscale(-sg, N[nu]);
                    double2 cm = (n - 1 >= m) ? cscale(-sg, P[nu]) : zero;
                    double2 pv, qv;
                    coef(n, RN[nu], nu, c0, cp, cm, pv, qv);
                    s4[0] = cadd(s4[0], cmul(pv, Y0[nu]));
                    s4[1] = cadd(s4[1], cmul(qv, Y1[nu]));
                    s4[2] = cadd(s4[2], cmul(qv, Y2[nu]));
                    s4[3] = cadd(s4[3], cmul(pv, Y3[nu]));
                }
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int o = 32; o >= 1; o >>= 1) {
                        s4[i].x += __shfl_xor(s4[i].x, o, 64);
                        s4[i].y += __shfl_xor(s4[i].y, o, 64);
                    }
                if (lane == 0)
#pragma unroll
                    for (int i = 0; i < 4; ++i) red[4 * w + i] = s4[i];
                __syncthreads();
                if (threadIdx.x == 0) {
                    double2 tot[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        tot[i] = red[i];
                        for (int ww = 1; ww < (int)(blockDim.x >> 6); ++ww) tot[i] = cadd(tot[i], red[4 * ww + i]);
                    }
                    const int64_t iz = rows_off(n, K) + m + (n < K ? n : K);
                    zp[iz] = tot[0];
                    zp[iz + nz] = tot[1];
                    if (m > 0) {
                        const int64_t izn = rows_off(n, K) - m + (n < K ? n : K);
                        zp[izn] = tot[2];
                        zp[izn + nz] = tot[3];
                    }
                }
            }
            if (c >= Lt) break;
            __syncthreads();
            double2 *tmp = P;
            P = C;
            C = N;
            N = tmp;
            top -= 1;
        }
        return;
    }
    double2 acc[FR][4];
#pragma unroll
    for (int j = 0; j < FR; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[j][i] = zero;
    for (int nu = m; nu <= Ls; ++nu) {
        if (nu >= 1) {
            const double inv_nu = RN[nu];
            const double2 ypp = Y0[nu], ypm = Y1[nu], ynp = Y2[nu], ynm = Y3[nu];
#pragma unroll
            for (int j = 0; j < FR; ++j) {
                const int n = q + threadIdx.x + j * blockDim.x;
                if (n > Lt) break;
                double2 c0 = C[n];
                double2 cp = C[n + 1];
                double2 cm = (n - 1 >= m) ? C[n - 1] : zero;
                double2 pv, qv;
                coef(n, inv_nu, nu, c0, cp, cm, pv, qv);
                acc[j][0] = cadd(acc[j][0], cmul(pv, ypp));
                acc[j][1] = cadd(acc[j][1], cmul(qv, ypm));
                acc[j][2] = cadd(acc[j][2], cmul(qv, ynp));
                acc[j][3] = cadd(acc[j][3], cmul(pv, ynm));
            }
        }
        if (nu == Ls) break;
        const double ianu = IKA[nu];
        const double anum1 = nu >= 1 ? KA[nu - 1] : 0.0;
        for (int n = m + threadIdx.x; n <= top - 1; n += blockDim.x) {
            double2 v = cscale(anum1, P[n]);
            v = csub(v, cscale(KA[n], C[n + 1]));
            if (n - 1 >= m) v = cadd(v, cscale(KA[n - 1], C[n - 1]));
            N[n] = cscale(ianu, v);
        }
        __syncthreads();
        double2 *tmp = P;
        P = C;
        C = N;
        N = tmp;
        top -= 1;
        for (int n = top + 1 + threadIdx.x; n <= nmax + 1; n += blockDim.x) C[n] = zero;
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < FR; ++j) {
        const int n = q + threadIdx.x + j * blockDim.x;
        if (n > Lt) break;
        const int kk = n < K ? n : K;
        const int64_t iz = rows_off(n, K) + m + kk;
        zp[iz] = acc[j][0];
        zp[iz + nz] = acc[j][1];
        if (m > 0) {
            const int64_t izn = rows_off(n, K) - m + kk;
            zp[izn] = acc[j][2];
            zp[izn + nz] = acc[j][3];
        }
    }
}

void launch_axial_fused(const Task *tasks, int ntasks, int maxK, int max_nmax, int max_Lt, const double2 *wsect,
                        const double2 *wy, double2 *wz, hipStream_t s) {
    if (ntasks <= 0) return;
    size_t shm = 168 * (size_t)(max_nmax + 2) + 256;
    // sphere-sphere orders (<= ~60 each) fill one wave; box orders take four
    const int threads = max_nmax + 2 <= 128 ? 64 : 256;
    if (max_Lt <= threads)
        hipLaunchKernelGGL(k_axial_fused<1>, dim3(maxK + 1, ntasks), dim3(threads), shm, s, tasks, max_nmax, wsect,
                           wy, wz);
    else
        hipLaunchKernelGGL(k_axial_fused<4>, dim3(maxK + 1, ntasks), dim3(threads), shm, s, tasks, max_nmax, wsect,
                           wy, wz);
}

void launch_gen_axial(const Task *tasks, int ntasks, int maxK, int max_nmax, const double2 *wsect, double2 *wax,
                      hipStream_t s) {
    if (ntasks <= 0) return;
    size_t shm = sizeof(double2) * 3 * (size_t)(max_nmax + 2);
    hipLaunchKernelGGL(k_gen_axial, dim3(maxK + 1, ntasks), dim3(256), shm, s, tasks, max_nmax, wsect, wax);
}

// ------------------------------------------------------------------------
// Forward rotation with the folded tables (E, O at fold_off, see ifs_internal.h):
//   S_k = sum_{m>=0} E[k][m] s_m,  T_k = sum_{m>=0} O[k][m] t_m,  k = 0..min(nu,K)
//   s_m = u_m + (-1)^m u_{-m} (s_0 = u_0), t_m = u_m - (-1)^m u_{-m}, u_m = e^{i m phi} x_m
//   y_k = (S_k + T_k)/2,  y_{-k} = (-1)^k (S_k - T_k)/2
// for both parts x_+- = x_a +- x_b.  One thread per (nu, k >= 0).
// ------------------------------------------------------------------------
__global__ void k_rot_fwd(const Task *__restrict__ tasks, const double2 *__restrict__ src_pool,
                          const double *__restrict__ wd, double2 *__restrict__ wy) {
    extern __shared__ double2 ph[];  // e^{i m phi}, m = 0..Ls
    const Task T = tasks[blockIdx.y];
    const int Ls = T.Ls, K = T.K;
    for (int m = threadIdx.x; m <= Ls; m += blockDim.x) {
        double sn, cs;
        sincos((double)m * T.phi, &sn, &cs);
        ph[m] = make_double2(cs, sn);
    }
    __syncthreads();
    const int64_t nf = frow_size(Ls, K);
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nf) return;
    int nu, k;
    {
        int64_t lim = frow_size(K, K);  // rows of degrees 1..K
        if (e < lim) {
            nu = (int)((sqrt(8.0 * (double)e + 9.0) - 3.0) * 0.5);  // nu(nu+3)/2 <= e
            while ((int64_t)(nu + 1) * (nu + 4) / 2 <= e) ++nu;
            while (nu > 0 && (int64_t)nu * (nu + 3) / 2 > e) --nu;
            k = (int)(e - (int64_t)nu * (nu + 3) / 2);
            nu += 1;
        } else {
            nu = K + 1 + (int)((e - lim) / (K + 1));
            k = (int)((e - lim) % (K + 1));
        }
    }
    const int kk = nu < K ? nu : K;
    const double *E = wd + T.off_d + fold_off(nu, K) + (int64_t)k * (nu + 1);
    const double *O = E + (int64_t)(kk + 1) * (nu + 1);
    const double2 *xa = src_pool + T.src + fidx(nu, 0);  // xa[m], m = -nu..nu
    const double2 *xb = xa + (int64_t)Ls * (Ls + 2);
    double2 Sp = make_double2(0.0, 0.0), Sm = Sp, Tp = Sp, Tm = Sp;
    for (int m = 0; m <= nu; ++m) {
        const double2 p = ph[m];
        double2 ap = xa[m], bp = xb[m];
        double2 xp = cadd(ap, bp), xm = csub(ap, bp);  // x_+, x_- at m
        double2 up = cmul(p, xp), um = cmul(p, xm);
        const double ev = E[m], ov = O[m];
        if (m == 0) {
            Sp = cadd(Sp, cscale(ev, up));
            Sm = cadd(Sm, cscale(ev, um));
            continue;
        }
        const double2 pc = make_double2(p.x, -p.y);
        double2 an = xa[-m], bn = xb[-m];
        double2 vp = cmul(pc, cadd(an, bn)), vm = cmul(pc, csub(an, bn));  // u_{-m}
        const double sg = (m & 1) ? -1.0 : 1.0;
        Sp = cadd(Sp, cscale(ev, cadd(up, cscale(sg, vp))));
        Sm = cadd(Sm, cscale(ev, cadd(um, cscale(sg, vm))));
        Tp = cadd(Tp, cscale(ov, csub(up, cscale(sg, vp))));
        Tm = cadd(Tm, cscale(ov, csub(um, cscale(sg, vm))));
    }
    const int64_t nent = rows_size(Ls, K);
    double2 *y = wy + T.off_y;
    const int64_t rp = rows_off(nu, K) + kk + k;
    y[rp] = cscale(0.5, cadd(Sp, Tp));            // y_+ at (nu, k)
    y[rp + nent] = cscale(0.5, cadd(Sm, Tm));     // y_- at (nu, k)
    if (k > 0) {
        const double f = (k & 1) ? -0.5 : 0.5;
        const int64_t rm = rows_off(nu, K) + kk - k;
        y[rm] = cscale(f, csub(Sp, Tp));
        y[rm + nent] = cscale(f, csub(Sm, Tm));
    }
}

// The same forward rotation for high source orders (box to sphere): one block per
// (task, degree nu).  The folded column entries s_m, t_m of both parts go to LDS once;
// each wave then takes rows k and reads its E and O rows with the lanes along m
// (coalesced), reducing over m with shuffles.
__global__ void k_rot_fwd_wide(const Task *__restrict__ tasks, const double2 *__restrict__ src_pool,
                               const double *__restrict__ wd, double2 *__restrict__ wy) {
    extern __shared__ double2 stm[];
    const Task T = tasks[blockIdx.x];
    const int nu = blockIdx.y + 1;
    if (nu > T.Ls) return;
    const int Ls = T.Ls, K = T.K;
    const int kk = nu < K ? nu : K;
    const int W = nu + 1;
    double2 *sp = stm, *sm = stm + W, *tp = stm + 2 * W, *tm = stm + 3 * W;
    const double2 *xa = src_pool + T.src + fidx(nu, 0);  // xa[m], m = -nu..nu
    const double2 *xb = xa + (int64_t)Ls * (Ls + 2);
    for (int m = threadIdx.x; m <= nu; m += blockDim.x) {
        double sn, cs;
        sincos((double)m * T.phi, &sn, &cs);
        const double2 p = make_double2(cs, sn);
        const double2 ap = xa[m], bp = xb[m];
        const double2 up = cmul(p, cadd(ap, bp)), um = cmul(p, csub(ap, bp));
        if (m == 0) {
            sp[0] = up;
            sm[0] = um;
            tp[0] = tm[0] = make_double2(0.0, 0.0);
            continue;
        }
        const double2 pc = make_double2(p.x, -p.y);
        const double2 an = xa[-m], bn = xb[-m];
        const double2 vp = cmul(pc, cadd(an, bn)), vm = cmul(pc, csub(an, bn));
        const double sg = (m & 1) ? -1.0 : 1.0;
        sp[m] = cadd(up, cscale(sg, vp));
        sm[m] = cadd(um, cscale(sg, vm));
        tp[m] = csub(up, cscale(sg, vp));
        tm[m] = csub(um, cscale(sg, vm));
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    const int64_t nent = rows_size(Ls, K);
    double2 *y = wy + T.off_y;
    for (int k = threadIdx.x >> 6; k <= kk; k += nw) {
        const double *E = wd + T.off_d + fold_off(nu, K) + (int64_t)k * W;
        const double *O = E + (int64_t)(kk + 1) * W;
        double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // S+, S-, T+, T- (re, im)
        for (int m = lane; m <= nu; m += 64) {
            const double e = E[m], o = O[m];
            a[0] += e * sp[m].x;
            a[1] += e * sp[m].y;
            a[2] += e * sm[m].x;
            a[3] += e * sm[m].y;
            a[4] += o * tp[m].x;
            a[5] += o * tp[m].y;
            a[6] += o * tm[m].x;
            a[7] += o * tm[m].y;
        }
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) a[i] += __shfl_xor(a[i], off, 64);
        if (lane == 0) {
            const double2 Sp = make_double2(a[0], a[1]), Sm = make_double2(a[2], a[3]);
            const double2 Tp = make_double2(a[4], a[5]), Tm = make_double2(a[6], a[7]);
            const int64_t rp = rows_off(nu, K) + kk + k;
            y[rp] = cscale(0.5, cadd(Sp, Tp));
            y[rp + nent] = cscale(0.5, cadd(Sm, Tm));
            if (k > 0) {
                const double f = (k & 1) ? -0.5 : 0.5;
                const int64_t rm = rows_off(nu, K) + kk - k;
                y[rm] = cscale(f, csub(Sp, Tp));
                y[rm + nent] = cscale(f, csub(Sm, Tm));
            }
        }
    }
}

void launch_rot_fwd_wide(const Task *tasks, int ntasks, int max_Ls, const double2 *src_pool, const double *wd,
                         double2 *wy, hipStream_t s) {
    if (ntasks <= 0 || max_Ls <= 0) return;
    size_t shm = sizeof(double2) * 4 * (size_t)(max_Ls + 1);
    hipLaunchKernelGGL(k_rot_fwd_wide, dim3(ntasks, max_Ls), dim3(256), shm, s, tasks, src_pool, wd, wy);
}

void launch_rot_fwd(const Task *tasks, int ntasks, int64_t max_rows, int max_Ls, const double2 *src_pool,
                    const double *wd, double2 *wy, hipStream_t s) {
    if (ntasks <= 0) return;
    // max_rows bounds the folded rows (about half the unfolded ones)
    dim3 grid((unsigned)((max_rows + 255) / 256), ntasks);
    size_t shm = sizeof(double2) * (size_t)(max_Ls + 1);
    hipLaunchKernelGGL(k_rot_fwd, grid, dim3(256), shm, s, tasks, src_pool, wd, wy);
}

// ------------------------------------------------------------------------
// Coaxial shift in the decoupled form: z_+(n,k) = sum_nu P_k(n,nu) y_+(nu,k),
//                                     z_-(n,k) = sum_nu Q_k(n,nu) y_-(nu,k)
// ------------------------------------------------------------------------
__global__ void k_axial_apply(const Task *__restrict__ tasks, const double2 *__restrict__ wax,
                              const double2 *__restrict__ wy, double2 *__restrict__ wz) {
    const Task T = tasks[blockIdx.y];
    const int Lt = T.Lt, Ls = T.Ls, K = T.K;
    const int64_t nent = rows_size(Lt, K);
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nent) return;
    int n, k;
    {
        int64_t lim = (int64_t)(K + 1) * (K + 1) - 1;
        if (e < lim) {
            n = (int)sqrt((double)(e + 1));
            while ((int64_t)(n + 1) * (n + 1) - 1 <= e) ++n;
            while ((int64_t)n * n - 1 > e) --n;
        } else {
            n = K + 1 + (int)((e - lim) / (2 * K + 1));
        }
        int kk = n < K ? n : K;
        k = (int)(e - rows_off(n, K)) - kk;
    }
    const int q = (k < 0 ? -k : k) < 1 ? 1 : (k < 0 ? -k : k);
    const int ncol = Ls - q + 1;
    const double2 *Ab = wax + T.off_ax + ax_off(k, Lt, Ls, K) + (int64_t)(n - q) * ncol;
    const double2 *Bb = Ab + ax_block(k, Lt, Ls);
    const int64_t ny = rows_size(Ls, K);
    const double2 *ya = wy + T.off_y;
    const double2 *yb = ya + ny;
    double2 za = make_double2(0.0, 0.0), zb = make_double2(0.0, 0.0);
    for (int nu = q; nu <= Ls; ++nu) {
        int kk = nu < K ? nu : K;
        int64_t iy = rows_off(nu, K) + k + kk;
        za = cadd(za, cmul(Ab[nu - q], ya[iy]));
        zb = cadd(zb, cmul(Bb[nu - q], yb[iy]));
    }
    double2 *z = wz + T.off_z;
    z[e] = za;
    z[e + nent] = zb;
}

void launch_axial_apply(const Task *tasks, int ntasks, int64_t max_rows, const double2 *wax, const double2 *wy,
                        double2 *wz, hipStream_t s) {
    if (ntasks <= 0) return;
    dim3 grid((unsigned)((max_rows + 255) / 256), ntasks);
    hipLaunchKernelGGL(k_axial_apply, grid, dim3(256), 0, s, tasks, wax, wy, wz);
}

// ------------------------------------------------------------------------
// Back rotation with the folded tables and accumulation into the target, one
// workgroup per target segment (fixed task order: deterministic, no atomics):
//   F_m = sum_{k>=0} E[k][m] s_k,  G_m = sum_{k>=0} O[k][m] t_k,  m = 0..n
//   s_k = z_k + (-1)^k z_{-k} (s_0 = z_0), t_k = z_k - (-1)^k z_{-k}
//   w_m = (F_m + G_m)/2,  w_{-m} = (-1)^m (F_m - G_m)/2
//   out_a/b(n, +-m) += e^{-+ i m phi} (w_+ +- w_-)/2
// One thread per (n, m >= 0).
// ------------------------------------------------------------------------
__global__ void k_rot_back(const Task *__restrict__ tasks, const Segment *__restrict__ segs,
                           const double *__restrict__ wd, const double2 *__restrict__ wz,
                           double2 *__restrict__ dst_pool) {
    const Segment S = segs[blockIdx.x];
    const int Lt = S.Lt;
    const int nent = Lt * (Lt + 2);
    const int nf = Lt * (Lt + 3) / 2;  // entries (n, m >= 0), n = 1..Lt
    for (int e = blockIdx.y * blockDim.x + threadIdx.x; e < nf; e += blockDim.x * gridDim.y) {
        int n = (int)((sqrt(8.0 * (double)e + 9.0) - 3.0) * 0.5);
        while ((n + 1) * (n + 4) / 2 <= e) ++n;
        while (n > 0 && n * (n + 3) / 2 > e) --n;
        const int m = e - n * (n + 3) / 2;
        n += 1;
        double2 ap = make_double2(0.0, 0.0), bp = ap, an = ap, bn = ap;  // target increments at +m, -m
        for (int ti = S.t_begin; ti < S.t_end; ++ti) {
            const Task T = tasks[ti];
            const int K = T.K;
            const int kk = n < K ? n : K;
            const double *E = wd + T.off_d + fold_off(n, K) + m;  // E[k][m] at k * (n+1)
            const double *O = E + (int64_t)(kk + 1) * (n + 1);
            const int64_t nz = rows_size(Lt, K);
            const double2 *zp = wz + T.off_z + rows_off(n, K) + kk;  // z_+(n, k) at zp[k]
            const double2 *zm = zp + nz;                              // z_-(n, k)
            double2 Fp = make_double2(0.0, 0.0), Fm = Fp, Gp = Fp, Gm = Fp;
            for (int k = 0; k <= kk; ++k) {
                const double ev = E[(int64_t)k * (n + 1)], ov = O[(int64_t)k * (n + 1)];
                if (k == 0) {
                    Fp = cadd(Fp, cscale(ev, zp[0]));
                    Fm = cadd(Fm, cscale(ev, zm[0]));
                    continue;
                }
                const double sg = (k & 1) ? -1.0 : 1.0;
                Fp = cadd(Fp, cscale(ev, cadd(zp[k], cscale(sg, zp[-k]))));
                Fm = cadd(Fm, cscale(ev, cadd(zm[k], cscale(sg, zm[-k]))));
                Gp = cadd(Gp, cscale(ov, csub(zp[k], cscale(sg, zp[-k]))));
                Gm = cadd(Gm, cscale(ov, csub(zm[k], cscale(sg, zm[-k]))));
            }
            double sn, cs;
            sincos((double)m * T.phi, &sn, &cs);
            const double2 pc = make_double2(cs, -sn), p = make_double2(cs, sn);  // e^{-i m phi}, e^{i m phi}
            // w at +m: (F + G)/2 per part; a = (w_+ + w_-)/2, b = (w_+ - w_-)/2
            double2 wpp = cscale(0.5, cadd(Fp, Gp)), wpm = cscale(0.5, cadd(Fm, Gm));
            ap = cadd(ap, cmul(pc, cscale(0.5, cadd(wpp, wpm))));
            bp = cadd(bp, cmul(pc, cscale(0.5, csub(wpp, wpm))));
            if (m > 0) {
                const double f = (m & 1) ? -0.5 : 0.5;
                double2 wnp = cscale(f, csub(Fp, Gp)), wnm = cscale(f, csub(Fm, Gm));
                an = cadd(an, cmul(p, cscale(0.5, cadd(wnp, wnm))));
                bn = cadd(bn, cmul(p, cscale(0.5, csub(wnp, wnm))));
            }
        }
        double2 *dst = dst_pool + S.dst;
        const int o = n * n - 1 + m + n;
        dst[o] = cadd(dst[o], ap);
        dst[o + nent] = cadd(dst[o + nent], bp);
        if (m > 0) {
            const int on = n * n - 1 - m + n;
            dst[on] = cadd(dst[on], an);
            dst[on + nent] = cadd(dst[on + nent], bn);
        }
    }
}

void launch_rot_back(const Task *tasks, const Segment *segs, int nsegs, int yblocks, const double *wd,
                     const double2 *wz, double2 *dst_pool, hipStream_t s) {
    if (nsegs <= 0) return;
    dim3 grid(nsegs, yblocks < 1 ? 1 : yblocks);
    hipLaunchKernelGGL(k_rot_back, grid, dim3(256), 0, s, tasks, segs, wd, wz, dst_pool);
}

}  // namespace ifs

namespace ifs {

// ------------------------------------------------------------------------
// Mie coefficients a_n, b_n (Barber & Hill log-derivative form), one block per
// sphere, sequential recurrences on thread 0 (FP64).  Restates
// FoggySim/src/lib/mie_theory/lib_mie_ss_helper_functions.f90:610-722 (a_n, b_n),
// :819-865 and :900-924 (A_n downward), :937-951 (n_c).
// ------------------------------------------------------------------------
__global__ void k_mie(const double *__restrict__ xsize, const double2 *__restrict__ mrel, int N,
                      double2 *__restrict__ an, double2 *__restrict__ bn) {
    extern __shared__ double2 shm[];
    if (threadIdx.x != 0) return;
    const int j = blockIdx.x;
    const double x = xsize[j];
    const double2 m = mrel[j];
    double2 *An = shm;            // [N+2]
    double2 *f = shm + (N + 2);   // [N+1]: j_n in .x, y_n in .y
    const double2 mx = make_double2(m.x * x, m.y * x);
    const double amx = sqrt(mx.x * mx.x + mx.y * mx.y);
    const int nc = (int)ceil(x + 4.05 * cbrt(x) + 2.0);
    int nstart = (nc > (int)amx ? nc : (int)amx) + 15;
    if (nstart < N + 1) nstart = N + 1;  // every A_n, n <= N, comes from the recurrence
    double2 buf = make_double2(0.0, 0.0);
    for (int i = nstart; i >= 2; --i) {
        double2 ndm = cdiv(make_double2((double)i, 0.0), mx);
        buf = csub(ndm, cdiv(make_double2(1.0, 0.0), cadd(buf, ndm)));
        if (i - 1 <= N) An[i - 1] = buf;
    }
    radial_into(x, N, 3, f);
    for (int n = 1; n <= N; ++n) {
        double2 jn = make_double2(f[n].x, 0.0), jn1 = make_double2(f[n - 1].x, 0.0);
        double2 hn = f[n], hn1 = f[n - 1];
        double2 ndx = make_double2((double)n / x, 0.0);
        double2 t = cadd(cdiv(An[n], m), ndx);
        an[(int64_t)j * N + n - 1] = cdiv(csub(cmul(t, jn), jn1), csub(cmul(t, hn), hn1));
        t = cadd(cmul(m, An[n]), ndx);
        bn[(int64_t)j * N + n - 1] = cdiv(csub(cmul(t, jn), jn1), csub(cmul(t, hn), hn1));
    }
}

void launch_mie(int ns, const double *xsize, const double2 *mrel, int N, double2 *an, double2 *bn, hipStream_t s) {
    if (ns <= 0) return;
    size_t shm = sizeof(double2) * (size_t)(2 * N + 4);
    hipLaunchKernelGGL(k_mie, dim3(ns), dim3(64), shm, s, xsize, mrel, N, an, bn);
}

// ------------------------------------------------------------------------
// Layout conversion.  Reference flat layout (Xu basis, global N) <-> compact
// normalised per-sphere blocks (order Leff_j).
// ------------------------------------------------------------------------
__global__ void k_pack_in(int N, const int32_t *__restrict__ Leff, const int64_t *__restrict__ off,
                          const double *__restrict__ snm, const double2 *__restrict__ x, double2 *__restrict__ xh,
                          int64_t ldx, int64_t ldxh) {
    const int j = blockIdx.x;
    const int r = blockIdx.z;
    const int L = Leff[j];
    const int blk = N * (N + 2);
    const int cnt = L * (L + 2);
    const double2 *xs = x + r * ldx + (int64_t)j * 2 * blk;
    double2 *xd = xh + r * ldxh + off[j];
    for (int e = blockIdx.y * blockDim.x + threadIdx.x; e < cnt; e += blockDim.x * gridDim.y) {
        double s = snm ? snm[e] : 1.0;
        xd[e] = cscale(s, xs[e]);
        xd[e + cnt] = cscale(s, xs[e + blk]);
    }
}

void launch_pack_in(int ns, int N, const int32_t *Leff, const int64_t *off, const double *snm, const double2 *x,
                    double2 *xh, int nrhs, int64_t ldx, int64_t ldxh, hipStream_t s) {
    if (ns <= 0) return;
    int yb = (N * (N + 2) + 255) / 256;
    hipLaunchKernelGGL(k_pack_in, dim3(ns, yb, nrhs), dim3(256), 0, s, N, Leff, off, snm, x, xh, ldx, ldxh);
}

// b = x + [active_j, n <= Leff_j] T_n(j) acc(n,m) / s(n,m)
__global__ void k_finalize(int N, const int32_t *__restrict__ Leff, const int32_t *__restrict__ active,
                           const int64_t *__restrict__ off, const double *__restrict__ snm,
                           const double2 *__restrict__ an, const double2 *__restrict__ bn,
                           const double2 *__restrict__ x, const double2 *__restrict__ acc, double2 *__restrict__ b,
                           int64_t ldx, int64_t ldb, int64_t ldacc) {
    const int j = blockIdx.x;
    const int r = blockIdx.z;
    const int L = Leff[j];
    const int blk = N * (N + 2);
    const int cnt = L * (L + 2);
    const double2 *xs = x + r * ldx + (int64_t)j * 2 * blk;
    double2 *bs = b + r * ldb + (int64_t)j * 2 * blk;
    const double2 *as = acc + r * ldacc + off[j];
    const int act = active[j];
    for (int e = blockIdx.y * blockDim.x + threadIdx.x; e < blk; e += blockDim.x * gridDim.y) {
        double2 va = xs[e], vb = xs[e + blk];
        if (act && e < cnt) {
            int n = (int)sqrt((double)(e + 1));
            while ((n + 1) * (n + 1) - 1 <= e) ++n;
            while (n * n - 1 > e) --n;
            double inv = snm ? 1.0 / snm[e] : 1.0;
            va = cadd(va, cmul(an[(int64_t)j * N + n - 1], cscale(inv, as[e])));
            vb = cadd(vb, cmul(bn[(int64_t)j * N + n - 1], cscale(inv, as[e + cnt])));
        }
        bs[e] = va;
        bs[e + blk] = vb;
    }
}

void launch_finalize(int ns, int N, const int32_t *Leff, const int32_t *active, const int64_t *off,
                     const double *snm, const double2 *an, const double2 *bn, const double2 *x, const double2 *acc,
                     double2 *b, int nrhs, int64_t ldx, int64_t ldb, int64_t ldacc, hipStream_t s) {
    if (ns <= 0) return;
    int yb = (N * (N + 2) + 255) / 256;
    hipLaunchKernelGGL(k_finalize, dim3(ns, yb, nrhs), dim3(256), 0, s, N, Leff, active, off, snm, an, bn, x, acc,
                       b, ldx, ldb, ldacc);
}

// y = s * x (or x / s) per flat block entry, all n <= N (reference layout)
__global__ void k_scale_flat(int N, const double *__restrict__ snm, const double2 *__restrict__ x,
                             double2 *__restrict__ y, int inverse) {
    const int j = blockIdx.x;
    const int blk = N * (N + 2);
    const double2 *xs = x + (int64_t)j * 2 * blk;
    double2 *ys = y + (int64_t)j * 2 * blk;
    for (int e = blockIdx.y * blockDim.x + threadIdx.x; e < blk; e += blockDim.x * gridDim.y) {
        double s = inverse ? 1.0 / snm[e] : snm[e];
        ys[e] = cscale(s, xs[e]);
        ys[e + blk] = cscale(s, xs[e + blk]);
    }
}

void launch_scale_flat(int ns, int N, const double *snm, const double2 *x, double2 *y, int inverse, hipStream_t s) {
    if (ns <= 0) return;
    int yb = (N * (N + 2) + 255) / 256;
    hipLaunchKernelGGL(k_scale_flat, dim3(ns, yb), dim3(256), 0, s, N, snm, x, y, inverse);
}

__global__ void k_zero(double2 *__restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)blockDim.x * gridDim.x)
        p[i] = make_double2(0.0, 0.0);
}

void launch_zero(double2 *p, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    int64_t nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)nb), dim3(256), 0, s, p, n);
}

// ------------------------------------------------------------------------
// Krylov BLAS-1 kernels (FP64 complex), fixed reduction order.
// ------------------------------------------------------------------------
// partial[b * nv + i] = sum over block b's slice of conj(V_i) w ; then out[i] = sum_b partial
__global__ void k_dots_partial(const double2 *__restrict__ V, int64_t ldv, int nv, const double2 *__restrict__ w,
                               int64_t n, double2 *__restrict__ partial) {
    __shared__ double2 red[256];
    for (int i = 0; i < nv; ++i) {
        const double2 *v = V + (int64_t)i * ldv;
        double2 acc = make_double2(0.0, 0.0);
        for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)blockDim.x * gridDim.x) {
            double2 a = v[e], b = w[e];
            acc.x += a.x * b.x + a.y * b.y;
            acc.y += a.x * b.y - a.y * b.x;
        }
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] = cadd(red[threadIdx.x], red[threadIdx.x + s]);
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * nv + i] = red[0];
        __syncthreads();
    }
}

__global__ void k_dots_final(const double2 *__restrict__ partial, int nblocks, int nv, double2 *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    double2 s = make_double2(0.0, 0.0);
    for (int b = 0; b < nblocks; ++b) s = cadd(s, partial[(int64_t)b * nv + i]);
    out[i] = s;
}

void launch_dots(const double2 *V, int64_t ldv, int nv, const double2 *w, int64_t n, double2 *out, double2 *partial,
                 int nblocks, hipStream_t s) {
    if (nv <= 0) return;
    hipLaunchKernelGGL(k_dots_partial, dim3(nblocks), dim3(256), 0, s, V, ldv, nv, w, n, partial);
    hipLaunchKernelGGL(k_dots_final, dim3((nv + 63) / 64), dim3(64), 0, s, partial, nblocks, nv, out);
}

// w -= sum_i h_i V_i
__global__ void k_gemv_sub(const double2 *__restrict__ V, int64_t ldv, int nv, const double2 *__restrict__ h,
                           double2 *__restrict__ w, int64_t n) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)blockDim.x * gridDim.x) {
        double2 acc = w[e];
        for (int i = 0; i < nv; ++i) acc = csub(acc, cmul(h[i], V[(int64_t)i * ldv + e]));
        w[e] = acc;
    }
}

void launch_gemv_sub(const double2 *V, int64_t ldv, int nv, const double2 *h, double2 *w, int64_t n, hipStream_t s) {
    if (nv <= 0) return;
    int64_t nb = (n + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(k_gemv_sub, dim3((unsigned)nb), dim3(256), 0, s, V, ldv, nv, h, w, n);
}

// x += sum_i c_i V_i
__global__ void k_gemv_add(const double2 *__restrict__ V, int64_t ldv, int nv, const double2 *__restrict__ c,
                           double2 *__restrict__ x, int64_t n) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)blockDim.x * gridDim.x) {
        double2 acc = x[e];
        for (int i = 0; i < nv; ++i) acc = cadd(acc, cmul(c[i], V[(int64_t)i * ldv + e]));
        x[e] = acc;
    }
}

void launch_gemv_add(const double2 *V, int64_t ldv, int nv, const double2 *c, double2 *x, int64_t n, hipStream_t s) {
    if (nv <= 0) return;
    int64_t nb = (n + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(k_gemv_add, dim3((unsigned)nb), dim3(256), 0, s, V, ldv, nv, c, x, n);
}

// y = a x + b y
__global__ void k_axpby(double2 a, const double2 *__restrict__ x, double2 b, double2 *__restrict__ y, int64_t n) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)blockDim.x * gridDim.x)
        y[e] = cadd(cmul(a, x[e]), cmul(b, y[e]));
}

void launch_axpby(double2 a, const double2 *x, double2 b, double2 *y, int64_t n, hipStream_t s) {
    int64_t nb = (n + 255) / 256;
    if (nb > 8192) nb = 8192;
    if (nb < 1) return;
    hipLaunchKernelGGL(k_axpby, dim3((unsigned)nb), dim3(256), 0, s, a, x, b, y, n);
}

// y = a x
__global__ void k_scale_copy(double2 a, const double2 *__restrict__ x, double2 *__restrict__ y, int64_t n) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)blockDim.x * gridDim.x)
        y[e] = cmul(a, x[e]);
}

void launch_scale_copy(double2 a, const double2 *x, double2 *y, int64_t n, hipStream_t s) {
    int64_t nb = (n + 255) / 256;
    if (nb > 8192) nb = 8192;
    if (nb < 1) return;
    hipLaunchKernelGGL(k_scale_copy, dim3((unsigned)nb), dim3(256), 0, s, a, x, y, n);
}

}  // namespace ifs
