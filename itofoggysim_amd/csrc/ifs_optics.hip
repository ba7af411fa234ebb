// Optics of a solved cluster (gfx950): plane-wave expansion coefficients, far
// field, amplitude and Mueller matrices, cross sections.
//
// Conventions (Xu's, as the reference's Mie code uses them):
//   incident wave along (alpha, beta), amplitude 1, at sphere centre d_j
//     p_mn = e^{i k khat.d_j} tau_mn(alpha) e^{-i m beta} / (n(n+1)),  q_mn likewise with pi_mn
//     (FoggySim/src/lib/mie_theory/lib_mie_illumination.f90:636-763);
//   far field E_s ~ e^{ikr}/(kr) F, with, in the normalised basis c_hat = s_nm c,
//     F_theta =  i sum_j e^{-ik rhat.d_j} sum_nm e^{i m phi} (a_hat T_nm + b_hat P_nm)
//     F_phi   = -  sum_j e^{-ik rhat.d_j} sum_nm e^{i m phi} (a_hat P_nm + b_hat T_nm)
//   where T_nm = s_nm tau_mn / (n(n+1)) and P_nm = s_nm pi_mn / (n(n+1)) are, through
//   the Wigner functions d^n_{m,+-1}(theta),
//     T_nm = (-1)^{m+1} sqrt(2n+1) (d^n_{m,1} - d^n_{m,-1}) / 2
//     P_nm = (-1)^{m+1} sqrt(2n+1) (d^n_{m,1} + d^n_{m,-1}) / 2   (bounded for every n);
//   cross sections (lib_mie_ss_helper_functions.f90:1380-1460, Xu eq. 57, summed over
//   the spheres; absorption from each sphere's exciting field a^j_mn / a_n^j):
//     C_ext = 4 pi/k^2 sum Re(conj(p_hat) a_hat + conj(q_hat) b_hat)
//     C_abs = 4 pi/k^2 sum |a_hat|^2 (Re(1/a_n) - 1) + |b_hat|^2 (Re(1/b_n) - 1),
//     C_sca = C_ext - C_abs.
#include <hip/hip_runtime.h>
#include <math.h>

#include <cmath>
#include <complex>
#include <map>
#include <mutex>
#include <vector>

#include "ifs_engine.h"
#include "ifs_wigner.h"

namespace ifs {

namespace {

__device__ __forceinline__ double2 cmul_(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// T_nm, P_nm at the polar angles theta[i]: tab[i * N(N+2) + fidx(n,m)] = (T, P).
// One block per direction, threads over m = -N..N, each running the recurrences
// of d^n_{m,1} and d^n_{m,-1} upward in n.
__global__ void k_ff_tables(int N, const double *__restrict__ theta, double2 *__restrict__ tab) {
    const int i = blockIdx.x;
    const double th = theta[i];
    const double cb = cos(th);
    const double ch = cos(0.5 * th), sh = sin(0.5 * th);
    const int64_t blk = (int64_t)N * (N + 2);
    double2 *out = tab + (int64_t)i * blk;
    for (int m = -N + (int)threadIdx.x; m <= N; m += blockDim.x) {
        const int am = m < 0 ? -m : m;
        const int n0 = am > 1 ? am : 1;
        const double sg = (am & 1) ? 1.0 : -1.0;  // (-1)^{m+1}
        double p1 = wigner_seed(m, 1, ch, sh), q1 = wigner_seed(m, -1, ch, sh);
        double p2 = 0.0, q2 = 0.0;
        for (int n = n0; n <= N; ++n) {
            if (n > n0) {
                double pn = wigner_step(n, m, 1, cb, p1, p2);
                double qn = wigner_step(n, m, -1, cb, q1, q2);
                p2 = p1;
                p1 = pn;
                q2 = q1;
                q1 = qn;
            }
            double f = 0.5 * sg * sqrt(2.0 * n + 1.0);
            out[n * n - 1 + m + n] = make_double2(f * (p1 - q1), f * (p1 + q1));
        }
    }
}

constexpr int FFD = 4;  // directions per block

// F_theta, F_phi for FFD directions per block; threads run over the entries of
// every sphere, then reduce through LDS.  out[i] = (F_theta, F_phi) complex.
__global__ __launch_bounds__(256) void k_far_field(int N, int ns, const int32_t *__restrict__ Leff,
                                                   const double *__restrict__ pos, double k,
                                                   const double *__restrict__ snm, const double2 *__restrict__ x,
                                                   int ndir, const double *__restrict__ theta,
                                                   const double *__restrict__ phi, const double2 *__restrict__ tab,
                                                   double2 *__restrict__ out) {
    extern __shared__ double2 sh[];
    const int W = 2 * N + 1;
    double2 *em = sh;                    // [FFD][W]: e^{i m phi_i}
    double2 *red = sh + FFD * W;         // [256] reduction scratch
    __shared__ double rh[FFD][3];
    const int i0 = blockIdx.x * FFD;
    const int nd = (ndir - i0) < FFD ? (ndir - i0) : FFD;
    for (int q = threadIdx.x; q < FFD * W; q += blockDim.x) {
        int d = q / W, m = q % W - N;
        double s = 0.0, c = 1.0;
        if (d < nd) sincos((double)m * phi[i0 + d], &s, &c);
        em[q] = make_double2(c, s);
    }
    if (threadIdx.x < FFD) {
        int d = threadIdx.x;
        double t = d < nd ? theta[i0 + d] : 0.0, f = d < nd ? phi[i0 + d] : 0.0;
        rh[d][0] = sin(t) * cos(f);
        rh[d][1] = sin(t) * sin(f);
        rh[d][2] = cos(t);
    }
    __syncthreads();
    const int64_t blk = (int64_t)N * (N + 2);
    double2 at[FFD], ap[FFD];
#pragma unroll
    for (int d = 0; d < FFD; ++d) at[d] = ap[d] = make_double2(0.0, 0.0);
    for (int j = 0; j < ns; ++j) {
        const int L = Leff[j];
        const int cnt = L * (L + 2);
        double2 ph[FFD];
#pragma unroll
        for (int d = 0; d < FFD; ++d) {
            double arg = -k * (rh[d][0] * pos[3 * j] + rh[d][1] * pos[3 * j + 1] + rh[d][2] * pos[3 * j + 2]);
            double s, c;
            sincos(arg, &s, &c);
            ph[d] = make_double2(c, s);
        }
        const double2 *xa = x + (int64_t)j * 2 * blk;
        for (int e = threadIdx.x; e < cnt; e += blockDim.x) {
            int n = (int)sqrt((double)(e + 1));
            while ((n + 1) * (n + 1) - 1 <= e) ++n;
            while (n * n - 1 > e) --n;
            const int m = e - (n * n - 1) - n;
            const double s = snm[e];
            const double2 a = make_double2(s * xa[e].x, s * xa[e].y);
            const double2 b = make_double2(s * xa[e + blk].x, s * xa[e + blk].y);
#pragma unroll
            for (int d = 0; d < FFD; ++d) {
                if (d >= nd) break;
                const double2 tp = tab[(int64_t)(i0 + d) * blk + e];  // (T, P)
                const double2 w = cmul_(ph[d], em[d * W + m + N]);
                double2 vt = make_double2(a.x * tp.x + b.x * tp.y, a.y * tp.x + b.y * tp.y);  // a T + b P
                double2 vp = make_double2(a.x * tp.y + b.x * tp.x, a.y * tp.y + b.y * tp.x);  // a P + b T
                vt = cmul_(w, vt);
                vp = cmul_(w, vp);
                at[d].x += vt.x;
                at[d].y += vt.y;
                ap[d].x += vp.x;
                ap[d].y += vp.y;
            }
        }
    }
    // block reduction, one direction and component at a time (fixed order)
    for (int d = 0; d < nd; ++d) {
        for (int comp = 0; comp < 2; ++comp) {
            red[threadIdx.x] = comp == 0 ? at[d] : ap[d];
            __syncthreads();
            for (int h = blockDim.x / 2; h > 0; h >>= 1) {
                if ((int)threadIdx.x < h) {
                    red[threadIdx.x].x += red[threadIdx.x + h].x;
                    red[threadIdx.x].y += red[threadIdx.x + h].y;
                }
                __syncthreads();
            }
            if (threadIdx.x == 0) {
                double2 v = red[0];
                // F_theta = i * sum, F_phi = - sum
                out[2 * (int64_t)(i0 + d) + comp] = comp == 0 ? make_double2(-v.y, v.x) : make_double2(-v.x, -v.y);
            }
            __syncthreads();
        }
    }
}

// Per-sphere partial sums for the cross sections: part[j] = (extinction, absorption)
// sums (without the 4 pi / k^2 factor).
__global__ __launch_bounds__(256) void k_xsec(int N, const int32_t *__restrict__ Leff, const double *__restrict__ snm,
                                              const double2 *__restrict__ x, const double2 *__restrict__ p,
                                              const double2 *__restrict__ an, const double2 *__restrict__ bn,
                                              double2 *__restrict__ part) {
    __shared__ double2 red[256];
    const int j = blockIdx.x;
    const int L = Leff[j];
    const int cnt = L * (L + 2);
    const int64_t blk = (int64_t)N * (N + 2);
    const double2 *xa = x + (int64_t)j * 2 * blk;
    const double2 *pa = p + (int64_t)j * 2 * blk;
    double ext = 0.0, abs_ = 0.0;
    for (int e = threadIdx.x; e < cnt; e += blockDim.x) {
        int n = (int)sqrt((double)(e + 1));
        while ((n + 1) * (n + 1) - 1 <= e) ++n;
        while (n * n - 1 > e) --n;
        const double s2 = snm[e] * snm[e];
        const double2 a = xa[e], b = xa[e + blk], pp = pa[e], qq = pa[e + blk];
        ext += s2 * (pp.x * a.x + pp.y * a.y + qq.x * b.x + qq.y * b.y);
        const double2 A = an[(int64_t)j * N + n - 1], B = bn[(int64_t)j * N + n - 1];
        const double na = A.x * A.x + A.y * A.y, nb = B.x * B.x + B.y * B.y;
        // Re(1/a_n) - 1 = Re(a_n)/|a_n|^2 - 1
        if (na > 0.0) abs_ += s2 * (a.x * a.x + a.y * a.y) * (A.x / na - 1.0);
        if (nb > 0.0) abs_ += s2 * (b.x * b.x + b.y * b.y) * (B.x / nb - 1.0);
    }
    red[threadIdx.x] = make_double2(ext, abs_);
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            red[threadIdx.x].x += red[threadIdx.x + h].x;
            red[threadIdx.x].y += red[threadIdx.x + h].y;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[j] = red[0];
}

// Host: d^n_{m,+-1}(beta) for n = 1..N, |m| <= n, in long double (seed in logs,
// three-term recurrence in n); dp/dm indexed at n*n-1+m+n.
void d_pm1_host(int N, long double beta, std::vector<long double> &dp, std::vector<long double> &dm) {
    const int64_t blk = (int64_t)N * (N + 2);
    dp.assign(blk, 0.0L);
    dm.assign(blk, 0.0L);
    const long double cb = cosl(beta), c = cosl(0.5L * beta), s = sinl(0.5L * beta);
    auto seed = [&](int k, int m) -> long double {
        long double sign = 1.0L;
        if (std::abs(k) < std::abs(m)) {
            std::swap(k, m);
            if ((k - m) & 1) sign = -sign;
        }
        int j = std::abs(k), ec, es;
        if (k >= 0) {
            ec = j + m;
            es = j - m;
            if (es & 1) sign = -sign;
        } else {
            ec = j - m;
            es = j + m;
        }
        long double lv = 0.5L * (lgammal(2.0L * j + 1.0L) - lgammal((long double)(j + m) + 1.0L) -
                                 lgammal((long double)(j - m) + 1.0L));
        if (ec > 0) {
            if (c <= 0.0L) return 0.0L;
            lv += ec * logl(c);
        }
        if (es > 0) {
            if (s <= 0.0L) return 0.0L;
            lv += es * logl(s);
        }
        return sign * expl(lv);
    };
    auto step = [&](int n, int k, int m, long double d1, long double d2) {
        long double nn = n, k2 = (long double)k * k, m2 = (long double)m * m;
        long double num = (2.0L * nn - 1.0L) * (nn * (nn - 1.0L) * cb - (long double)k * m) * d1 -
                          nn * sqrtl(((nn - 1) * (nn - 1) - k2) * ((nn - 1) * (nn - 1) - m2)) * d2;
        long double den = (nn - 1.0L) * sqrtl((nn * nn - k2) * (nn * nn - m2));
        return num / den;
    };
    for (int m = -N; m <= N; ++m) {
        const int n0 = std::max(std::abs(m), 1);
        long double p1 = seed(m, 1), q1 = seed(m, -1), p2 = 0.0L, q2 = 0.0L;
        for (int n = n0; n <= N; ++n) {
            if (n > n0) {
                long double pn = step(n, m, 1, p1, p2), qn = step(n, m, -1, q1, q2);
                p2 = p1;
                p1 = pn;
                q2 = q1;
                q1 = qn;
            }
            dp[n * n - 1 + m + n] = p1;
            dm[n * n - 1 + m + n] = q1;
        }
    }
}

}  // namespace

}  // namespace ifs

// Scratch of ifs_far_field_device_noalloc: the per-direction T/P tables, in a
// fixed array of this library on every GPU it runs on (held until the process
// ends).  Host code never uses the array's name as an address: its device address
// is looked up once per device (far_field_scratch) and passed to the kernels.
static __device__ __align__(256) double2 scratch[IFS_FF_SCRATCH_ENTRIES];

namespace ifs {
int far_field_scratch(int device, void **out) {
    static std::mutex mu;
    static std::map<int, void *> per_device;
    std::lock_guard<std::mutex> lock(mu);
    auto it = per_device.find(device);
    if (it != per_device.end()) {
        *out = it->second;
        return IFS_OK;
    }
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return IFS_E_HIP;
    per_device[device] = p;
    *out = p;
    return IFS_OK;
}
}  // namespace ifs

using cplx = std::complex<double>;

extern "C" int ifs_plane_wave(ifs_ctx *c, double alpha, double beta, double *p_out, int64_t n) {
    if (!c || !p_out || n != ifs_vector_length(c)) return IFS_E_ARG;
    const int N = c->N, ns = c->ns;
    const int64_t blk = (int64_t)N * (N + 2);
    std::vector<long double> dp, dm;
    ifs::d_pm1_host(N, (long double)alpha, dp, dm);
    // angular part tau_mn / (n(n+1)) and pi_mn / (n(n+1)) (flat layout, unnormalised)
    std::vector<cplx> ta(blk), pa(blk);
    for (int nn = 1; nn <= N; ++nn)
        for (int m = -nn; m <= nn; ++m) {
            int64_t e = (int64_t)nn * nn - 1 + m + nn;
            // s_nm = sqrt(n(n+1)(2n+1)(n-m)!/(n+m)!)
            long double ls = 0.5L * (logl((long double)nn * (nn + 1) * (2 * nn + 1)) + lgammal(nn - m + 1.0L) -
                                     lgammal(nn + m + 1.0L));
            long double sg = (std::abs(m) & 1) ? 1.0L : -1.0L;  // (-1)^{m+1}
            long double f = 0.5L * sg * sqrtl(2.0L * nn + 1.0L) / expl(ls);
            long double t = f * (dp[e] - dm[e]), p = f * (dp[e] + dm[e]);
            long double a = -(long double)m * beta;
            cplx em((double)cosl(a), (double)sinl(a));
            ta[e] = (double)t * em;
            pa[e] = (double)p * em;
        }
    const double kh[3] = {std::sin(alpha) * std::cos(beta), std::sin(alpha) * std::sin(beta), std::cos(alpha)};
    cplx *out = (cplx *)p_out;
    for (int j = 0; j < ns; ++j) {
        double kd = c->prm.k * (kh[0] * c->pos[3 * j] + kh[1] * c->pos[3 * j + 1] + kh[2] * c->pos[3 * j + 2]);
        cplx ph(std::cos(kd), std::sin(kd));
        for (int64_t e = 0; e < blk; ++e) {
            out[(int64_t)j * 2 * blk + e] = ph * ta[e];
            out[(int64_t)j * 2 * blk + blk + e] = ph * pa[e];
        }
    }
    return IFS_OK;
}

// TE mode of the reference's illumination: the TM coefficients with p and q exchanged
// (lib_mie_illumination.f90:347-352; the polarisation turned by 90 degrees about the
// propagation direction).
extern "C" int ifs_plane_wave_mode(ifs_ctx *c, double alpha, double beta, int32_t te, double *p_out, int64_t n) {
    if (te != 0 && te != 1) return IFS_E_ARG;
    int rc = ifs_plane_wave(c, alpha, beta, p_out, n);
    if (rc || !te) return rc;
    const int64_t blk = (int64_t)c->N * (c->N + 2);
    cplx *out = (cplx *)p_out;
    for (int j = 0; j < c->ns; ++j)
        for (int64_t e = 0; e < blk; ++e) std::swap(out[(int64_t)j * 2 * blk + e], out[(int64_t)j * 2 * blk + blk + e]);
    return IFS_OK;
}

extern "C" int ifs_plane_wave_rhs_dir(ifs_ctx *c, double alpha, double beta, double *rhs, int64_t n) {
    return ifs_plane_wave_rhs_mode(c, alpha, beta, 0, rhs, n);
}

extern "C" int ifs_plane_wave_rhs_mode(ifs_ctx *c, double alpha, double beta, int32_t te, double *rhs, int64_t n) {
    int rc = ifs_plane_wave_mode(c, alpha, beta, te, rhs, n);
    if (rc) return rc;
    const int N = c->N, ns = c->ns;
    const int64_t blk = (int64_t)N * (N + 2);
    std::vector<cplx> an((size_t)ns * N), bn((size_t)ns * N);
    if ((rc = ifs_get_mie(c, (double *)an.data(), (double *)bn.data()))) return rc;
    cplx *out = (cplx *)rhs;
    for (int j = 0; j < ns; ++j)
        for (int nn = 1; nn <= N; ++nn)
            for (int m = -nn; m <= nn; ++m) {
                int64_t e = (int64_t)nn * nn - 1 + m + nn;
                bool keep = nn <= c->Leff[j] && c->active[j];
                out[(int64_t)j * 2 * blk + e] *= keep ? an[(size_t)j * N + nn - 1] : 0.0;
                out[(int64_t)j * 2 * blk + blk + e] *= keep ? bn[(size_t)j * N + nn - 1] : 0.0;
            }
    return IFS_OK;
}

extern "C" int ifs_cross_sections_device(ifs_ctx *c, const void *x_dev, const void *p_dev, int64_t n,
                                         double out[3]) {
    if (!c || !x_dev || !p_dev || !out || n != ifs_vector_length(c)) return IFS_E_ARG;
    ifs::DevBuf part;
    int rc;
    if ((rc = part.alloc(16ull * c->ns))) return rc;
    hipLaunchKernelGGL(ifs::k_xsec, dim3(c->ns), dim3(256), 0, c->stream, c->N, c->d_Leff.as<int32_t>(),
                       c->d_snm.as<double>(), (const double2 *)x_dev, (const double2 *)p_dev,
                       c->d_an.as<double2>(), c->d_bn.as<double2>(), part.as<double2>());
    if (hipGetLastError() != hipSuccess) return IFS_E_HIP;
    std::vector<double2> h(c->ns);
    if (hipMemcpyAsync(h.data(), part.p, 16ull * c->ns, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return IFS_E_HIP;
    double ext = 0.0, ab = 0.0;
    for (int j = 0; j < c->ns; ++j) {
        ext += h[j].x;
        ab += h[j].y;
    }
    const double f = 4.0 * M_PI / (c->prm.k * c->prm.k);
    out[0] = f * (ext - ab);
    out[1] = f * ext;
    out[2] = f * ab;
    return IFS_OK;
}

extern "C" int ifs_cross_sections(ifs_ctx *c, const double *x, const double *p, int64_t n, double out[3]) {
    if (!c || !x || !p || !out || n != ifs_vector_length(c)) return IFS_E_ARG;
    ifs::DevBuf dx, dp;
    int rc;
    if ((rc = dx.alloc(16ull * n)) || (rc = dp.alloc(16ull * n))) return rc;
    if (hipMemcpy(dx.p, x, 16ull * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dp.p, p, 16ull * n, hipMemcpyHostToDevice) != hipSuccess)
        return IFS_E_HIP;
    return ifs_cross_sections_device(c, dx.p, dp.p, n, out);
}

extern "C" int ifs_far_field_device(ifs_ctx *c, const void *x_dev, int64_t n, int32_t ndir, const void *theta_dev,
                                    const void *phi_dev, void *out_dev) {
    if (!c || !x_dev || !theta_dev || !phi_dev || !out_dev || n != ifs_vector_length(c) || ndir < 0)
        return IFS_E_ARG;
    if (ndir == 0) return IFS_OK;
    const int N = c->N;
    int rc;
    ifs::DevBuf tab, dpos;
    if ((rc = tab.alloc(16ull * ndir * N * (N + 2))) || (rc = dpos.alloc(24ull * c->ns))) return rc;
    if (hipMemcpyAsync(dpos.p, c->pos.data(), 24ull * c->ns, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return IFS_E_HIP;
    hipLaunchKernelGGL(ifs::k_ff_tables, dim3(ndir), dim3(256), 0, c->stream, N, (const double *)theta_dev,
                       tab.as<double2>());
    const size_t shm = sizeof(double2) * ((size_t)ifs::FFD * (2 * N + 1) + 256);
    hipLaunchKernelGGL(ifs::k_far_field, dim3((ndir + ifs::FFD - 1) / ifs::FFD), dim3(256), shm, c->stream, N, c->ns,
                       c->d_Leff.as<int32_t>(), dpos.as<double>(), c->prm.k, c->d_snm.as<double>(),
                       (const double2 *)x_dev, ndir, (const double *)theta_dev, (const double *)phi_dev,
                       tab.as<double2>(), (double2 *)out_dev);
    if (hipGetLastError() != hipSuccess) return IFS_E_HIP;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return IFS_E_HIP;  // tab, dpos are freed on return
    return IFS_OK;
}

extern "C" int ifs_far_field_device_noalloc(ifs_ctx *c, const void *x_dev, int64_t n, int32_t ndir,
                                            const void *theta_dev, const void *phi_dev, void *out_dev) {
    if (!c || !x_dev || !theta_dev || !phi_dev || !out_dev || n != ifs_vector_length(c) || ndir < 0)
        return IFS_E_ARG;
    if (ndir == 0) return IFS_OK;
    const int N = c->N;
    if ((int64_t)ndir * N * (N + 2) > (int64_t)IFS_FF_SCRATCH_ENTRIES) return IFS_E_ARG;
    if (!c->ff_scratch || !c->d_pos.p) return IFS_E_ARG;
    double2 *tab = (double2 *)c->ff_scratch;  // every entry the far-field kernel reads is written first
    hipLaunchKernelGGL(ifs::k_ff_tables, dim3(ndir), dim3(256), 0, c->stream, N, (const double *)theta_dev, tab);
    const size_t shm = sizeof(double2) * ((size_t)ifs::FFD * (2 * N + 1) + 256);
    hipLaunchKernelGGL(ifs::k_far_field, dim3((ndir + ifs::FFD - 1) / ifs::FFD), dim3(256), shm, c->stream, N, c->ns,
                       c->d_Leff.as<int32_t>(), c->d_pos.as<double>(), c->prm.k, c->d_snm.as<double>(),
                       (const double2 *)x_dev, ndir, (const double *)theta_dev, (const double *)phi_dev, tab,
                       (double2 *)out_dev);
    if (hipGetLastError() != hipSuccess) return IFS_E_HIP;
    return IFS_OK;
}

extern "C" int ifs_far_field(ifs_ctx *c, const double *x, int64_t n, int32_t ndir, const double *theta,
                             const double *phi, double *out) {
    if (!c || !x || !theta || !phi || !out || n != ifs_vector_length(c) || ndir < 0) return IFS_E_ARG;
    if (ndir == 0) return IFS_OK;
    ifs::DevBuf dx, dt, dph, dout;
    int rc;
    if ((rc = dx.alloc(16ull * n)) || (rc = dt.alloc(8ull * ndir)) || (rc = dph.alloc(8ull * ndir)) ||
        (rc = dout.alloc(32ull * ndir)))
        return rc;
    if (hipMemcpy(dx.p, x, 16ull * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dt.p, theta, 8ull * ndir, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dph.p, phi, 8ull * ndir, hipMemcpyHostToDevice) != hipSuccess)
        return IFS_E_HIP;
    if ((rc = ifs_far_field_device(c, dx.p, n, ndir, dt.p, dph.p, dout.p))) return rc;
    if (hipMemcpy(out, dout.p, 32ull * ndir, hipMemcpyDeviceToHost) != hipSuccess) return IFS_E_HIP;
    return IFS_OK;
}

extern "C" int ifs_amplitude_matrix(int32_t ndir, const double *phi, const double *Fx, const double *Fy, double *S) {
    if (ndir < 0 || !phi || !Fx || !Fy || !S) return IFS_E_ARG;
    const cplx I(0.0, 1.0);
    const cplx *fx = (const cplx *)Fx, *fy = (const cplx *)Fy;
    cplx *s = (cplx *)S;
    for (int i = 0; i < ndir; ++i) {
        double c = std::cos(phi[i]), sn = std::sin(phi[i]);
        cplx xt = fx[2 * i], xp = fx[2 * i + 1], yt = fy[2 * i], yp = fy[2 * i + 1];
        s[4 * i + 0] = I * (xp * sn - yp * c);         // S1
        s[4 * i + 1] = -I * (xt * c + yt * sn);        // S2
        s[4 * i + 2] = -I * (xt * sn - yt * c);        // S3
        s[4 * i + 3] = I * (xp * c + yp * sn);         // S4
    }
    return IFS_OK;
}

extern "C" int ifs_mueller(int32_t ndir, const double *S, double *M) {
    if (ndir < 0 || !S || !M) return IFS_E_ARG;
    const cplx *s = (const cplx *)S;
    for (int i = 0; i < ndir; ++i) {
        cplx S1 = s[4 * i], S2 = s[4 * i + 1], S3 = s[4 * i + 2], S4 = s[4 * i + 3];
        double a1 = std::norm(S1), a2 = std::norm(S2), a3 = std::norm(S3), a4 = std::norm(S4);
        double *m = M + 16 * i;
        m[0] = 0.5 * (a2 + a3 + a4 + a1);
        m[1] = 0.5 * (a2 - a3 + a4 - a1);
        m[2] = std::real(S2 * std::conj(S3) + S1 * std::conj(S4));
        m[3] = std::imag(S2 * std::conj(S3) - S1 * std::conj(S4));
        m[4] = 0.5 * (a2 - a4 + a3 - a1);
        m[5] = 0.5 * (a2 - a3 - a4 + a1);
        m[6] = std::real(S2 * std::conj(S3) - S1 * std::conj(S4));
        m[7] = std::imag(S2 * std::conj(S3) + S1 * std::conj(S4));
        m[8] = std::real(S2 * std::conj(S4) + S1 * std::conj(S3));
        m[9] = std::real(S2 * std::conj(S4) - S1 * std::conj(S3));
        m[10] = std::real(S1 * std::conj(S2) + S3 * std::conj(S4));
        m[11] = std::imag(S2 * std::conj(S1) + S4 * std::conj(S3));
        m[12] = std::imag(S4 * std::conj(S2) + S1 * std::conj(S3));
        m[13] = std::imag(S4 * std::conj(S2) - S1 * std::conj(S3));
        m[14] = std::imag(S1 * std::conj(S2) - S3 * std::conj(S4));
        m[15] = std::real(S1 * std::conj(S2) - S3 * std::conj(S4));
    }
    return IFS_OK;
}
