/*
 * ifs_oracle.c -- FP64 CPU restatement of the ItoFoggySim multi-sphere matvec.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  The product never links this file.
 *
 * The restatement follows the reference's algorithm: full Xu translation
 * matrices A/B by the Gaunt / Wigner-3j formula, regenerated for every ordered
 * sphere pair on every matvec, applied densely, scaled by the Mie coefficients,
 * plus the identity.  Intermediate sums run in long double so the oracle is
 * more accurate than the FP64 product it checks.
 *
 * Deviations from the literal Fortran, each stated where it happens:
 *   - 3j symbols come from the Schulten-Gordon three-term recurrence in j3
 *     (two-sided, Luscombe & Luban, Phys. Rev. E 57, 7274 (1998)), not from
 *     wigxjpf's exact prime factorisation.
 *   - factorial ratios are formed in long double, so N up to ~100 does not
 *     overflow where the reference's double products would.
 */
#include "ifs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef long double ld;
typedef long double complex ldc;

static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int fidx(int n, int m) { return n * n - 1 + m + n; }

/* ------------------------------------------------------------------ */
/* spherical Bessel functions                                          */
/* ------------------------------------------------------------------ */

/* j_n by Miller's downward recurrence, normalised with j_0 or j_1,
 * whichever is larger in magnitude.  Restates the role of Zhang-Jin SPHJ
 * called from lib_math_bessel.f90:479-510. */
static void sph_j_ld(ld x, int nmax, ld *out)
{
    if (x == 0.0L) {
        out[0] = 1.0L;
        for (int n = 1; n <= nmax; ++n) out[n] = 0.0L;
        return;
    }
    ld big = (ld)nmax > x ? (ld)nmax : x;
    int nstart = (int)big + 30 + (int)sqrtl(40.0L * big);
    ld *f = (ld *)malloc(sizeof(ld) * (nstart + 2));
    f[nstart + 1] = 0.0L;
    f[nstart] = 1e-300L;
    for (int n = nstart; n >= 1; --n) {
        f[n - 1] = (2.0L * n + 1.0L) / x * f[n] - f[n + 1];
        if (fabsl(f[n - 1]) > 1e3000L) {
            for (int k = n - 1; k <= nstart + 1; ++k) f[k] *= 1e-3000L;
        }
    }
    ld j0 = sinl(x) / x;
    ld j1 = sinl(x) / (x * x) - cosl(x) / x;
    ld scale = fabsl(j0) >= fabsl(j1) ? j0 / f[0] : j1 / f[1];
    for (int n = 0; n <= nmax; ++n) out[n] = f[n] * scale;
    free(f);
}

/* y_n by upward recurrence (stable upward); restates Zhang-Jin SPHY
 * called from lib_math_bessel.f90:593-624. */
static void sph_y_ld(ld x, int nmax, ld *out)
{
    out[0] = -cosl(x) / x;
    if (nmax >= 1) out[1] = -cosl(x) / (x * x) - sinl(x) / x;
    for (int n = 1; n < nmax; ++n) out[n + 1] = (2.0L * n + 1.0L) / x * out[n] - out[n - 1];
}

int ifs_or_sph_j(double x, int nmax, double *out)
{
    if (nmax < 0 || x < 0.0) return -1;
    ld *t = (ld *)malloc(sizeof(ld) * (nmax + 1));
    sph_j_ld((ld)x, nmax, t);
    for (int n = 0; n <= nmax; ++n) out[n] = (double)t[n];
    free(t);
    return 0;
}

int ifs_or_sph_y(double x, int nmax, double *out)
{
    if (nmax < 0 || x <= 0.0) return -1;
    ld *t = (ld *)malloc(sizeof(ld) * (nmax + 1));
    sph_y_ld((ld)x, nmax, t);
    for (int n = 0; n <= nmax; ++n) out[n] = (double)t[n];
    free(t);
    return 0;
}

/* ------------------------------------------------------------------ */
/* associated Legendre without Condon-Shortley phase                   */
/* lib_math_legendre.f90:128-188 (negative m rule :164-172), :210-271  */
/* ------------------------------------------------------------------ */
static void legendre_ld(ld x, int nmax, ld *out /* [(nmax+1)^2], out[n*n+m+n] */)
{
    ld s = sqrtl(fmaxl(0.0L, 1.0L - x * x));
    ld pmm = 1.0L;
    for (int m = 0; m <= nmax; ++m) {
        if (m > 0) pmm *= (2.0L * m - 1.0L) * s; /* no (-1)^m: Zhang-Jin convention */
        out[m * m + 2 * m] = pmm;                /* P_m^m at n = m: index m*m + m + m */
        if (m + 1 <= nmax) {
            int n = m + 1;
            out[n * n + m + n] = x * (2.0L * m + 1.0L) * pmm;
        }
        for (int n = m + 2; n <= nmax; ++n) {
            ld p1 = out[(n - 1) * (n - 1) + m + (n - 1)];
            ld p2 = out[(n - 2) * (n - 2) + m + (n - 2)];
            out[n * n + m + n] = ((2.0L * n - 1.0L) * x * p1 - (ld)(n + m - 1) * p2) / (ld)(n - m);
        }
    }
    /* negative orders: P_n^{-m} = (-1)^m (n-m)!/(n+m)! P_n^m */
    for (int n = 1; n <= nmax; ++n) {
        for (int m = 1; m <= n; ++m) {
            ld r = 1.0L;
            for (int i = n - m + 1; i <= n + m; ++i) r *= (ld)i;
            ld v = out[n * n + m + n] / r;
            if (m & 1) v = -v;
            out[n * n - m + n] = v;
        }
    }
}

int ifs_or_legendre(double x, int nmax, double *out)
{
    if (nmax < 0) return -1;
    int cnt = (nmax + 1) * (nmax + 1);
    ld *t = (ld *)malloc(sizeof(ld) * cnt);
    legendre_ld((ld)x, nmax, t);
    for (int i = 0; i < cnt; ++i) out[i] = (double)t[i];
    free(t);
    return 0;
}

/* (n+m)!/(n-m)!, any sign of m (lib_math_factorial.f90:102-290) */
static ld fact_ratio_ld(int n, int m)
{
    ld r = 1.0L;
    if (m >= 0) {
        for (int i = n - m + 1; i <= n + m; ++i) r *= (ld)i;
        return r;
    }
    for (int i = n + m + 1; i <= n - m; ++i) r *= (ld)i;
    return 1.0L / r;
}

double ifs_or_fact_ratio(int n, int m) { return (double)fact_ratio_ld(n, m); }

/* ------------------------------------------------------------------ */
/* Wigner 3j: Schulten-Gordon recurrence in j3, two-sided              */
/* (K. Schulten, R. G. Gordon, J. Math. Phys. 16 (1975) 1961; the      */
/* region logic of Luscombe & Luban, Phys. Rev. E 57 (1998) 7274).     */
/* The family f(j) = (j1 j2 j; m1 m2 m3), j = jmin..jmax, obeys        */
/*   j A(j+1) f(j+1) + B(j) f(j) + (j+1) A(j) f(j-1) = 0.              */
/* Near each end of the range f decays towards that end (non-classical */
/* regions); a recurrence is only run in the direction in which f      */
/* grows, or inside the oscillating classical region in between:       */
/* backward from jmax until |f| peaks, forward from jmin until |f|     */
/* peaks and on through the classical region, matched where both are   */
/* accurate.  (An earlier version ran the forward pass past a          */
/* monotonically decaying start, e.g. (20 22 j; -20 22 -2), and lost   */
/* the sign of the whole family.)                                      */
/* ------------------------------------------------------------------ */
static int w3j_all_ld(int j1, int j2, int m1, int m2, ld *f, int *j3min_out)
{
    int m3 = -m1 - m2;
    if (abs(m1) > j1 || abs(m2) > j2) return 0;
    int jmin = imax(abs(j1 - j2), abs(m3));
    int jmax = j1 + j2;
    if (jmin > jmax) return 0;
    int nv = jmax - jmin + 1;
    *j3min_out = jmin;

#define W3A(j) sqrtl(((ld)(j) * (j) - (ld)(j1 - j2) * (j1 - j2)) * ((ld)(j1 + j2 + 1) * (j1 + j2 + 1) - (ld)(j) * (j)) * ((ld)(j) * (j) - (ld)m3 * m3))
#define W3B(j) (-(2.0L * (j) + 1.0L) * ((ld)j1 * (j1 + 1) * m3 - (ld)j2 * (j2 + 1) * m3 - (ld)(j) * ((j) + 1) * (ld)(m2 - m1)))

    ld sgn_top = ((j1 - j2 - m3) & 1) ? -1.0L : 1.0L;
    if (nv == 1) {
        f[0] = sgn_top / sqrtl(2.0L * jmin + 1.0L);
        return 1;
    }
    ld *g = (ld *)calloc(nv, sizeof(ld)); /* backward values, valid on [gb, nv-1] */
    ld *h = (ld *)calloc(nv, sizeof(ld)); /* forward values, valid on [0, hf] */

    /* backward from jmax while |g| grows; two more steps into the classical region */
    g[nv - 1] = 1.0L;
    int gb = nv - 1, peak_b = nv - 1, extra = -1;
    for (int i = nv - 1; i >= 1; --i) {
        int j = jmin + i;
        ld gp1 = (i + 1 < nv) ? g[i + 1] : 0.0L;
        ld a = W3A(j);
        if (a == 0.0L) break;
        g[i - 1] = -((ld)j * (i + 1 < nv ? W3A(j + 1) : 0.0L) * gp1 + W3B(j) * g[i]) / ((ld)(j + 1) * a);
        gb = i - 1;
        if (fabsl(g[i - 1]) > 1e2000L) {
            for (int k = i - 1; k < nv; ++k) g[k] *= 1e-2000L;
        }
        if (extra < 0) {
            if (fabsl(g[i - 1]) >= fabsl(g[i])) {
                peak_b = i - 1;
            } else {
                extra = 2;
            }
        }
        if (extra >= 0 && extra-- == 0) break;
    }

    /* forward from jmin while |h| grows, then on to the backward peak */
    h[0] = 1.0L;
    int hf = 0, peak_f = 0;
    if (jmin == 0) {
        /* j1 = j2, m3 = 0: (j j 1; m -m 0) / (j j 0; m -m 0) = -m1 ... up to a sign
           fixed by the recurrence at j = 1; take the ratio from the exact forms
           (j j 0; m -m 0) = (-1)^(j-m)/sqrt(2j+1), (j j 1; m -m 0) = (-1)^(j-m) m/sqrt(j(j+1)(2j+1)) */
        h[1] = (j1 > 0) ? (ld)m1 / sqrtl((ld)j1 * (j1 + 1)) : 0.0L;
        hf = 1;
    } else {
        h[1] = -W3B(jmin) * h[0] / ((ld)jmin * W3A(jmin + 1));
        hf = 1;
    }
    if (fabsl(h[1]) >= fabsl(h[0])) peak_f = 1;
    int growing = fabsl(h[1]) >= fabsl(h[0]);
    int target = peak_b; /* the forward pass must reach the backward peak (or its own) */
    for (int i = 1; i < nv - 1; ++i) {
        if (!growing && i >= target) break;
        int j = jmin + i;
        ld den = (ld)j * W3A(j + 1);
        if (den == 0.0L) break;
        h[i + 1] = -(W3B(j) * h[i] + (ld)(j + 1) * W3A(j) * h[i - 1]) / den;
        hf = i + 1;
        if (fabsl(h[i + 1]) > 1e2000L) {
            for (int k = 0; k <= i + 1; ++k) h[k] *= 1e-2000L;
        }
        if (growing) {
            if (fabsl(h[i + 1]) >= fabsl(h[i])) {
                peak_f = i + 1;
            } else {
                growing = 0;
                if (peak_f >= peak_b) break; /* the two accurate ranges overlap */
            }
        }
    }

    /* match: both passes are accurate on [max(gb, ...), hf] around the backward peak */
    int im;
    if (peak_f >= peak_b) {
        im = peak_b; /* overlap: both accurate from peak_b to peak_f */
    } else {
        im = imin(peak_b, hf);
    }
    if (im < gb) im = gb;
    if (im > hf) im = hf;
    ld num = 0.0L, den = 0.0L;
    for (int i = imax(gb, im - 2); i <= imin(hf, im + 2); ++i) {
        num += h[i] * g[i];
        den += g[i] * g[i];
    }
    ld lam = den > 0.0L ? num / den : 0.0L;
    if (lam == 0.0L || hf < gb) {
        for (int i = 0; i < nv; ++i) f[i] = (i >= gb) ? g[i] : 0.0L;
    } else {
        for (int i = 0; i < nv; ++i) f[i] = (i <= im) ? h[i] : lam * g[i];
    }
    free(g);
    free(h);

    /* normalise: sum (2 j3 + 1) f^2 = 1, sign fixed at j3max */
    ld s = 0.0L;
    for (int i = 0; i < nv; ++i) s += (2.0L * (jmin + i) + 1.0L) * f[i] * f[i];
    ld nrm = 1.0L / sqrtl(s);
    if ((f[nv - 1] < 0.0L) != (sgn_top < 0.0L)) nrm = -nrm;
    for (int i = 0; i < nv; ++i) f[i] *= nrm;
#undef W3A
#undef W3B
    return nv;
}

int ifs_or_wigner3j_all(int j1, int j2, int m1, int m2, double *out, int *j3min)
{
    if (j1 < 0 || j2 < 0) return -1;
    ld *t = (ld *)malloc(sizeof(ld) * (j1 + j2 + 2));
    int nv = w3j_all_ld(j1, j2, m1, m2, t, j3min);
    for (int i = 0; i < nv; ++i) out[i] = (double)t[i];
    free(t);
    return nv;
}

static ld w3j_single_ld(int j1, int j2, int j3, int m1, int m2, int m3)
{
    if (m1 + m2 + m3 != 0) return 0.0L;
    if (j3 < abs(j1 - j2) || j3 > j1 + j2 || abs(m3) > j3) return 0.0L;
    ld *t = (ld *)malloc(sizeof(ld) * (j1 + j2 + 2));
    int jmin = 0;
    int nv = w3j_all_ld(j1, j2, m1, m2, t, &jmin);
    ld v = 0.0L;
    if (nv > 0 && j3 >= jmin && j3 < jmin + nv) v = t[j3 - jmin];
    free(t);
    return v;
}

double ifs_or_wigner3j(int j1, int j2, int j3, int m1, int m2, int m3)
{
    return (double)w3j_single_ld(j1, j2, j3, m1, m2, m3);
}

/* ------------------------------------------------------------------ */
/* Xu's a and b (lib_math_vector_spherical_harmonics.f90:2283-2349)     */
/* ------------------------------------------------------------------ */
static void xu_ab_ld(int m, int n, int mu, int nu, int p, const ld *w_mmu, int jmin_mmu, int nv_mmu,
                     ld w000, int calc_a, int calc_b, ld *a, ld *b)
{
    ld sgn = ((m + mu) & 1) ? -1.0L : 1.0L;
    ld f1 = fact_ratio_ld(n, m);
    ld f2 = fact_ratio_ld(nu, mu);
    *a = 0.0L;
    *b = 0.0L;
    if (calc_a) {
        ld f3 = fact_ratio_ld(p, -(m + mu)); /* (p-m-mu)!/(p+m+mu)! */
        ld w1 = (p >= jmin_mmu && p < jmin_mmu + nv_mmu) ? w_mmu[p - jmin_mmu] : 0.0L;
        *a = sgn * (2.0L * p + 1.0L) * sqrtl(f1 * f2 * f3) * w1 * w000;
    }
    if (calc_b) {
        ld f3 = fact_ratio_ld(p + 1, -(m + mu));
        int pp = p + 1;
        ld w1 = (pp >= jmin_mmu && pp < jmin_mmu + nv_mmu) ? w_mmu[pp - jmin_mmu] : 0.0L;
        *b = sgn * (2.0L * p + 3.0L) * sqrtl(f1 * f2 * f3) * w1 * w000;
    }
}

void ifs_or_xu_ab(int m, int n, int mu, int nu, int p, double *a, double *b)
{
    ld *w = (ld *)malloc(sizeof(ld) * (n + nu + 2));
    ld *w0 = (ld *)malloc(sizeof(ld) * (n + nu + 2));
    int jm = 0, j0 = 0;
    int nv = w3j_all_ld(n, nu, m, mu, w, &jm);
    int nv0 = w3j_all_ld(n, nu, 0, 0, w0, &j0);
    ld w000 = (p >= j0 && p < j0 + nv0) ? w0[p - j0] : 0.0L;
    ld av, bv;
    int ca = (p >= abs(m + mu));
    int cb = (p + 1 >= abs(m + mu));
    xu_ab_ld(m, n, mu, nu, p, w, jm, nv, w000, ca, cb, &av, &bv);
    *a = (double)av;
    *b = (double)bv;
    free(w);
    free(w0);
}

/* ------------------------------------------------------------------ */
/* Translation matrices A/B (lib_math_vector_spherical_harmonics.f90:1354-1585) */
/* ------------------------------------------------------------------ */
static void radial_ld(int zsel, ld x, int pmax, ldc *z)
{
    ld *jj = (ld *)malloc(sizeof(ld) * (pmax + 1));
    ld *yy = (ld *)malloc(sizeof(ld) * (pmax + 1));
    sph_j_ld(x, pmax, jj);
    if (zsel != 1) sph_y_ld(x, pmax, yy);
    for (int p = 0; p <= pmax; ++p) {
        switch (zsel) {
        case 1: z[p] = jj[p]; break;
        case 2: z[p] = yy[p]; break;
        case 3: z[p] = jj[p] + I * yy[p]; break;
        default: z[p] = jj[p] - I * yy[p]; break;
        }
    }
    free(jj);
    free(yy);
}

static inline ldc ipow(int p)
{
    switch (((p % 4) + 4) % 4) {
    case 0: return 1.0L;
    case 1: return I;
    case 2: return -1.0L;
    default: return -I;
    }
}

/* Core: fills A/B rows for n = 1..nmax, cols nu = 1..numax, using
 * precomputed radial z[0..pmax] and Legendre table P[(pmax+1)^2]. */
static void translation_core(int nmax, int numax, int zsel_real, const ldc *z, const ld *P, double phi,
                             double complex *A, double complex *B)
{
    (void)zsel_real;
    int ncol = numax * (numax + 2);
    int jmax3 = nmax + numax + 2;
    ld *w = (ld *)malloc(sizeof(ld) * jmax3);
    ld *w0 = (ld *)malloc(sizeof(ld) * jmax3);
    for (int n = 1; n <= nmax; ++n) {
        for (int nu = 1; nu <= numax; ++nu) {
            int j0 = 0;
            int nv0 = w3j_all_ld(n, nu, 0, 0, w0, &j0);
            for (int m = -n; m <= n; ++m) {
                for (int mu = -nu; mu <= nu; ++mu) {
                    int M = -m; /* the reference calls ab_xu_cruzan_eq34(-m, n, mu, nu, p) */
                    int jm = 0;
                    int nv = w3j_all_ld(n, nu, M, mu, w, &jm);
                    int dm = mu - m;
                    int qa = imin(imin(n, nu), (n + nu - abs(M + mu)) / 2);
                    int qb = imin(imin(n, nu), (n + nu + 1 - abs(M + mu)) / 2);
                    int qmx = imax(qa, qb);
                    ld fac = (2.0L * nu + 1.0L) * fact_ratio_ld(n, m) * fact_ratio_ld(nu, -mu) / (2.0L * n * (n + 1));
                    ldc sa = 0.0L, sb = 0.0L;
                    for (int q = 0; q <= qmx; ++q) {
                        int p = n + nu - 2 * q;
                        int ca = (q <= qa);
                        int cb = (q >= 1 && q <= qb);
                        ld w000 = (p >= j0 && p < j0 + nv0) ? w0[p - j0] : 0.0L;
                        ld av, bv;
                        xu_ab_ld(M, n, mu, nu, p, w, jm, nv, w000, ca, cb, &av, &bv);
                        if (ca) {
                            ld t = (ld)(n * (n + 1) + nu * (nu + 1) - p * (p + 1)) * av * P[p * p + dm + p];
                            sa += ipow(p) * t * z[p];
                        }
                        if (cb) {
                            ld r = sqrtl(((ld)(p + 1) * (p + 1) - (ld)(n - nu) * (n - nu)) *
                                         ((ld)(n + nu + 1) * (n + nu + 1) - (ld)(p + 1) * (p + 1)));
                            ld t = r * bv * P[(p + 1) * (p + 1) + dm + (p + 1)];
                            sb += ipow(p + 1) * t * z[p + 1];
                        }
                    }
                    ld ang = (ld)dm * (ld)phi;
                    ldc e = cosl(ang) + I * sinl(ang);
                    ld sgnm = (m & 1) ? -1.0L : 1.0L; /* (-1)^(-m) */
                    ldc Av = sgnm * fac * e * sa;
                    ldc Bv = -sgnm * fac * e * sb;
                    size_t o = (size_t)fidx(n, m) * ncol + fidx(nu, mu);
                    A[o] = (double complex)Av;
                    B[o] = (double complex)Bv;
                }
            }
        }
    }
    free(w);
    free(w0);
}

int ifs_or_translation_ab(double x, double theta, double phi, int nmax, int numax, int zsel,
                          double complex *A, double complex *B)
{
    if (nmax < 1 || numax < 1 || zsel < 1 || zsel > 4) return -1;
    int pmax = nmax + numax + 1;
    ldc *z = (ldc *)malloc(sizeof(ldc) * (pmax + 1));
    ld *P = (ld *)malloc(sizeof(ld) * (pmax + 1) * (pmax + 1));
    if (x <= 0.0 && zsel != 1) {
        free(z);
        free(P);
        return -2;
    }
    radial_ld(zsel, (ld)x, pmax, z);
    legendre_ld(cosl((ld)theta), pmax, P);
    translation_core(nmax, numax, zsel, z, P, phi, A, B);
    free(z);
    free(P);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Mie coefficients (Barber & Hill), lib_mie_ss_helper_functions.f90   */
/* :610-722 (a_n, b_n), :819-865 (A_n downward), :900-924, :937-951   */
/* ------------------------------------------------------------------ */
int ifs_or_mie_ab(double x, double m_re, double m_im, int nmax, double complex *a, double complex *b)
{
    if (nmax < 1 || x <= 0.0) return -1;
    double complex m = m_re + I * m_im;
    double complex mx = m * x;
    int nc = (int)ceil(x + 4.05 * pow(x, 1.0 / 3.0) + 2.0);
    int nstart = imax(nc, (int)cabs(mx)) + 15;
    /* the reference leaves A_n above nstart-1 unset; start high enough that
     * every A_n, n <= nmax, comes from the recurrence */
    if (nstart < nmax + 1) nstart = nmax + 1;
    double complex *An = (double complex *)malloc(sizeof(double complex) * (imax(nstart, nmax) + 2));
    double complex buf = 0.0;
    for (int i = nstart; i >= 2; --i) {
        double complex ndm = (double)i / mx;
        buf = ndm - 1.0 / (buf + ndm);
        if (i - 1 <= nmax) An[i - 1] = buf;
    }
    ld *jj = (ld *)malloc(sizeof(ld) * (nmax + 1));
    ld *yy = (ld *)malloc(sizeof(ld) * (nmax + 1));
    sph_j_ld((ld)x, nmax, jj);
    sph_y_ld((ld)x, nmax, yy);
    for (int n = 1; n <= nmax; ++n) {
        double complex jn = (double)jj[n], jn1 = (double)jj[n - 1];
        double complex hn = (double)jj[n] + I * (double)yy[n];
        double complex hn1 = (double)jj[n - 1] + I * (double)yy[n - 1];
        double complex ndx = (double)n / x;
        double complex t = An[n] / m + ndx;
        a[n - 1] = (t * jn - jn1) / (t * hn - hn1);
        t = m * An[n] + ndx;
        b[n - 1] = (t * jn - jn1) / (t * hn - hn1);
    }
    free(An);
    free(jj);
    free(yy);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Direct matvec (lib_mie_ms_solver_interface.f90:806-998)             */
/* ------------------------------------------------------------------ */

/* interaction of source l on target j, accumulated into bj (one block of the
 * flat global-N layout).  Rows n = 1..nrow, cols nu = 1..ncolN. */
static void pair_apply(const double *pj, const double *pl, int nrow, int ncolN, int nglob,
                       const double complex *anj, const double complex *bnj, double k, int zsel,
                       const double complex *xl, double complex *bj, double complex *A, double complex *B)
{
    double d[3] = {pl[0] - pj[0], pl[1] - pj[1], pl[2] - pj[2]}; /* d_0_l - d_0_j (:941-943) */
    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double theta = r > 0.0 ? acos(fmax(-1.0, fmin(1.0, d[2] / r))) : 0.0;
    double phi = atan2(d[1], d[0]);
    ifs_or_translation_ab(k * r, theta, phi, nrow, ncolN, zsel, A, B);
    int blk = nglob * (nglob + 2);
    int ncol = ncolN * (ncolN + 2);
    for (int n = 1; n <= nrow; ++n) {
        for (int m = -n; m <= n; ++m) {
            int row = fidx(n, m);
            ldc sa = 0.0L, sb = 0.0L;
            for (int c = 0; c < ncol; ++c) {
                ldc A_ = A[(size_t)row * ncol + c];
                ldc B_ = B[(size_t)row * ncol + c];
                ldc xa = xl[c], xb = xl[blk + c];
                sa += A_ * xa + B_ * xb;
                sb += B_ * xa + A_ * xb;
            }
            bj[row] += anj[n - 1] * (double complex)sa;
            bj[blk + row] += bnj[n - 1] * (double complex)sb;
        }
    }
}

static void set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

int ifs_or_matvec_direct(int ns, const double *pos, const int *lmax, int nglob,
                         const double complex *an, const double complex *bn, double k, int zsel,
                         int ref_row_zeroing, const double complex *x, double complex *b, int nthreads)
{
    if (ns < 1 || nglob < 1) return -1;
    int blk = nglob * (nglob + 2);
    set_threads(nthreads);
    int err = 0;
#pragma omp parallel
    {
        double complex *A = (double complex *)malloc(sizeof(double complex) * (size_t)blk * blk);
        double complex *B = (double complex *)malloc(sizeof(double complex) * (size_t)blk * blk);
        if (!A || !B) {
#pragma omp atomic write
            err = -3;
        }
#pragma omp for schedule(dynamic, 1)
        for (int j = 0; j < ns; ++j) {
            double complex *bj = b + (size_t)j * 2 * blk;
            memset(bj, 0, sizeof(double complex) * 2 * blk);
            if (!A || !B) continue;
            int nrow = imin(lmax[j], nglob);
            int active = ref_row_zeroing ? (lmax[j] >= nglob) : 1;
            for (int l = 0; l < ns; ++l) {
                if (l == j) {
                    const double complex *xj = x + (size_t)j * 2 * blk;
                    for (int i = 0; i < 2 * blk; ++i) bj[i] += xj[i];
                    continue;
                }
                if (!active) continue;
                int ncolN = imin(lmax[l], nglob);
                pair_apply(pos + 3 * j, pos + 3 * l, nrow, ncolN, nglob, an + (size_t)j * nglob,
                           bn + (size_t)j * nglob, k, zsel, x + (size_t)l * 2 * blk, bj, A, B);
            }
        }
        free(A);
        free(B);
    }
    return err;
}

int ifs_or_matvec_pairs(int ns, const double *pos, const int *lmax, int nglob,
                        const double complex *an, const double complex *bn, double k, int zsel,
                        int npairs, const int *tgt, const int *src,
                        const double complex *x, double complex *b, int nthreads)
{
    if (ns < 1 || nglob < 1 || npairs < 0) return -1;
    int blk = nglob * (nglob + 2);
    set_threads(nthreads);
    int err = 0;
#pragma omp parallel
    {
        double complex *A = (double complex *)malloc(sizeof(double complex) * (size_t)blk * blk);
        double complex *B = (double complex *)malloc(sizeof(double complex) * (size_t)blk * blk);
        double complex *acc = (double complex *)calloc((size_t)ns * 2 * blk, sizeof(double complex));
        if (!A || !B || !acc) {
#pragma omp atomic write
            err = -3;
        }
#pragma omp for schedule(dynamic, 1)
        for (int i = 0; i < npairs; ++i) {
            if (!A || !B || !acc) continue;
            int j = tgt[i], l = src[i];
            if (j == l) continue;
            pair_apply(pos + 3 * j, pos + 3 * l, imin(lmax[j], nglob), imin(lmax[l], nglob), nglob,
                       an + (size_t)j * nglob, bn + (size_t)j * nglob, k, zsel,
                       x + (size_t)l * 2 * blk, acc + (size_t)j * 2 * blk, A, B);
        }
#pragma omp critical
        {
            if (acc)
                for (size_t i = 0; i < (size_t)ns * 2 * blk; ++i) b[i] += acc[i];
        }
        free(A);
        free(B);
        free(acc);
    }
    return err;
}
