/*
 * ifs_oracle.h -- FP64 CPU restatement of the ItoFoggySim multi-sphere matvec.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (itofoggysim_amd/) links,
 * loads or calls this library; only tests/, the smoke check and the bench's
 * CPU-baseline leg use it, and only as the checker.
 *
 * Every function cites the reference file:line it restates.  Paths are
 * relative to the reference repository root.
 *
 * Flat index of (n, m) inside one a- or b-block (n >= 1, -n <= m <= n):
 *     idx(n, m) = n*n - 1 + (m + n)
 * (lib_math_type_operator.f90:2590-2645, make_array).
 */
#ifndef IFS_ORACLE_H
#define IFS_ORACLE_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

/* spherical Bessel j_n(x), n = 0..nmax (lib_math_bessel.f90:479-510, SPHJ) */
int ifs_or_sph_j(double x, int nmax, double *out);
/* spherical Neumann y_n(x), n = 0..nmax (lib_math_bessel.f90:593-624, SPHY) */
int ifs_or_sph_y(double x, int nmax, double *out);

/* associated Legendre P_n^m(x) without Condon-Shortley phase, -n <= m <= n,
 * n = 0..nmax; out[n*n + m + n] (lib_math_legendre.f90:128-271). */
int ifs_or_legendre(double x, int nmax, double *out);

/* (n+m)!/(n-m)! (lib_math_factorial.f90:233-290) */
double ifs_or_fact_ratio(int n, int m);

/* Wigner 3j (j1 j2 j3; m1 m2 m3) for integer arguments
 * (lib_math_wigner.f90:45-72 -> wigxjpf fwig3jj). */
double ifs_or_wigner3j(int j1, int j2, int j3, int m1, int m2, int m3);
/* all j3 for fixed (j1 j2 m1 m2); returns count, *j3min set. */
int ifs_or_wigner3j_all(int j1, int j2, int m1, int m2, double *out, int *j3min);

/* Xu's a(m,n,mu,nu,p) and b(...) (lib_math_vector_spherical_harmonics.f90:2283-2349) */
void ifs_or_xu_ab(int m, int n, int mu, int nu, int p, double *a, double *b);

/* Full translation matrices A, B (lib_math_vector_spherical_harmonics.f90:1354-1585).
 * x = k*|d|, (theta, phi) direction of d.  Rows (n,m) for n = 1..nmax, cols
 * (nu,mu) for nu = 1..numax.  A[idx(n,m) * ncol + idx(nu,mu)], ncol = numax(numax+2).
 * zsel: 1 = j_n, 2 = y_n, 3 = h1_n, 4 = h2_n. */
int ifs_or_translation_ab(double x, double theta, double phi, int nmax, int numax, int zsel,
                          double complex *A, double complex *B);

/* Mie coefficients a_n, b_n for n = 1..nmax, complex relative index m
 * (lib_mie_ss_helper_functions.f90:610-722, 819-924, 937-951). */
int ifs_or_mie_ab(double x, double m_re, double m_im, int nmax, double complex *a, double complex *b);

/* Direct matvec b = x + T A x in the reference's flat global-N layout
 * (FoggySim/src/lib/mie_theory/lib_mie_ms_solver_interface.f90:806-998).
 *   pos[3*ns]: centres, same length unit as 1/k
 *   lmax[ns] : per-sphere truncation L_j
 *   nglob    : global N (layout)
 *   an, bn   : [ns * nglob], T-matrix coefficients a_n^j, b_n^j (n = 1..nglob)
 *   zsel     : radial kind of the pair translation (3 = h1 is physical)
 *   ref_row_zeroing: 1 -> zero every interaction row of a sphere with L_j < N
 *   x, b     : [ns * 2 * nglob(nglob+2)]
 *   nthreads : OpenMP threads (<=0: library default) */
int ifs_or_matvec_direct(int ns, const double *pos, const int *lmax, int nglob,
                         const double complex *an, const double complex *bn, double k, int zsel,
                         int ref_row_zeroing, const double complex *x, double complex *b, int nthreads);

/* Same as above, restricted to the ordered pairs (tgt[i], src[i]), i < npairs,
 * accumulating T A x into b (no identity).  Used to time bounded samples. */
int ifs_or_matvec_pairs(int ns, const double *pos, const int *lmax, int nglob,
                        const double complex *an, const double complex *bn, double k, int zsel,
                        int npairs, const int *tgt, const int *src,
                        const double complex *x, double complex *b, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
