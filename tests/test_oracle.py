"""The CPU oracle against the reference's golden vectors and against what the
mathematics fixes (closed forms, independent evaluations, invariants)."""
import csv
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.special as sp

GOLD = os.path.join(os.path.dirname(__file__), "golden")
REFV = json.load(open(os.path.join(GOLD, "reference_vectors.json")))


# ---------------------------------------------------------------- Bessel
def test_sph_j_golden(oracle):
    g = REFV["sph_j_x5"]
    j = oracle.sph_j(g["x"], 40)
    assert np.max(np.abs((j - g["values"]) / np.array(g["values"]))) < g["rtol"]


def test_sph_y_golden(oracle):
    g = REFV["sph_y_x5"]
    y = oracle.sph_y(g["x"], 40)
    assert np.max(np.abs(1 - np.abs(y / np.array(g["values"])))) < g["rtol"]


def test_sph_h1_golden(oracle):
    g = REFV["sph_h1_x2"]
    j, y = oracle.sph_j(2.0, 40), oracle.sph_y(2.0, 40)
    for n, (re_, im_) in g["values"].items():
        n = int(n)
        # the reference compares logarithms of the real and imaginary parts
        # (magnitudes here, with equal signs)
        assert abs(math.log(abs(j[n])) - math.log(abs(re_))) < 2e-13 and np.sign(j[n]) == np.sign(re_)
        assert abs(math.log(abs(y[n])) - math.log(abs(im_))) < 2e-13 and np.sign(y[n]) == np.sign(im_)


@pytest.mark.parametrize("x", [0.3, 2.0, 7.5, 40.0, 300.0])
def test_bessel_independent(oracle, x):
    n = np.arange(0, 60)
    j, y = oracle.sph_j(x, 59), oracle.sph_y(x, 59)
    jr, yr = sp.spherical_jn(n, x), sp.spherical_yn(n, x)
    mask = np.abs(jr) > 1e-280
    assert np.max(np.abs(j[mask] - jr[mask]) / np.abs(jr[mask])) < 1e-11
    assert np.max(np.abs(y - yr) / np.abs(yr)) < 1e-11
    # Wronskian j_n y_{n-1} - j_{n-1} y_n = x^-2
    w = j[1:] * y[:-1] - j[:-1] * y[1:]
    assert np.max(np.abs(w * x * x - 1.0)) < 1e-10


# ---------------------------------------------------------------- Legendre
def test_legendre_no_condon_shortley(oracle):
    x = 0.37
    P = oracle.legendre(x, 12)
    for n in range(13):
        for m in range(0, n + 1):
            # scipy's lpmv carries the Condon-Shortley phase
            assert abs(P[n * n + m + n] - (-1) ** m * sp.lpmv(m, n, x)) < 1e-10 * max(1, abs(sp.lpmv(m, n, x)))
            neg = (-1) ** m * math.factorial(n - m) / math.factorial(n + m) * P[n * n + m + n]
            assert abs(P[n * n - m + n] - neg) < 1e-14 * max(1, abs(neg))
    P1 = oracle.legendre(1.0, 8)
    for n in range(9):
        for m in range(-n, n + 1):
            assert P1[n * n + m + n] == (1.0 if m == 0 else 0.0)


# ---------------------------------------------------------------- 3j
def _w3j_exact(j1, j2, j3, m1, m2, m3):
    if m1 + m2 + m3 != 0 or j3 < abs(j1 - j2) or j3 > j1 + j2:
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    f = math.factorial
    t1, t2, t3, t4, t5 = j2 - m1 - j3, j1 + m2 - j3, j1 + j2 - j3, j1 - m1, j2 + m2
    s = Fraction(0)
    for t in range(max(0, t1, t2), min(t3, t4, t5) + 1):
        s += Fraction((-1) ** t, f(t) * f(t - t1) * f(t - t2) * f(t3 - t) * f(t4 - t) * f(t5 - t))
    tri = Fraction(f(j1 + j2 - j3) * f(j1 - j2 + j3) * f(-j1 + j2 + j3), f(j1 + j2 + j3 + 1))
    pre = tri * f(j1 + m1) * f(j1 - m1) * f(j2 + m2) * f(j2 - m2) * f(j3 + m3) * f(j3 - m3)
    val = float(pre * s * s) ** 0.5
    return (-1) ** (j1 - j2 - m3) * (1 if s > 0 else -1) * val


def test_wigner3j_golden(oracle):
    g = REFV["wigner3j"]
    for j3, v in zip(g["j3"], g["values"]):
        assert abs(oracle.wigner3j(g["j1"], g["j2"], j3, g["m1"], g["m2"], g["m3"]) - v) < g["atol"]


def test_wigner3j_exact_racah(oracle):
    rng = np.random.default_rng(5)
    for _ in range(400):
        j1, j2 = (int(v) for v in rng.integers(0, 22, 2))
        j3 = int(rng.integers(abs(j1 - j2), j1 + j2 + 1))
        m1 = int(rng.integers(-j1, j1 + 1))
        m2 = int(rng.integers(-j2, j2 + 1))
        m3 = -m1 - m2
        if abs(m3) > j3:
            continue
        assert abs(oracle.wigner3j(j1, j2, j3, m1, m2, m3) - _w3j_exact(j1, j2, j3, m1, m2, m3)) < 1e-15


def test_wigner3j_monotone_families_exact(oracle):
    """Families whose lower non-classical region decays monotonically (stretched or
    nearly stretched m's, e.g. (20 22 j; -20 22 -2)): a two-sided recurrence that runs
    forward through such a region returns them with the wrong sign.  Exact Racah values
    up to j1 + j2 = 100."""
    fams = [(20, 22, -20, 22), (20, 22, 20, -22), (40, 12, -40, 12), (48, 50, -48, -2), (33, 17, 33, -17),
            (25, 25, 25, -25), (30, 45, -29, 44), (50, 50, -50, 3), (12, 40, 12, -40), (7, 31, -7, 0)]
    for j1, j2, m1, m2 in fams:
        m3 = -m1 - m2
        for j3 in range(max(abs(j1 - j2), abs(m3)), j1 + j2 + 1):
            ex = _w3j_exact(j1, j2, j3, m1, m2, m3)
            assert abs(oracle.wigner3j(j1, j2, j3, m1, m2, m3) - ex) < 1e-15 + 1e-13 * abs(ex), (j1, j2, j3, m1, m2)


# ---------------------------------------------------------------- Xu a, b
def test_xu_ab_golden(oracle):
    g = REFV["xu_ab"]
    for i, q in enumerate(g["q"]):
        p = g["n"] + g["nu"] - 2 * q
        a, b = oracle.xu_ab(g["m"], g["n"], g["mu"], g["nu"], p)
        assert abs(a - g["a"][i]) < g["atol"]
        assert abs(b - g["b"][i]) < g["atol"]


# ---------------------------------------------------------------- A/B translation coefficients
def _mantissa_err(erg, gt):
    if erg == 0.0:
        return abs(gt)
    e = int(math.log(abs(erg)) / math.log(10.0))
    return abs(gt / 10.0 ** e - erg / 10.0 ** e)


@pytest.mark.parametrize("fname,phi", [("ground_truth_AB_mnmunu.csv", 0.5),
                                       ("ground_truth_AB_mnmunu_minus_phi.csv", 2 * math.pi - 0.5)])
def test_translation_csv(oracle, fname, phi):
    """The reference's Mathematica CSVs (x=2, theta=0.5, h1, n,nu<=5), reference metric.
    The CSV itself carries ~1e-13 errors (checked against 40-digit arithmetic below),
    so the bound is 5e-13 on the mantissa instead of the reference's 1e-13."""
    rows = list(csv.reader(open(os.path.join(GOLD, fname))))[1:]
    A, B = oracle.translation_ab(2.0, 0.5, phi, 5, 5, 3)
    assert len(rows) == 35 * 35
    worst = 0.0
    for r in rows:
        m, n, mu, nu = (int(float(v)) for v in r[:4])
        v = [float(t) for t in r[4:]]
        a = A[oracle.fidx(n, m), oracle.fidx(nu, mu)]
        b = B[oracle.fidx(n, m), oracle.fidx(nu, mu)]
        for c, gt in ((a.real, v[0]), (a.imag, v[1]), (b.real, v[2]), (b.imag, v[3])):
            worst = max(worst, _mantissa_err(c, gt))
    assert worst < 5e-13


def test_translation_high_precision_spot():
    """Spot entries where the CSV and the oracle differ most, recomputed with the
    same Gaunt formula in 40-digit arithmetic (mpmath + sympy 3j)."""
    mp = pytest.importorskip("mpmath")
    wig = pytest.importorskip("sympy.physics.wigner")
    from oracle import oracle as O
    mp.mp.dps = 40

    def w3j(*a):
        return mp.mpf(str(wig.wigner_3j(*a).evalf(50)))

    def fr(n, m):
        return mp.factorial(n + m) / mp.factorial(n - m)

    def P(p, k, x):
        if k >= 0:
            return (-1) ** k * mp.legenp(p, k, x)
        return (-1) ** (-k) * mp.factorial(p + k) / mp.factorial(p - k) * P(p, -k, x)

    def h(p, x):
        return mp.sqrt(mp.pi / (2 * x)) * (mp.besselj(p + 0.5, x) + 1j * mp.bessely(p + 0.5, x))

    def AB(x, th, ph, n, m, nu, mu):
        M = -m
        qa = min(n, nu, (n + nu - abs(M + mu)) // 2)
        qb = min(n, nu, (n + nu + 1 - abs(M + mu)) // 2)
        fac = (2 * nu + 1) * fr(n, m) / fr(nu, mu) / (2 * n * (n + 1))
        sa = sb = 0
        c = mp.cos(th)
        for q in range(0, max(qa, qb) + 1):
            p = n + nu - 2 * q
            w0 = w3j(n, nu, p, 0, 0, 0)
            if q <= qa:
                a = (-1) ** (M + mu) * (2 * p + 1) * mp.sqrt(fr(n, M) * fr(nu, mu) * mp.factorial(p - M - mu)
                                                             / mp.factorial(p + M + mu)) * w3j(n, nu, p, M, mu, -M - mu) * w0
                sa += (1j) ** p * (n * (n + 1) + nu * (nu + 1) - p * (p + 1)) * a * P(p, mu - m, c) * h(p, x)
            if 1 <= q <= qb:
                b = (-1) ** (M + mu) * (2 * p + 3) * mp.sqrt(fr(n, M) * fr(nu, mu) * mp.factorial(p + 1 - M - mu)
                                                             / mp.factorial(p + 1 + M + mu)) * w3j(n, nu, p + 1, M, mu, -M - mu) * w0
                sb += (1j) ** (p + 1) * mp.sqrt(((p + 1) ** 2 - (n - nu) ** 2) * ((n + nu + 1) ** 2 - (p + 1) ** 2)) \
                    * b * P(p + 1, mu - m, c) * h(p + 1, x)
        e = mp.exp(1j * (mu - m) * ph)
        return complex((-1) ** m * fac * e * sa), complex(-(-1) ** m * fac * e * sb)

    A, B = O.translation_ab(2.0, 0.5, 0.5, 5, 5, 3)
    for (m, n, mu, nu) in [(-3, 3, -4, 4), (0, 3, -1, 5), (1, 3, 2, 5)]:
        a, b = AB(mp.mpf(2), mp.mpf("0.5"), mp.mpf("0.5"), n, m, nu, mu)
        ia, ib = O.fidx(n, m), O.fidx(nu, mu)
        assert abs(A[ia, ib] - a) <= 1e-14 * abs(a)
        assert abs(B[ia, ib] - b) <= 1e-14 * abs(b)


def _snm(N):
    s = []
    for n in range(1, N + 1):
        for m in range(-n, n + 1):
            s.append(math.sqrt(n * (n + 1) * (2 * n + 1) * math.factorial(n - m) / math.factorial(n + m)))
    return np.array(s)


def test_translation_regular_unitary(oracle):
    """Regular (j_n) translation is unitary in the normalised basis: low rows of a
    well-truncated operator have unit norm."""
    N = 22
    A, B = oracle.translation_ab(2.0, 0.7, 0.4, N, N, 1)
    S = _snm(N)
    An = S[:, None] * A / S[None, :]
    Bn = S[:, None] * B / S[None, :]
    T = np.block([[An, Bn], [Bn, An]])
    rows = np.sqrt(np.sum(np.abs(T) ** 2, axis=1))
    assert np.max(np.abs(rows[:15] - 1.0)) < 1e-10


def test_translation_regular_composition(oracle):
    """T(d1) T(d2) = T(d1 + d2) for regular waves (converged truncation)."""
    def T(d, N, Nin):
        r = np.linalg.norm(d)
        A, B = oracle.translation_ab(r, math.acos(d[2] / r), math.atan2(d[1], d[0]), N, Nin, 1)
        return np.block([[A, B], [B, A]])
    d1 = np.array([0.3, -0.2, 0.5])
    d2 = np.array([-0.1, 0.4, 0.2])
    N, Nm = 4, 24
    lhs = T(d1, N, Nm) @ T(d2, Nm, N)
    rhs = T(d1 + d2, N, N)
    assert np.max(np.abs(lhs - rhs)) / np.max(np.abs(rhs)) < 1e-11


def test_translation_plane_wave_high_order(oracle):
    """Closed form at high order: a plane wave expanded about c1 and translated with the
    regular coefficients to c2 is the plane wave expanded about c2,
        (A p1 + B q1, B p1 + A q1) = (p2, q2),
    exact once the source order exceeds the target order plus k|c2 - c1| by a margin.
    Source orders up to 75 and t = 20 exercise 3j families far beyond the CSV fixtures;
    compared in the normalised basis."""
    from oracle import farfield as FF
    Lt, Ls, t = 20, 75, 20.0
    c1 = np.zeros(3)
    d = t * np.array([math.sqrt(1 - 0.09) * math.cos(0.4), math.sqrt(1 - 0.09) * math.sin(0.4), 0.3])
    P1 = FF.plane_wave(c1[None, :], Ls, 1.0, 0.7, 0.3)
    P2 = FF.plane_wave(d[None, :], Lt, 1.0, 0.7, 0.3)
    nS, nT = Ls * (Ls + 2), Lt * (Lt + 2)
    A, B = oracle.translation_ab(t, math.acos(0.3), 0.4, Lt, Ls, 1)
    ra = A @ P1[:nS] + B @ P1[nS:]
    rb = B @ P1[:nS] + A @ P1[nS:]
    S = _snm(Lt)
    num = np.linalg.norm(np.concatenate([(ra - P2[:nT]) * S, (rb - P2[nT:]) * S]))
    den = np.linalg.norm(np.concatenate([P2[:nT] * S, P2[nT:] * S]))
    assert num / den < 1e-12


# ---------------------------------------------------------------- Mie
def test_mie_golden(oracle):
    g = REFV["mie_s2_real"]
    a, b = oracle.mie_ab(g["x"], complex(*g["m"]), 4)
    ga = np.array([complex(*v) for v in g["a"]])
    gb = np.array([complex(*v) for v in g["b"]])
    assert np.max(np.abs(a / ga - 1)) < g["rtol"]
    assert np.max(np.abs(b / gb - 1)) < g["rtol"]
    g = REFV["mie_s2_cmplx"]
    a, b = oracle.mie_ab(g["x"], complex(*g["m"]), 4)
    assert np.max(np.abs(a - np.array([complex(*v) for v in g["a"]]))) < g["atol"]
    assert np.max(np.abs(b - np.array([complex(*v) for v in g["b"]]))) < g["atol"]


@pytest.mark.parametrize("x,m", [(10.0, 1.5), (3.7, 1.333 + 1e-9j), (25.0, 1.318 + 1e-4j), (1.2, 1.6 + 0.2j)])
def test_mie_textbook(oracle, x, m):
    """Lorenz-Mie from Riccati-Bessel functions evaluated by scipy (independent path)."""
    N = int(x + 4 * x ** (1 / 3) + 2)
    a, b = oracle.mie_ab(x, m, N)
    n = np.arange(1, N + 1)
    mx = m * x
    psi = lambda z: z * sp.spherical_jn(n, z)
    dpsi = lambda z: sp.spherical_jn(n, z) + z * sp.spherical_jn(n, z, derivative=True)
    h = lambda z: sp.spherical_jn(n, z) + 1j * sp.spherical_yn(n, z)
    dh = lambda z: sp.spherical_jn(n, z, True) + 1j * sp.spherical_yn(n, z, True)
    xi, dxi = x * h(x), h(x) + x * dh(x)
    ar = (m * psi(mx) * dpsi(x) - psi(x) * dpsi(mx)) / (m * psi(mx) * dxi - xi * dpsi(mx))
    br = (psi(mx) * dpsi(x) - m * psi(x) * dpsi(mx)) / (psi(mx) * dxi - m * xi * dpsi(mx))
    assert np.max(np.abs(a - ar)) < 1e-9
    assert np.max(np.abs(b - br)) < 1e-9
    if np.imag(m) == 0:
        assert np.all(np.abs(a - 0.5) <= 0.5 + 1e-12)


# ---------------------------------------------------------------- matvec structure
def test_matvec_identity_when_T_zero(oracle):
    rng = np.random.default_rng(1)
    pos = rng.uniform(0, 5, (3, 3))
    L = np.array([3, 2, 4], dtype=np.int32)
    N = 4
    z = np.zeros((3, N), complex)
    x = rng.standard_normal(3 * 2 * N * (N + 2)) + 0j
    assert np.array_equal(oracle.matvec_direct(pos, L, N, z, z, 1.0, x), x)


def test_matvec_two_sphere_brute_force(oracle):
    """Two spheres: b_j = x_j + a_n (A x_l^a + B x_l^b) assembled from the oracle's
    own A/B of d = d_l - d_j; row zeroing of the reference for L_j < N."""
    pos = np.array([[0.0, 0.0, 0.0], [0.4, -0.3, 1.1]])
    L = np.array([3, 2], dtype=np.int32)
    N = 3
    k = 2.1
    rng = np.random.default_rng(2)
    an = rng.standard_normal((2, N)) + 1j * rng.standard_normal((2, N))
    bn = rng.standard_normal((2, N)) + 1j * rng.standard_normal((2, N))
    x = rng.standard_normal(2 * 2 * N * (N + 2)) + 1j * rng.standard_normal(2 * 2 * N * (N + 2))
    blk = N * (N + 2)
    out = oracle.matvec_direct(pos, L, N, an, bn, k, x)
    d = pos[1] - pos[0]
    r = np.linalg.norm(d)
    A, B = oracle.translation_ab(k * r, math.acos(d[2] / r), math.atan2(d[1], d[0]), 3, 2, 3)
    xl_a, xl_b = x[2 * blk:2 * blk + 8], x[2 * blk + blk:2 * blk + blk + 8]
    nidx = np.concatenate([[n] * (2 * n + 1) for n in range(1, 4)]) - 1
    exp_a = x[:blk] + an[0][nidx] * (A @ xl_a + B @ xl_b)
    exp_b = x[blk:2 * blk] + bn[0][nidx] * (B @ xl_a + A @ xl_b)
    assert np.allclose(out[:blk], exp_a, rtol=1e-13, atol=1e-13 * np.abs(exp_a).max())
    assert np.allclose(out[blk:2 * blk], exp_b, rtol=1e-13, atol=1e-13 * np.abs(exp_b).max())
    zr = oracle.matvec_direct(pos, L, N, an, bn, k, x, ref_row_zeroing=True)
    assert np.array_equal(zr[2 * blk:], x[2 * blk:])  # sphere 2 has L < N: rows stay identity


def test_fmm_restatement_converges_to_direct(oracle):
    """The oracle's MLFMM approaches its direct matvec: ~6 digits with the box rule
    at d0 = 6 on a two-level tree, and a deeper (4-level) tree improves with d0."""
    from oracle import fmm_ref
    rng = np.random.default_rng(11)
    N, k = 2, 0.6
    for ns, s, d0s, bound in ((24, 2, (6.0,), 1e-5), (40, 1, (3.0, 6.0), None)):
        pos = rng.uniform(0.0, 8.0, (ns, 3))
        L = np.full(ns, 2, dtype=np.int32)
        an = rng.standard_normal((ns, N)) * 0.3 + 0j
        bn = rng.standard_normal((ns, N)) * 0.3 + 0j
        x = rng.standard_normal(ns * 2 * N * (N + 2)) + 1j * rng.standard_normal(ns * 2 * N * (N + 2))
        direct = oracle.matvec_direct(pos, L, N, an, bn, k, x)
        errs = []
        for d0 in d0s:
            f, tr = fmm_ref.fmm_matvec(pos, L, N, an, bn, k, x, s=s, d0=d0)
            errs.append(np.linalg.norm(f - direct) / np.linalg.norm(direct - x))
        if bound is not None:
            assert errs[0] < bound
        else:
            assert tr["lmax"] >= 3
            assert errs[1] < 0.5 * errs[0]
