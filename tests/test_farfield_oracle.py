"""The far-field / cross-section oracle (oracle/farfield.py) against closed forms,
an independent evaluation, Lorenz-Mie, the optical theorem and energy conservation."""
import math

import mpmath as mp
import numpy as np
import pytest

from oracle import farfield as FF


def _P(n, m, x):
    """P_n^m without Condon-Shortley phase for m >= 0; P_n^{-m} = (-1)^m (n-m)!/(n+m)! P_n^m."""
    if m >= 0:
        return (-1) ** m * mp.legenp(n, m, x)
    return (-1) ** (-m) * mp.factorial(n + m) / mp.factorial(n - m) * _P(n, -m, x)


def test_pitau_closed_forms_alpha0():
    pi, tau = FF.pitau(8, 0.0)
    for n in range(1, 9):
        row_p, row_t = pi[n - 1], tau[n - 1]
        for m in range(-n, n + 1):
            ep = {1: n * (n + 1) / 2, -1: 0.5}.get(m, 0.0)
            et = {1: n * (n + 1) / 2, -1: -0.5}.get(m, 0.0)
            assert abs(row_p[m + n] - ep) < 1e-12 * max(1, abs(ep))
            assert abs(row_t[m + n] - et) < 1e-12 * max(1, abs(et))


def test_pitau_brute_force():
    mp.mp.dps = 30
    th = 0.7
    pi, tau = FF.pitau(6, th)
    for n in range(1, 7):
        for m in range(-n, n + 1):
            ep = m * _P(n, m, mp.cos(th)) / mp.sin(th)
            et = mp.diff(lambda t: _P(n, m, mp.cos(t)), th)
            assert abs(pi[n - 1][m + n] - float(ep)) < 1e-11 * max(1.0, abs(float(ep)))
            assert abs(tau[n - 1][m + n] - float(et)) < 1e-11 * max(1.0, abs(float(et)))


def _mie_s12(an, bn, th):
    """Lorenz-Mie S1, S2 by the textbook pi_n, tau_n recurrences (Bohren & Huffman 4.74)."""
    mu = math.cos(th)
    pis = [0.0, 1.0]
    S1 = S2 = 0.0
    for n in range(1, len(an) + 1):
        if n >= 2:
            pis.append((2 * n - 1) / (n - 1) * mu * pis[n - 1] - n / (n - 1) * pis[n - 2])
        tau = n * mu * pis[n] - (n + 1) * pis[n - 1]
        f = (2 * n + 1) / (n * (n + 1))
        S1 += f * (an[n - 1] * pis[n] + bn[n - 1] * tau)
        S2 += f * (an[n - 1] * tau + bn[n - 1] * pis[n])
    return S1, S2


def _single(oracle, N, beta):
    an, bn = oracle.mie_ab(3.0, 1.5 + 0.01j, N)
    pos = np.zeros((1, 3))
    p = FF.plane_wave(pos, N, 1.0, 0.0, beta)
    blk = N * (N + 2)
    nn = np.array([n for n in range(1, N + 1) for m in range(-n, n + 1)])
    c = p.copy()
    c[:blk] *= an[nn - 1]
    c[blk:] *= bn[nn - 1]
    return an, bn, pos, p, c


def test_single_sphere_is_lorenz_mie(oracle):
    N = 14
    an, bn, pos, p, cx = _single(oracle, N, 0.0)
    _, _, _, _, cy = _single(oracle, N, math.pi / 2)
    th = np.array([0.0, 0.3, 1.1, 2.5, math.pi])
    ph = np.array([0.0, 0.2, 2.0, -1.0, 0.4])
    Fx = FF.far_field(cx, pos, N, 1.0, th, ph)
    Fy = FF.far_field(cy, pos, N, 1.0, th, ph)
    S1, S2, S3, S4 = FF.amplitude_matrix(Fx, Fy, ph)
    for i, t in enumerate(th):
        e1, e2 = _mie_s12(an, bn, t)
        assert abs(S1[i] - e1) < 1e-12 * abs(e1)
        assert abs(S2[i] - e2) < 1e-12 * abs(e2)
        assert abs(S3[i]) < 1e-12 * abs(e1) and abs(S4[i]) < 1e-12 * abs(e1)
    Cs, Ce, Ca = FF.cross_sections(cx, p, [an], [bn], N, 1.0)
    n = np.arange(1, N + 1)
    assert abs(Ce - 2 * math.pi * np.sum((2 * n + 1) * np.real(an + bn))) < 1e-12 * Ce
    assert abs(Cs - 2 * math.pi * np.sum((2 * n + 1) * (np.abs(an) ** 2 + np.abs(bn) ** 2))) < 1e-12 * Cs
    # Mueller of a sphere: M12 = M21, M33 = M44, M34 = -M43, zero off-block
    M = FF.mueller(S1, S2, S3, S4)
    assert np.allclose(M[:, 0, 1], M[:, 1, 0]) and np.allclose(M[:, 2, 2], M[:, 3, 3])
    assert np.allclose(M[:, 2, 3], -M[:, 3, 2]) and np.allclose(M[:, 0, 2:], 0, atol=1e-12 * M[:, 0, 0].max())


def _dense_solve(oracle, pos, lmax, N, k, m_rel, radii, alpha=0.0, beta=0.0, dsign=-1.0):
    """Dense solve of (I + T A) x = T p.  The oracle's matvec uses d = origin(source) -
    origin(target) (the reference's direct path); it sees positions only through d, so
    dsign = -1 (d = origin(target) - origin(source)) evaluates it at -pos."""
    ns = len(pos)
    an = np.zeros((ns, N), complex)
    bn = np.zeros((ns, N), complex)
    for j in range(ns):
        a, b = oracle.mie_ab(k * radii[j], m_rel, N)
        an[j], bn[j] = a, b
    n = ns * 2 * N * (N + 2)
    M = np.zeros((n, n), complex)
    for i in range(n):
        e = np.zeros(n, complex)
        e[i] = 1.0
        M[:, i] = oracle.matvec_direct(dsign * pos, lmax, N, an, bn, k, e)
    p = FF.plane_wave(pos, N, k, alpha, beta)
    blk = N * (N + 2)
    nn = np.array([q for q in range(1, N + 1) for _ in range(-q, q + 1)])
    rhs = p.copy()
    for j in range(ns):
        rhs[j * 2 * blk:j * 2 * blk + blk] *= an[j][nn - 1]
        rhs[j * 2 * blk + blk:(j + 1) * 2 * blk] *= bn[j][nn - 1]
    return np.linalg.solve(M, rhs), p, an, bn


def _csca_far_field(x, pos, N, k, ng=40, nphi=48):
    """Far-field integral (1/k^2) int |F|^2 dOmega: Gauss-Legendre in cos(theta) x trapezoid in phi."""
    xg, wg = np.polynomial.legendre.leggauss(ng)
    th = np.arccos(xg)
    phis = 2 * math.pi * np.arange(nphi) / nphi
    T, P = np.meshgrid(th, phis, indexing="ij")
    Ft, Fp = FF.far_field(x, pos, N, k, T.ravel(), P.ravel())
    I = (np.abs(Ft) ** 2 + np.abs(Fp) ** 2).reshape(ng, nphi)
    return np.sum(wg[:, None] * I) * (2 * math.pi / nphi) / k ** 2


POS3 = np.array([[0.0, 0.0, 0.0], [1.9, 0.4, -0.3], [-0.5, 1.6, 1.2]])
RAD3 = np.array([0.7, 0.6, 0.5])


@pytest.mark.parametrize("m_rel", [1.5 + 0.0j, 1.4 + 0.05j])
def test_cluster_optical_theorem_and_energy(oracle, m_rel):
    """Three coupled spheres: C_ext from the coefficients equals the optical theorem
    4 pi/k^2 Im F_theta(0, 0); C_ext - C_abs equals the far-field integral of |F|^2."""
    pos, radii = POS3, RAD3
    N = 6
    k = 1.3
    x, p, an, bn = _dense_solve(oracle, pos, np.full(3, N, np.int32), N, k, m_rel, radii)
    Cs, Ce, Ca = FF.cross_sections(x, p, an, bn, N, k)
    Ft, Fp = FF.far_field(x, pos, N, k, [0.0], [0.0])
    assert abs(4 * math.pi / k ** 2 * Ft[0].imag - Ce) < 1e-10 * Ce
    csca_ff = _csca_far_field(x, pos, N, k)
    assert abs(csca_ff - Cs) < 1e-8 * Cs
    if m_rel.imag == 0:
        assert abs(Ca) < 1e-10 * Ce
    else:
        assert Ca > 0


def test_reference_sign_breaks_energy_balance(oracle):
    """The translation sign the reference's direct path uses (d = origin(source) -
    origin(target), lib_mie_ms_solver_interface.f90:941-943) gives a lossless cluster
    a far-field integral that differs from its extinction; the opposite sign (the
    default of the product) conserves energy (previous test)."""
    N, k = 6, 1.3
    x, p, an, bn = _dense_solve(oracle, POS3, np.full(3, N, np.int32), N, k, 1.5, RAD3, dsign=1.0)
    Cs, Ce, Ca = FF.cross_sections(x, p, an, bn, N, k)
    assert abs(_csca_far_field(x, POS3, N, k) - Cs) > 1e-2 * Cs
