"""GPU optics of solved clusters (plane-wave coefficients, far field, amplitude and
Mueller matrices, cross sections) against the oracle restatement (oracle/farfield.py)
and the physics it is pinned by (Lorenz-Mie, optical theorem, energy conservation)."""
import math

import numpy as np
import pytest

from itofoggysim_amd import MultiSphereSolver, config_cluster
from oracle import farfield as FF

pytestmark = pytest.mark.gpu


def _solver(c, **kw):
    return MultiSphereSolver(c["centres"], c["radii"], c["m_rel"], c["k"], lmax=c["lmax"], **kw)


def test_plane_wave_vs_oracle():
    c = config_cluster(1)
    s = _solver(c)
    for alpha, beta in [(0.0, 0.0), (0.0, math.pi / 2), (0.7, 2.1), (2.9, -0.4)]:
        p = s.plane_wave(alpha, beta)
        ref = FF.plane_wave(c["centres"], s.n_global, c["k"], alpha, beta)
        err = np.linalg.norm(p - ref) / np.linalg.norm(ref)
        print(f"plane wave alpha={alpha} beta={beta}: rel err {err:.2e}")
        assert err < 1e-12
    # the alpha = beta = 0 right-hand side agrees with the general one on the rows
    # n <= L_j (the general one sets the rows beyond a sphere's L_j to 0)
    r0 = s.get_vector_b()
    r1 = s.get_vector_b_dir(0.0, 0.0)
    N = s.n_global
    nn = np.array([n for n in range(1, N + 1) for _ in range(-n, n + 1)])
    mask = np.concatenate([np.tile(nn <= min(L, N), 2) for L in s.lmax])
    assert np.all(r1[~mask] == 0)
    assert np.linalg.norm((r0 - r1)[mask]) <= 1e-14 * np.linalg.norm(r0)
    s.close()


def test_single_sphere_lorenz_mie(oracle):
    k = 2 * math.pi / 0.55
    pos = np.zeros((1, 3))
    s = MultiSphereSolver(pos, [0.8], 1.333 + 1e-3j, k)
    N = s.n_global
    an, bn = s.mie_coefficients()
    th = np.linspace(0.0, math.pi, 7)
    ph = np.linspace(-1.0, 2.0, 7)
    F = []
    for beta in (0.0, math.pi / 2):
        x = s.get_vector_b_dir(0.0, beta)  # one sphere: x = T p solves the system
        F.append(s.far_field(x, th, ph))
        Fo = FF.far_field(x, pos, N, k, th, ph)
        for a, b in zip(F[-1], Fo):
            assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))
    S1, S2, S3, S4 = s.amplitude_matrix(ph, F[0], F[1])
    for i, t in enumerate(th):
        mu = math.cos(t)
        pis = [0.0, 1.0]
        e1 = e2 = 0.0
        for n in range(1, N + 1):
            if n >= 2:
                pis.append((2 * n - 1) / (n - 1) * mu * pis[n - 1] - n / (n - 1) * pis[n - 2])
            tau = n * mu * pis[n] - (n + 1) * pis[n - 1]
            f = (2 * n + 1) / (n * (n + 1))
            e1 += f * (an[0, n - 1] * pis[n] + bn[0, n - 1] * tau)
            e2 += f * (an[0, n - 1] * tau + bn[0, n - 1] * pis[n])
        assert abs(S1[i] - e1) <= 1e-11 * abs(e1) and abs(S2[i] - e2) <= 1e-11 * abs(e2)
        assert abs(S3[i]) <= 1e-12 * abs(e1) and abs(S4[i]) <= 1e-12 * abs(e1)
    x = s.get_vector_b_dir()
    Cs, Ce, Ca = s.cross_sections(x, s.plane_wave())
    n = np.arange(1, N + 1)
    ce = 2 * math.pi / k ** 2 * np.sum((2 * n + 1) * np.real(an[0] + bn[0]))
    cs = 2 * math.pi / k ** 2 * np.sum((2 * n + 1) * (np.abs(an[0]) ** 2 + np.abs(bn[0]) ** 2))
    assert abs(Ce - ce) <= 1e-12 * ce and abs(Cs - cs) <= 1e-10 * cs
    s.close()


def test_config1_far_field_and_cross_sections(oracle):
    """20-sphere cluster: solve for x- and y-polarised waves, far field on a grid vs the
    oracle, optical theorem, energy conservation, Mueller matrix on the 181 x 36 grid."""
    c = config_cluster(1)
    s = _solver(c)
    N = s.n_global
    k = c["k"]
    an, bn = s.mie_coefficients()
    sols = []
    for beta in (0.0, math.pi / 2):
        x, info = s.solve(s.get_vector_b_dir(0.0, beta), restart=50, max_iter=600, tol=1e-11)
        assert info["converged"]
        sols.append(x)
    p = s.plane_wave()
    Cs, Ce, Ca = s.cross_sections(sols[0], p)
    So, Eo, Ao = FF.cross_sections(sols[0], p, an, bn, N, k)
    assert abs(Ce - Eo) <= 1e-12 * Eo and abs(Cs - So) <= 1e-10 * So
    # far field parity on a coarse grid
    rng = np.random.default_rng(3)
    th = rng.uniform(0, math.pi, 9)
    ph = rng.uniform(0, 2 * math.pi, 9)
    Ft, Fp = s.far_field(sols[0], th, ph)
    Ot, Op = FF.far_field(sols[0], c["centres"], N, k, th, ph)
    assert np.max(np.abs(Ft - Ot)) <= 1e-11 * np.max(np.abs(Ot))
    assert np.max(np.abs(Fp - Op)) <= 1e-11 * np.max(np.abs(Op))
    # optical theorem (an identity of the formulas) and energy conservation (physics of the solve)
    F0t, _ = s.far_field(sols[0], [0.0], [0.0])
    assert abs(4 * math.pi / k ** 2 * F0t[0].imag - Ce) <= 1e-10 * Ce
    # |F|^2 is band-limited by about 2 (k R + L) ~ 250 (cluster radius R ~ 9 um): 200
    # Gauss-Legendre nodes in cos(theta) and 512 in phi integrate it to rounding
    xg, wg = np.polynomial.legendre.leggauss(200)
    nphi = 512
    T, P = np.meshgrid(np.arccos(xg), 2 * math.pi * np.arange(nphi) / nphi, indexing="ij")
    Ft, Fp = s.far_field(sols[0], T.ravel(), P.ravel())
    I = (np.abs(Ft) ** 2 + np.abs(Fp) ** 2).reshape(T.shape)
    cs_ff = np.sum(wg[:, None] * I) * (2 * math.pi / nphi) / k ** 2
    print(f"config1 C_ext {Ce:.6e} C_sca {Cs:.6e} far-field integral {cs_ff:.6e} C_abs {Ca:.3e}")
    assert abs(cs_ff - Cs) <= 1e-6 * Cs
    # amplitude and Mueller matrices on the 181 x 36 grid
    th = np.repeat(np.linspace(0, math.pi, 181), 36)
    ph = np.tile(np.arange(36) * 2 * math.pi / 36, 181)
    Fx = s.far_field(sols[0], th, ph)
    Fy = s.far_field(sols[1], th, ph)
    S = s.amplitude_matrix(ph, Fx, Fy)
    M = s.mueller(*S)
    Mo = FF.mueller(*FF.amplitude_matrix(Fx, Fy, ph))
    assert np.max(np.abs(M - Mo)) <= 1e-12 * np.max(np.abs(Mo))
    assert np.all(M[:, 0, 0] > 0)
    s.close()
