"""GPU parity of the direct (all-pairs) matvec, the Mie T-matrix and the solve,
against the CPU oracle on the same seeded inputs."""
import math

import numpy as np
import pytest

from itofoggysim_amd import MultiSphereSolver, config_cluster

pytestmark = pytest.mark.gpu

RTOL = 1e-10  # FP64 mode: relative 2-norm error vs the oracle

# The oracle restates the reference's direct path, d = origin(source) - origin(target).
# The default (physical) convention is d = origin(target) - origin(source); the oracle
# evaluates it at the mirrored positions -pos (it sees positions only through d).


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def _small_cluster(ns, lmax_hi, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.0, 6.0, (ns, 3))
    radii = np.full(ns, 0.3)
    lmax = rng.integers(1, lmax_hi + 1, ns).astype(np.int32)
    return pos, radii, lmax


@pytest.mark.parametrize("cache", [0, -1])
@pytest.mark.parametrize("conv", ["physical", "reference"])
@pytest.mark.parametrize("zsel", [3, 1])
@pytest.mark.parametrize("ns,lhi,seed", [(2, 3, 0), (5, 6, 1), (7, 9, 2)])
def test_direct_matvec_small(oracle, conv, zsel, ns, lhi, seed, cache):
    """cache 0: tables generated once and kept; cache -1: coaxial coefficients generated
    and applied on the fly in every matvec (per-pair orders Ls > Lt and Ls <= Lt both occur)."""
    pos, radii, lmax = _small_cluster(ns, lhi, seed)
    k = 1.3
    s = MultiSphereSolver(pos, radii, 1.5 + 0.01j, k, lmax=lmax, z_selector=zsel, d_convention=conv, cache_mb=cache)
    an, bn = s.mie_coefficients()
    x = _rand(s.n, seed + 10)
    b = s.calculate_vector_b(x)
    opos = -pos if conv == "physical" else pos
    ref = oracle.matvec_direct(opos, lmax, s.n_global, an, bn, k, x, zsel=zsel)
    err = np.linalg.norm(b - ref) / np.linalg.norm(ref)
    print(f"ns={ns} lhi={lhi} zsel={zsel} {conv} rel err {err:.3e}")
    assert err < RTOL
    s.close()


def test_mie_vs_oracle(oracle):
    c = config_cluster(1)
    s = MultiSphereSolver(c["centres"], c["radii"], c["m_rel"], c["k"], lmax=c["lmax"])
    an, bn = s.mie_coefficients()
    for j in range(len(c["radii"])):
        ao, bo = oracle.mie_ab(c["k"] * c["radii"][j], c["m_rel"][j], s.n_global)
        assert np.max(np.abs(an[j] - ao)) < 1e-12
        assert np.max(np.abs(bn[j] - bo)) < 1e-12
    s.close()


@pytest.mark.parametrize("cache", [0, -1])
def test_direct_matvec_config1(oracle, cache):
    """20 water spheres, U[0.3,1] um in a 10 um cube, lambda 550 nm (L_j 14..23)."""
    c = config_cluster(1)
    s = MultiSphereSolver(c["centres"], c["radii"], c["m_rel"], c["k"], lmax=c["lmax"], cache_mb=cache)
    an, bn = s.mie_coefficients()
    x = _rand(s.n, 7)
    b = s.calculate_vector_b(x)
    ref = oracle.matvec_direct(-c["centres"], c["lmax"], s.n_global, an, bn, c["k"], x, zsel=3)
    err = np.linalg.norm(b - ref) / np.linalg.norm(ref)
    print(f"config1 rel err {err:.3e}")
    assert err < RTOL
    s.close()


def test_ref_row_zeroing(oracle):
    pos, radii, lmax = _small_cluster(5, 5, 3)
    k = 1.1
    s = MultiSphereSolver(pos, radii, 1.4, k, lmax=lmax, ref_row_zeroing=True, d_convention="reference")
    an, bn = s.mie_coefficients()
    x = _rand(s.n, 4)
    b = s.calculate_vector_b(x)
    ref = oracle.matvec_direct(pos, lmax, s.n_global, an, bn, k, x, zsel=3, ref_row_zeroing=True)
    assert np.linalg.norm(b - ref) / np.linalg.norm(ref) < RTOL
    s.close()


def test_solve_config1(oracle):
    """GMRES(50) to 1e-8 on config 1; the converged x satisfies the oracle's operator."""
    c = config_cluster(1)
    s = MultiSphereSolver(c["centres"], c["radii"], c["m_rel"], c["k"], lmax=c["lmax"])
    rhs = s.get_vector_b()
    x, info = s.solve(rhs, restart=50, max_iter=500, tol=1e-8)
    print(info)
    assert info["converged"]
    an, bn = s.mie_coefficients()
    Mx = oracle.matvec_direct(-c["centres"], c["lmax"], s.n_global, an, bn, c["k"], x, zsel=3)
    # residual in the normalised basis s(n,m) = sqrt(n(n+1)(2n+1)(n-m)!/(n+m)!)
    N = s.n_global
    snm = np.array([np.sqrt(n * (n + 1) * (2 * n + 1) * math.factorial(n - m) / math.factorial(n + m))
                    for n in range(1, N + 1) for m in range(-n, n + 1)])
    w = np.tile(np.concatenate([snm, snm]), s.n_spheres)
    res = np.linalg.norm(w * (Mx - rhs)) / np.linalg.norm(w * rhs)
    print("oracle residual (normalised basis)", res)
    assert res < 1e-7
    s.close()
