"""Batched right-hand sides: one matvec pass for several vectors equals the single
matvecs (direct and MLFMM, including a multi-chunk workspace), and the lock-step
GMRES returns the solutions of the single solves."""
import math

import numpy as np
import pytest

from itofoggysim_amd import MultiSphereSolver, config_cluster

pytestmark = pytest.mark.gpu


def _rand(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("fmm,ws", [(False, 0), (True, 0), (True, 1)])
def test_batched_matvec_equals_single(fmm, ws):
    c = config_cluster(1)
    kw = dict(lmax=c["lmax"], use_fmm=fmm, leaf_size=2, workspace_mb=ws)
    s = MultiSphereSolver(c["centres"], c["radii"], c["m_rel"], c["k"], **kw)
    X = _rand((3, s.n), 5)
    B = s.calculate_vector_b_batch(X)
    for r in range(3):
        b = s.calculate_vector_b(X[r])
        err = np.linalg.norm(B[r] - b) / np.linalg.norm(b)
        print(f"fmm={fmm} ws={ws} rhs {r}: rel diff {err:.2e}")
        assert err < 1e-13
    s.close()


def test_batched_solve_equals_single():
    c = config_cluster(1)
    s = MultiSphereSolver(c["centres"], c["radii"], c["m_rel"], c["k"], lmax=c["lmax"])
    dirs = [(0.0, 0.0), (0.0, math.pi / 2), (1.1, 0.3), (2.5, 4.0)]
    R = np.stack([s.get_vector_b_dir(a, b) for a, b in dirs])
    X, infos = s.solve_batch(R, restart=30, max_iter=400, tol=1e-9)
    for r, (a, b) in enumerate(dirs):
        assert infos[r]["converged"], infos[r]
        x, inf = s.solve(R[r], restart=30, max_iter=400, tol=1e-9)
        err = np.linalg.norm(X[r] - x) / np.linalg.norm(x)
        print(f"dir {a:.2f},{b:.2f}: batched {infos[r]['iterations']} it, single {inf['iterations']} it, diff {err:.1e}")
        assert err < 1e-7
    s.close()


@pytest.mark.parametrize("fmm,precision", [(False, "fp64"), (True, "fp64"), (True, "mixed")])
def test_block_gmres_equals_single(fmm, precision):
    """Block GMRES (one Krylov space for all right-hand sides: 3 directions x TM/TE)
    returns the single solves' solutions (normalised basis), with FP64 true residuals
    below the tolerance, in no more products per right-hand side than the slowest
    single solve."""
    from helpers import snm_flat
    if fmm:
        c = config_cluster(2, n=200)
        s = MultiSphereSolver(c["centres"], c["radii"], c["m_rel"], c["k"], lmax=c["lmax"], use_fmm=True,
                              leaf_size=16, fmm_d0=6.0)
        tol = 1e-8
    else:
        c = config_cluster(1)
        s = MultiSphereSolver(c["centres"], c["radii"], c["m_rel"], c["k"], lmax=c["lmax"])
        tol = 1e-10
    try:
        S, _ = snm_flat(s.n_global, c["lmax"])
        dirs = [(0.0, 0.0), (1.1, 0.3), (2.5, 4.0)]
        R = np.stack([s.get_vector_b_dir(a, b, te=te) for a, b in dirs for te in (False, True)])
        X, infos = s.solve_block(R, restart=20, max_iter=600, tol=tol, precision=precision)
        worst_single = 0
        for r in range(len(R)):
            assert infos[r]["converged"] and infos[r]["rel_residual"] < tol, infos[r]
            x, inf = s.solve(R[r], restart=20, max_iter=600, tol=tol, precision=precision)
            worst_single = max(worst_single, inf["iterations"])
            err = np.linalg.norm((X[r] - x) * S) / np.linalg.norm(x * S)
            print(f"fmm={fmm} {precision} rhs {r}: block {infos[r]['iterations']} products, single "
                  f"{inf['iterations']}, difference {err:.1e}")
            assert err < 100 * tol
        assert infos[0]["iterations"] <= worst_single
    finally:
        s.close()
