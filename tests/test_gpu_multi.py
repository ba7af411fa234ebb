"""Many independent clusters in one context (ifs_create_multi / ifs_solve_multi), the
HIP-graph matvec (ifs_matvec_graph_device, IFS_SOLVE_GRAPH) and the TE plane wave.

  * the block-diagonal matvec equals each cluster's own matvec;
  * the lock-step per-cluster GMRES equals each cluster solved alone, and the smallest
    cluster's cross sections equal the oracle's dense solve
    (lib_mie_ms_solver_interface.f90:806-998, lib_mie_ss_helper_functions.f90:1380-1475);
  * graph replays equal the eager matvec bit for bit, replay after replay with eager
    products of other vectors in between, on a direct and on an MLFMM context (FP64
    and mixed); a graph-mode solve equals the eager one;
  * TE mode = TM coefficients with p and q exchanged (lib_mie_illumination.f90:347-352)."""
import math

import numpy as np
import pytest

from helpers import Hip, snm_flat
from itofoggysim_amd import ClusterBatch, MultiSphereSolver, config_cluster, make_cluster

pytestmark = pytest.mark.gpu


def _clusters(sizes, seed0=4000):
    return [make_cluster(n, ("uniform", 0.3, 1.0), side=10.0 * (n / 20.0) ** (1.0 / 3.0), seed=seed0 + i)
            for i, n in enumerate(sizes)]


def _single(c, N, **kw):
    return MultiSphereSolver(c["centres"], c["radii"], c["m_rel"], c["k"], lmax=c["lmax"], n_global=N, **kw)


def test_multi_matvec_equals_single():
    cls = _clusters([3, 7, 12, 25, 5, 18])
    mb = ClusterBatch(cls, cls[0]["k"])
    N = mb.n_global
    S, mask = snm_flat(N, mb.lmax)
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(mb.n) + 1j * rng.standard_normal(mb.n)) * mask / S
    b = mb.calculate_vector_b(x) * S
    assert mb.n_clusters == len(cls)
    worst = 0.0
    for g, c in enumerate(cls):
        s = _single(c, N)
        bs = s.calculate_vector_b(mb.segment(x, g)) * mb.segment(S, g)
        s.close()
        e = np.linalg.norm(mb.segment(b, g) - bs) / np.linalg.norm(bs)
        worst = max(worst, e)
    mb.close()
    print(f"multi-cluster matvec vs per-cluster matvecs: {worst:.1e}")
    assert worst < 1e-13


def _oracle_csec(c, oracle, N, alpha, beta):
    """Cross sections of the dense oracle solve of one cluster."""
    from oracle import farfield as FF
    pos, rad, k, lm = c["centres"], c["radii"], c["k"], c["lmax"]
    ns = len(rad)
    blk = N * (N + 2)
    an = np.zeros((ns, N), complex)
    bn = np.zeros_like(an)
    for j in range(ns):
        an[j], bn[j] = oracle.mie_ab(k * rad[j], c["m_rel"][j], N)
    nn = np.array([q for q in range(1, N + 1) for _ in range(-q, q + 1)])
    M = np.eye(2 * ns * blk, dtype=complex)
    for j in range(ns):
        rt, nt = j * 2 * blk, lm[j] * (lm[j] + 2)
        ta, tb = an[j][nn[:nt] - 1], bn[j][nn[:nt] - 1]
        for l in range(ns):
            if l == j:
                continue
            d = pos[j] - pos[l]
            dd = np.linalg.norm(d)
            A, B = oracle.translation_ab(k * dd, np.arccos(d[2] / dd), np.arctan2(d[1], d[0]), lm[j], lm[l], 3)
            cs, nc = l * 2 * blk, lm[l] * (lm[l] + 2)
            M[rt:rt + nt, cs:cs + nc] = ta[:, None] * A
            M[rt:rt + nt, cs + blk:cs + blk + nc] = ta[:, None] * B
            M[rt + blk:rt + blk + nt, cs:cs + nc] = tb[:, None] * B
            M[rt + blk:rt + blk + nt, cs + blk:cs + blk + nc] = tb[:, None] * A
    p = FF.plane_wave(pos, N, k, alpha, beta)
    rhs = p.copy()
    for j in range(ns):
        keep = nn <= lm[j]
        rhs[2 * j * blk:2 * j * blk + blk] *= np.where(keep, an[j][nn - 1], 0)
        rhs[2 * j * blk + blk:2 * (j + 1) * blk] *= np.where(keep, bn[j][nn - 1], 0)
    # solved in the normalised basis (flat entries span 1e+-30 at these orders)
    S, _ = snm_flat(N, lm)
    xh = np.linalg.solve(S[:, None] * M / S[None, :], S * rhs)
    return np.array(FF.cross_sections(xh / S, p, an, bn, N, k))


def test_multi_solve_equals_single_and_oracle(oracle):
    cls = _clusters([3, 9, 16, 6], seed0=4100)
    mb = ClusterBatch(cls, cls[0]["k"])
    N = mb.n_global
    a, bt = 0.4, 1.1
    rhs = mb.get_vector_b_dir(a, bt)
    p = mb.plane_wave(a, bt)
    x, infos = mb.solve(rhs, restart=30, tol=1e-10, max_iter=500)
    xg, infos_g = mb.solve(rhs, restart=30, tol=1e-10, max_iter=500, graph=True)
    print("multi", infos)
    assert all(i["converged"] for i in infos)
    assert np.array_equal(x, xg) and [i["iterations"] for i in infos] == [i["iterations"] for i in infos_g]
    S, _ = snm_flat(N, mb.lmax)
    for g, c in enumerate(cls):
        s = _single(c, N)
        xs, inf = s.solve(s.get_vector_b_dir(a, bt), restart=30, tol=1e-10, max_iter=500)
        s.close()
        Sg = mb.segment(S, g)
        e = np.linalg.norm((mb.segment(x, g) - xs) * Sg) / np.linalg.norm(xs * Sg)
        print(f"cluster {g} ({len(c['radii'])} spheres): multi vs alone {e:.1e}, "
              f"iterations {infos[g]['iterations']} / {inf['iterations']}")
        assert abs(infos[g]["iterations"] - inf["iterations"]) <= 1 and e < 1e-9
    g = 0  # the 3-sphere cluster against the oracle's dense solve
    cs = np.array(mb.cross_sections_cluster(x, p, g))
    ref = _oracle_csec(cls[g], oracle, N, a, bt)
    err = np.max(np.abs(cs - ref)) / abs(ref[1])
    print(f"cluster 0 cross sections vs oracle dense solve: {err:.1e} (C_ext {ref[1]:.6g})")
    mb.close()
    assert err < 1e-8


def _graph_replays(s, precision):
    hip = Hip()
    rng = np.random.default_rng(7)
    S, mask = snm_flat(s.n_global, s.lmax)
    x = (rng.standard_normal(s.n) + 1j * rng.standard_normal(s.n)) * mask / S
    y = (rng.standard_normal(s.n) + 1j * rng.standard_normal(s.n)) * mask / S
    s.set_precision(precision)
    dx, dy = hip.upload(x), hip.upload(y)
    db, dg, dz = hip.alloc(16 * s.n), hip.alloc(16 * s.n), hip.alloc(16 * s.n)
    s.calculate_vector_b_device(dx.value, db.value)
    hip.sync(s.stream())
    ref = hip.download(db, x)
    refy = None
    for rep in range(6):  # 1: eager, 2: capture and launch, 3-6: replays
        s.calculate_vector_b_graph_device(dx.value, dg.value)
        s.calculate_vector_b_device(dy.value, dz.value)  # eager work on the same buffers in between
        hip.sync(s.stream())
        got = hip.download(dg, x)
        gy = hip.download(dz, y)
        refy = gy if refy is None else refy
        assert np.array_equal(got, ref), f"graph call {rep + 1} differs from the eager matvec"
        assert np.array_equal(gy, refy)
    hip.free_all()
    return ref


@pytest.mark.parametrize("fmm", [False, True])
def test_graph_matvec_replays(fmm):
    if fmm:
        c = config_cluster(2, n=150)
        s = MultiSphereSolver(c["centres"], c["radii"], c["m_rel"], c["k"], lmax=c["lmax"], use_fmm=True,
                              leaf_size=16, fmm_d0=6.0)
    else:
        c = config_cluster(1)
        s = MultiSphereSolver(c["centres"], c["radii"], c["m_rel"], c["k"], lmax=c["lmax"])
    try:
        for prec in (("fp64", "mixed") if fmm else ("fp64",)):
            _graph_replays(s, prec)
        s.set_precision("fp64")
        rhs = s.get_vector_b()
        x1, i1 = s.solve(rhs, restart=30, tol=1e-8, x0=rhs)
        x2, i2 = s.solve(rhs, restart=30, tol=1e-8, x0=rhs, graph=True)
        print("eager", i1, "\ngraph", i2)
        assert i1["converged"] and i1["iterations"] == i2["iterations"] and np.array_equal(x1, x2)
    finally:
        s.close()


def test_te_mode():
    c = config_cluster(1)
    s = MultiSphereSolver(c["centres"], c["radii"], c["m_rel"], c["k"], lmax=c["lmax"])
    try:
        blk = s.n_global * (s.n_global + 2)
        tm = s.plane_wave_mode(0.7, 2.0, te=False)
        te = s.plane_wave_mode(0.7, 2.0, te=True)
        assert np.array_equal(tm, s.plane_wave(0.7, 2.0))
        for j in range(s.n_spheres):
            o = 2 * j * blk
            assert np.array_equal(te[o:o + blk], tm[o + blk:o + 2 * blk])
            assert np.array_equal(te[o + blk:o + 2 * blk], tm[o:o + blk])
        an, bn = s.mie_coefficients()
        rte = s.get_vector_b_dir(0.7, 2.0, te=True)
        nn = np.array([q for q in range(1, s.n_global + 1) for _ in range(-q, q + 1)])
        for j in range(s.n_spheres):
            o = 2 * j * blk
            keep = nn <= s.lmax[j]
            assert np.allclose(rte[o:o + blk], np.where(keep, an[j][nn - 1], 0) * te[o:o + blk], rtol=1e-14, atol=0)
    finally:
        s.close()
