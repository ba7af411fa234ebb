"""FP64 parity of the direct (sphere-to-sphere) translations at the orders of the
2000-sphere benchmark cluster: sphere orders up to 50 (radius 3 um at 550 nm) and
k|d| from near contact (t ~ 40) to t ~ 2300, polar angles down to 10 degrees, against
the oracle's Gaunt/3j matvec (lib_mie_ms_solver_interface.f90:806-998), per target
sphere at 1e-10.

Inputs are random in the normalised basis (x = x_hat / s_nm) and results are compared
there: in the flat basis one sphere's entries span 1e+-80 at these orders, and a flat
random vector makes the comparison hinge on couplings of relative size 1e-100."""
import math

import numpy as np
import pytest

from helpers import per_sphere_rel, snm_flat
from itofoggysim_amd import MultiSphereSolver, config_cluster

pytestmark = pytest.mark.gpu

LAM = 0.55
K = 2.0 * math.pi / LAM
M_WATER = 1.333 + 1e-9j


def _geometry():
    """Spheres 0 and 1 of radius 3.0 / 2.9 um (orders 50 and 48) at near contact; the
    others at separations spanning the near field and beyond, two of them within 10
    and 25 degrees of the polar axis through sphere 0."""
    r = np.array([3.0, 2.9, 2.5, 1.8, 1.0, 0.6, 0.3, 2.0, 1.5, 2.2])
    dist = np.array([0.0, 3.0 + 2.9 + 0.05, 10.0, 30.0, 60.0, 110.0, 3.0 + 0.3 + 0.05, 200.0, 14.0, 40.0])
    rng = np.random.default_rng(2024)
    pos = np.zeros((len(r), 3))
    for j in range(1, len(r)):
        v = rng.standard_normal(3)
        pos[j] = dist[j] * v / np.linalg.norm(v)
    for j, ang in ((8, 10.0), (9, 25.0)):
        a = math.radians(ang)
        pos[j] = dist[j] * np.array([math.sin(a) * math.cos(0.3 * j), math.sin(a) * math.sin(0.3 * j), math.cos(a)])
    for i in range(len(r)):
        for j in range(i):
            assert np.linalg.norm(pos[i] - pos[j]) >= r[i] + r[j] + 0.049
    lmax = np.ceil(K * r + 4.0 * np.cbrt(K * r) + 2.0).astype(np.int32)
    return pos, r, lmax


@pytest.mark.parametrize("cache", [-1, 0])
def test_direct_high_order_vs_oracle(oracle, cache):
    """cache -1: coaxial coefficients generated and applied on the fly;
    cache 0: tables generated once and kept."""
    pos, r, lmax = _geometry()
    ns = len(r)
    assert lmax.max() == 50 and lmax.min() < 15
    s = MultiSphereSolver(pos, r, M_WATER, K, lmax=lmax, cache_mb=cache)
    an, bn = s.mie_coefficients()
    ng = s.n_global
    blk = 2 * ng * (ng + 2)
    S, mask = snm_flat(ng, lmax)
    rng = np.random.default_rng(5)
    xh = np.zeros(s.n, dtype=np.complex128)
    # sources: the two order-50/48 spheres
    for j in (0, 1):
        sl = slice(j * blk, (j + 1) * blk)
        xh[sl] = (rng.standard_normal(blk) + 1j * rng.standard_normal(blk)) * mask[sl]
    x = xh / S
    b = s.calculate_vector_b(x) * S
    tgt = [j for src in (0, 1) for j in range(ns) if j != src]
    src = [src for src in (0, 1) for j in range(ns) if j != src]
    ref = (x + oracle.matvec_pairs(-pos, lmax, ng, an, bn, K, x, tgt, src)) * S
    errs = per_sphere_rel(b, ref, xh, ns, ng)
    t = [K * np.linalg.norm(pos[j] - pos[0]) for j in range(ns)]
    print("cache", cache, "per-target rel err", " ".join(f"{e:.1e}(t={tt:.0f})" for e, tt in zip(errs, t)))
    assert errs.max() < 1e-10, errs


def test_direct_config2_sample_vs_oracle(oracle):
    """Eight spheres of the 2000-sphere generator (the largest radii and their
    nearest neighbours in the drawn cluster), all pairs, against the oracle."""
    c = config_cluster(2)
    pos, r, lm = c["centres"], c["radii"], c["lmax"]
    big = np.argsort(r)[-3:]
    pick = list(big)
    for j in big:
        d = np.linalg.norm(pos - pos[j], axis=1)
        for q in np.argsort(d)[1:3]:
            if q not in pick:
                pick.append(int(q))
    pick = np.array(pick[:8])
    P, R, Lm = pos[pick], r[pick], lm[pick]
    s = MultiSphereSolver(P, R, c["m_rel"][pick], c["k"], lmax=Lm, cache_mb=-1)
    an, bn = s.mie_coefficients()
    S, mask = snm_flat(s.n_global, Lm)
    rng = np.random.default_rng(9)
    xh = (rng.standard_normal(s.n) + 1j * rng.standard_normal(s.n)) * mask
    x = xh / S
    b = s.calculate_vector_b(x) * S
    ref = oracle.matvec_direct(-P, Lm, s.n_global, an, bn, c["k"], x) * S
    errs = per_sphere_rel(b, ref, xh, len(pick), s.n_global)
    print("config-2 sample lmax", list(Lm), "per-target rel err", " ".join(f"{e:.1e}" for e in errs))
    assert errs.max() < 1e-10, errs
