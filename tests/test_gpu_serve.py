"""Serving loop: requests of different sizes and wavelengths admitted while others
run give the same cross sections as each request run alone, the admitted requests
overlap, and the loop reports p50/p99 latency."""
import math

import numpy as np
import pytest

from itofoggysim_amd import MultiSphereSolver, Server, make_cluster

pytestmark = pytest.mark.gpu


def _requests():
    rng = np.random.default_rng(55)
    out = []
    for i in range(6):
        ns = int(rng.integers(3, 12))
        lam = float(rng.uniform(0.4, 1.55))
        nw = 1.3199 + 3500e-6 / lam ** 2 + 1e-8j  # water, lambda in um
        k = 2 * math.pi / lam
        rad = rng.uniform(0.1, 0.3, ns)
        pos = rng.uniform(0, 4.0, (ns, 3))
        ok = False
        while not ok:  # redraw overlapping spheres
            ok = True
            for a in range(ns):
                for b in range(a):
                    if np.linalg.norm(pos[a] - pos[b]) < rad[a] + rad[b] + 0.05:
                        pos[a] = rng.uniform(0, 4.0, 3)
                        ok = False
        nd = int(rng.integers(1, 5))
        dirs = [(float(rng.uniform(0, math.pi)), float(rng.uniform(0, 2 * math.pi))) for _ in range(nd)]
        out.append(dict(centres=pos, radii=rad, m_rel=nw, k=k, directions=dirs))
    return out


def _alone(r):
    s = MultiSphereSolver(r["centres"], r["radii"], r["m_rel"], r["k"])
    res = []
    for a, b in r["directions"]:
        x, info = s.solve(s.get_vector_b_dir(a, b), restart=30, max_iter=1000, tol=1e-10)
        res.append(s.cross_sections(x, s.plane_wave(a, b)))
    s.close()
    return np.array(res)


def test_serving_same_alone_and_in_the_mix():
    reqs = _requests()
    srv = Server(slots=3, hbm_budget_gb=32.0)
    ids = [srv.submit(r["centres"], r["radii"], r["m_rel"], r["k"], r["directions"], tol=1e-10) for r in reqs]
    got = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    print("serving stats", st)
    assert st["completed"] == len(reqs) and st["p99_s"] >= st["p50_s"] > 0
    assert st["peak_concurrent"] >= 2  # requests overlapped on the device
    for r, (cs, info) in zip(reqs, got):
        assert info["status"] == 0
        ref = _alone(r)
        err = np.max(np.abs(cs - ref) / np.abs(ref))
        print(f"ns={len(r['radii'])} ndir={len(r['directions'])} latency {info['latency_s']:.3f} s, "
              f"rel diff vs alone {err:.1e}")
        assert err < 1e-8
