"""The serving loop's admission plan (ifs_server_plan, host only): the device memory a
request is admitted with and the limits its context gets.

  * MLFMM pools equal the multipole and local pools of the octree the library's own
    tree functions build (non-empty boxes per level, 2 L(L+2) complex per box and
    direction, L = kD + 1.8 d0^(2/3) (kD)^(1/3));
  * the Krylov vectors of GMRES(restart) over ndir right-hand sides;
  * the workspace is a quarter of a slot's share, at least 256 MiB and at most 8 GiB
    (plus 64 MiB), and the context is created with exactly that (workspace_mb);
  * translation tables are kept only when they fit the rest of the share
    (cache_mb -1 otherwise); the footprint grows with the directions."""
import ctypes
import math

import numpy as np

from itofoggysim_amd import make_cluster, plan_request
from itofoggysim_amd._lib import lib


def _tree_pools(c, leaf, d0, ndir):
    L = lib()
    pts = np.ascontiguousarray(c["centres"])
    n = len(pts)
    sc = np.zeros_like(pts)
    xmd = np.zeros(4)
    L.ifs_tree_scale(pts.ctypes.data, n, sc.ctypes.data, xmd.ctypes.data)
    u = np.zeros(n, np.int64)
    L.ifs_tree_universal_index(sc.ctypes.data, n, 21, u.ctypes.data)
    u.sort()
    lm = L.ifs_tree_level_max(u.ctypes.data, ctypes.c_int64(n), ctypes.c_int64(leaf))
    total = 0.0
    for lev in range(2, max(2, lm) + 1):
        boxes = np.unique(u >> (3 * (21 - lev)))
        kD = c["k"] * xmd[3] / 2 ** lev
        Lb = math.ceil(kD + 1.8 * d0 ** (2.0 / 3.0) * kD ** (1.0 / 3.0))
        total += 2.0 * len(boxes) * 2.0 * Lb * (Lb + 2.0) * 16.0 * ndir
    return total


def test_plan_mlfmm_pools_vectors_workspace():
    c = make_cluster(400, ("uniform", 0.3, 1.0), side=10.0 * (400 / 20.0) ** (1.0 / 3.0), seed=77)
    for ndir in (1, 3):
        p = plan_request(c["centres"], c["radii"], c["k"], ndir, lmax=c["lmax"], restart=30, use_fmm=True,
                         hbm_budget_gb=200.0, slots=4)
        pools = _tree_pools(c, 32, 6.0, ndir)
        assert abs(p["pools"] - pools) <= 1e-9 * pools + 16, (p["pools"], pools)
        N = int(c["lmax"].max())
        n = 2 * len(c["radii"]) * N * (N + 2)
        assert p["vectors"] == 16 * n * (31 * ndir + 6 * ndir + 4)
        share = 200 * 2 ** 30 / 4
        cap = min(max(share / 4, 256 * 2 ** 20), 8 * 2 ** 30)
        assert p["workspace"] <= cap and p["workspace_mb"] == int(p["workspace"]) >> 20
        assert p["total"] == p["vectors"] + p["pools"] + p["workspace"] + p["tables"] + p["fixed"]


def test_plan_tables_kept_only_when_they_fit():
    c = make_cluster(60, ("uniform", 0.3, 1.0), side=10.0 * (60 / 20.0) ** (1.0 / 3.0), seed=78)
    big = plan_request(c["centres"], c["radii"], c["k"], 2, lmax=c["lmax"], use_fmm=False, hbm_budget_gb=200.0,
                       slots=4)
    small = plan_request(c["centres"], c["radii"], c["k"], 2, lmax=c["lmax"], use_fmm=False, hbm_budget_gb=1.0,
                         slots=8)
    assert big["tables"] > 0 and big["cache_mb"] > 0 and big["cache_mb"] * 2 ** 20 == big["tables"]
    assert small["tables"] == 0 and small["cache_mb"] == -1
    more = plan_request(c["centres"], c["radii"], c["k"], 6, lmax=c["lmax"], use_fmm=False, hbm_budget_gb=200.0,
                        slots=4)
    assert more["vectors"] > big["vectors"] and more["total"] > big["total"]
