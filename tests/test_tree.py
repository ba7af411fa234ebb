"""Octree: the product's C++ octree (host-only C ABI, no device work) against the
reference's libtree fixtures and against the oracle's independent restatement."""
import ctypes
import json
import os

import numpy as np
import pytest

from oracle import fmm_ref

REFV = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))["tree"]


@pytest.fixture(scope="module")
def L():
    from itofoggysim_amd._lib import lib
    return lib()


def _uindex(L, pts, level):
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    out = np.empty(len(pts), dtype=np.int64)
    assert L.ifs_tree_universal_index(pts.ctypes.data, len(pts), level, out.ctypes.data) == 0
    return out


def _neigh(L, k, n, level):
    out = np.empty(26, dtype=np.int64)
    assert L.ifs_tree_neighbours(k, n, level, out.ctypes.data) == 0
    return out


def _i4(L, n, level):
    out = np.empty(400, dtype=np.int64)
    cnt = L.ifs_tree_domain_i4(n, level, out.ctypes.data, 400)
    return out[:cnt]


def test_reference_fixtures(L):
    f = REFV["uindex"]
    assert _uindex(L, f["point"], f["level"])[0] == f["n"]
    f = REFV["uindex_lmax"]
    assert _uindex(L, f["point"], f["level"])[0] == f["n"]
    assert list(_neigh(L, 1, 35, 2)) == REFV["neighbours_35"]["out"]
    assert list(_neigh(L, 1, 5, 1)) == REFV["neighbours_5"]["out"]
    c = np.empty(3)
    assert L.ifs_tree_centre(1, 1, c.ctypes.data) == 0
    assert list(c) == REFV["centre"]["centre"]
    nb = _neigh(L, 1, 0, 2)
    assert 0 + nb[nb >= 0].sum() == REFV["i2_sum"]["sum"]
    # level max for 5 diagonal points (lib_tree.f90:2595-2681); the points are
    # single-precision expressions in the reference test
    pts = np.array(REFV["level_max"]["points"], dtype=np.float32).astype(np.float64)
    u = _uindex(L, pts, 21)
    order = np.argsort(u, kind="stable")
    assert list(order + 1) == REFV["level_max"]["sorted"]
    us = np.ascontiguousarray(np.sort(u))
    for s, lm in zip(REFV["level_max"]["s"], REFV["level_max"]["l_max"]):
        assert L.ifs_tree_level_max(us.ctypes.data, len(us), s) == lm


def test_scaling_matches_restatement(L):
    rng = np.random.default_rng(3)
    pts = rng.uniform(-7.0, 13.0, (500, 3))
    sc = np.empty_like(pts)
    md = np.empty(4)
    assert L.ifs_tree_scale(pts.ctypes.data, len(pts), sc.ctypes.data, md.ctypes.data) == 0
    ref, lo, D = fmm_ref.scale_points(pts)
    assert np.array_equal(sc, ref)
    assert md[3] == D and np.array_equal(md[:3], lo)
    assert sc.max() < 1.0 and sc.min() >= 0.0


@pytest.mark.parametrize("level", [1, 2, 3, 5, 8, 13, 21])
def test_morton_bit_exact(L, level):
    rng = np.random.default_rng(level)
    pts = rng.uniform(0.0, 1.0, (2000, 3))
    pts[:5] = [[0.0, 0.0, 0.0], [0.5, 0.5, 0.5], [1 - 2 ** -53, 0.25, 0.125], [2 ** -30, 0.75, 0.0],
               [0.999999, 0.000001, 0.5]]
    u = _uindex(L, pts, level)
    ref = np.array([fmm_ref.morton(p, level) for p in pts])
    assert np.array_equal(u, ref)


@pytest.mark.parametrize("level", [2, 3, 4])
def test_interaction_lists(L, level):
    rng = np.random.default_rng(10 + level)
    for n in rng.integers(0, 8 ** level, 60):
        n = int(n)
        assert list(_i4(L, n, level)) == fmm_ref.i4(n, level)
        nb = _neigh(L, 1, n, level)
        assert sorted(nb[nb >= 0]) == sorted(fmm_ref.adjacent(n, level))
    # interior box at level 3 has 189 partners
    assert len(_i4(L, fmm_ref.from_coords([3, 4, 3], 3), 3)) == 189


def test_level_max_random(L):
    rng = np.random.default_rng(8)
    for s in (1, 4, 10, 32):
        pts = rng.uniform(0.0, 1.0, (300, 3))
        u = np.sort(_uindex(L, pts, 21))
        assert L.ifs_tree_level_max(u.ctypes.data, len(u), s) == fmm_ref.level_max(pts, s)
