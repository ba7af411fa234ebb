"""GPU octree build (ifs_tree_build_gpu, ifs_tree_i4_gpu) bit-exact against the host
octree (ifs_tree_*), which the reference's libtree fixtures pin (test_tree.py)."""
import numpy as np
import pytest

from itofoggysim_amd import config_cluster

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    from itofoggysim_amd._lib import lib
    return lib()


def _host_tree(L, pts, s):
    n = len(pts)
    scaled = np.empty_like(pts)
    xd = np.empty(4)
    assert L.ifs_tree_scale(pts.ctypes.data, n, scaled.ctypes.data, xd.ctypes.data) == 0
    u = np.empty(n, dtype=np.int64)
    assert L.ifs_tree_universal_index(scaled.ctypes.data, n, 21, u.ctypes.data) == 0
    us = np.sort(u)
    lmax = max(L.ifs_tree_level_max(us.ctypes.data, n, s), 2)
    return u, us, lmax


def _gpu_tree(L, pts, s):
    n = len(pts)
    us = np.empty(n, dtype=np.int64)
    perm = np.empty(n, dtype=np.int32)
    lv = np.empty(2, dtype=np.int32)
    assert L.ifs_tree_build_gpu(pts.ctypes.data, n, s, us.ctypes.data, perm.ctypes.data, lv.ctypes.data) == 0
    return us, perm, lv


def _clouds():
    rng = np.random.default_rng(21)
    yield "uniform", rng.uniform(-3.0, 5.0, (4000, 3))
    c = rng.uniform(0, 1, (40, 3))
    yield "clustered", np.concatenate([cc + 1e-3 * rng.standard_normal((100, 3)) for cc in c])
    yield "config2", config_cluster(2)["centres"]
    g = np.stack(np.meshgrid(*[np.arange(8.0)] * 3, indexing="ij"), -1).reshape(-1, 3)
    yield "grid with duplicates", np.concatenate([g, g[:50]])


@pytest.mark.parametrize("s", [1, 10, 32])
def test_build_bit_exact(L, s):
    for name, pts in _clouds():
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        u, us, lmax = _host_tree(L, pts, s)
        gu, perm, lv = _gpu_tree(L, pts, s)
        assert np.array_equal(gu, us), name
        assert np.array_equal(u[perm], gu), name
        for i in range(1, len(perm)):  # stable: equal keys keep ascending point order
            if gu[i] == gu[i - 1]:
                assert perm[i] > perm[i - 1]
        assert lv[0] == 2 and lv[1] == lmax, (name, lv, lmax)
        print(f"{name}: n={len(pts)} s={s} l_max={lmax}")


def test_i4_bit_exact(L):
    pts = np.ascontiguousarray(config_cluster(2)["centres"], dtype=np.float64)
    _, us, lmax = _host_tree(L, pts, 10)
    for l in range(2, lmax + 1):
        boxes = np.unique(us >> (3 * (21 - l))).astype(np.int64)
        out = np.empty(189 * len(boxes), dtype=np.int64)
        cnt = np.empty(len(boxes), dtype=np.int32)
        assert L.ifs_tree_i4_gpu(boxes.ctypes.data, len(boxes), l, out.ctypes.data, cnt.ctypes.data) == 0
        out = out.reshape(-1, 189)
        present = set(boxes.tolist())
        for q, b in enumerate(boxes):
            ref = np.empty(400, dtype=np.int64)
            c = L.ifs_tree_domain_i4(int(b), l, ref.ctypes.data, 400)
            ref = [v for v in ref[:c] if v in present]
            assert cnt[q] == len(ref) and list(out[q, :cnt[q]]) == ref, (l, b)
            assert np.all(out[q, cnt[q]:] == -1)


@pytest.mark.parametrize("s", [10, 32])
def test_build_vs_oracle_restatement(L, s):
    """The GPU build against the oracle's independent restatement of libtree
    (oracle/fmm_ref.py: floor(x 2^l) bit interleave, l_max by counting, I4 as the
    children of the parent's neighbourhood minus the own neighbourhood), on the same
    clouds: Morton keys at level 21, l_max and the I4 lists of the occupied boxes."""
    from oracle import fmm_ref
    for name, pts in _clouds():
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        gu, perm, lv = _gpu_tree(L, pts, s)
        scaled, _, _ = fmm_ref.scale_points(pts)
        keys = np.array([fmm_ref.morton(p, 21) for p in scaled], dtype=np.int64)
        assert np.array_equal(np.sort(keys), gu), name
        lmax = max(fmm_ref.level_max(scaled, s), 2)
        assert lv[1] == lmax, (name, lv, lmax)
        l = int(lv[1])
        boxes = np.unique(gu >> (3 * (21 - l))).astype(np.int64)
        out = np.empty(189 * len(boxes), dtype=np.int64)
        cnt = np.empty(len(boxes), dtype=np.int32)
        assert L.ifs_tree_i4_gpu(boxes.ctypes.data, len(boxes), l, out.ctypes.data, cnt.ctypes.data) == 0
        out = out.reshape(-1, 189)
        present = set(boxes.tolist())
        for q, b in enumerate(boxes):
            ref = [v for v in fmm_ref.i4(int(b), l) if v in present]
            assert list(out[q, :cnt[q]]) == ref, (name, l, b)
        print(f"{name}: n={len(pts)} s={s} l_max={l}, {len(boxes)} leaf boxes: GPU = oracle restatement")
