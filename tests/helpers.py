"""Shared helpers of the tests (no test functions here)."""
import numpy as np


def snm_flat(N, lmax):
    """Per-entry factors s_nm = sqrt(n(n+1)(2n+1)(n-m)!/(n+m)!) of the normalised basis
    over the flat vector (a- and b-block of every sphere), and the mask of the rows
    n <= L_j.  In the flat basis the entries of one sphere span 1e+-80 at N = 50, so
    comparisons there measure the relative accuracy of the tiniest couplings (rotation
    elements d^n_{n,-n} ~ (theta/2)^(2n)); physical vectors are O(1) in the normalised
    basis, where the comparisons below are made."""
    from scipy.special import gammaln
    nn = np.array([q for q in range(1, N + 1) for m in range(-q, q + 1)], dtype=float)
    mm = np.array([m for q in range(1, N + 1) for m in range(-q, q + 1)], dtype=float)
    s = np.exp(0.5 * (np.log(nn * (nn + 1) * (2 * nn + 1)) + gammaln(nn - mm + 1) - gammaln(nn + mm + 1)))
    mask = np.concatenate([np.tile(nn <= min(l, N), 2) for l in lmax])
    return np.tile(np.concatenate([s, s]), len(lmax)), mask


def per_sphere_rel(b, ref, x, ns, nglob):
    """|b - ref| / |ref - x| per sphere block (the translated part only)."""
    blk = 2 * nglob * (nglob + 2)
    out = []
    for j in range(ns):
        s = slice(j * blk, (j + 1) * blk)
        den = np.linalg.norm(ref[s] - x[s])
        out.append(np.linalg.norm(b[s] - ref[s]) / den if den > 0 else 0.0)
    return np.array(out)


class Hip:
    """Device buffers through the HIP runtime that libifs.so itself links (ctypes),
    for GPU tests that hand device pointers to the C ABI.  (torch bundles a second
    HIP runtime; a process that initialises it after libifs.so has created and
    destroyed contexts finds no GPU, so these tests stay off torch.)"""

    def __init__(self):
        import ctypes
        self.ct = ctypes
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        self.live = []

    def alloc(self, nbytes):
        p = self.ct.c_void_p()
        assert self.rt.hipMalloc(self.ct.byref(p), max(int(nbytes), 16)) == 0
        self.live.append(p)
        return p

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        assert self.rt.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def sync(self, stream):
        assert self.rt.hipStreamSynchronize(stream) == 0

    def free_all(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []
