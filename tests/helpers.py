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
