"""CPU restatement of the octree and the MLFMM matvec (test infrastructure only).

The octree is restated from its definitions, independently of the product's
C++ code:
  * scaling to [0,1)^3 with one edge D = max extent * (1 + 2^-52)
    (libtree/src/lib/lib_tree.f90:320-374);
  * Morton index at level l = interleave of floor(x * 2^l) per axis, x the most
    significant bit of every triple (lib_tree_helper_functions.f90:202-364);
  * l_max = smallest level at which no box holds more than s elements
    (lib_tree.f90:1188-1242);
  * I2 = own box and the boxes adjacent to it, I4 = children of the parent's I2
    minus own I2 (lib_tree.f90:898-1143), E2 = elements of I2.

The MLFMM passes follow libmlfmm (libmlfmm/src/lib/lib_ml_fmm.f90:231-1204) with
full Xu translation matrices from the C oracle (Gaunt / 3j formula) in Xu's basis,
translation vector d = dsign * (origin(source) - origin(target)) throughout, j_n for
outgoing->outgoing and regular->regular, h1_n for outgoing->regular.
"""
import math

import numpy as np

from . import oracle as O

L_TH = 21


def scale_points(pts):
    pts = np.asarray(pts, dtype=np.float64)
    lo = pts.min(axis=0)
    hi = pts.max(axis=0)
    D = ((hi - lo) * (1.0 + 2.0 ** -52)).max()
    if not D > 0.0:
        D = 1.0
    return (pts - lo) / D, lo, D


def morton(scaled_point, level):
    n = 0
    q = [int(math.floor(v * (2 ** level))) for v in scaled_point]
    for i in range(level - 1, -1, -1):
        n = (n << 3) | (((q[0] >> i) & 1) << 2) | (((q[1] >> i) & 1) << 1) | ((q[2] >> i) & 1)
    return n


def demorton(n, level):
    v = [0, 0, 0]
    for i in range(level - 1, -1, -1):
        t = (n >> (3 * i)) & 7
        v[0] = (v[0] << 1) | ((t >> 2) & 1)
        v[1] = (v[1] << 1) | ((t >> 1) & 1)
        v[2] = (v[2] << 1) | (t & 1)
    return v


def from_coords(v, level):
    n = 0
    for i in range(level - 1, -1, -1):
        n = (n << 3) | (((v[0] >> i) & 1) << 2) | (((v[1] >> i) & 1) << 1) | ((v[2] >> i) & 1)
    return n


def level_max(scaled, s):
    """Smallest level >= 1 with at most s points per box."""
    for level in range(1, L_TH + 1):
        counts = {}
        for p in scaled:
            b = morton(p, level)
            counts[b] = counts.get(b, 0) + 1
        if max(counts.values()) <= s:
            return level
    return L_TH


def adjacent(n, level):
    """Boxes at Chebyshev distance exactly 1 (inside the domain)."""
    v = demorton(n, level)
    top = 2 ** level - 1
    out = []
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                if dx == dy == dz == 0:
                    continue
                w = [v[0] + dx, v[1] + dy, v[2] + dz]
                if min(w) < 0 or max(w) > top:
                    continue
                out.append(from_coords(w, level))
    return out


def i4(n, level):
    own = set(adjacent(n, level)) | {n}
    p = n >> 3
    par = set(adjacent(p, level - 1)) | {p}
    kids = {8 * q + j for q in par for j in range(8)}
    return sorted(kids - own)


def centre(n, level):
    v = demorton(n, level)
    h = 0.5 ** level
    return np.array([h * (x + 0.5) for x in v])


def box_L(k, edge, d0, cap):
    kD = k * edge
    L = int(math.ceil(kD + 1.8 * d0 ** (2.0 / 3.0) * kD ** (1.0 / 3.0)))
    L = max(L, 1)
    if cap > 0:
        L = min(L, cap)
    return L


def build_tree(pos, s):
    scaled, lo, D = scale_points(pos)
    lmin = 2
    lmax = max(level_max(scaled, s), lmin)
    leaf = [morton(p, lmax) for p in scaled]
    return dict(scaled=scaled, lo=lo, D=D, lmin=lmin, lmax=lmax, leaf=leaf)


def _translate(A, B, xa, xb):
    return A @ xa + B @ xb, B @ xa + A @ xb


def _T(src, dst, k, Ls, Lt, zsel, dsign=1.0):
    d = dsign * (np.asarray(src, dtype=np.float64) - np.asarray(dst, dtype=np.float64))
    r = float(np.linalg.norm(d))
    th = math.acos(max(-1.0, min(1.0, d[2] / r))) if r > 0 else 0.0
    ph = math.atan2(d[1], d[0]) if (d[0] != 0.0 or d[1] != 0.0) else 0.0
    return O.translation_ab(k * r, th, ph, Lt, Ls, zsel)


def fmm_matvec(pos, lmax_j, nglob, an, bn, k, x, s=32, d0=6.0, cap=0, zsel=3, dsign=1.0):
    """b = x + T (far field + near field) x in the reference flat layout.

    dsign = +1: d = origin(source) - origin(target) as the reference's direct path
    passes it; dsign = -1: the opposite sign."""
    pos = np.asarray(pos, dtype=np.float64)
    ns = len(pos)
    tr = build_tree(pos, s)
    lmin, lmax, D, lo = tr["lmin"], tr["lmax"], tr["D"], tr["lo"]
    Leff = [min(int(L), nglob) for L in lmax_j]
    blk = nglob * (nglob + 2)

    def xs(i):
        c = Leff[i] * (Leff[i] + 2)
        base = i * 2 * blk
        return x[base:base + c], x[base + blk:base + blk + c]

    boxes = {}
    for l in range(lmin, lmax + 1):
        bl = sorted({b >> (3 * (lmax - l)) for b in tr["leaf"]})
        Lb = box_L(k, D / 2 ** l, d0, cap)
        cen = {b: centre(b, l) * D + lo for b in bl}
        boxes[l] = dict(boxes=bl, L=Lb, centre=cen)
    C = {l: {b: [np.zeros(boxes[l]["L"] * (boxes[l]["L"] + 2), complex) for _ in range(2)]
             for b in boxes[l]["boxes"]} for l in boxes}
    Dl = {l: {b: [np.zeros(boxes[l]["L"] * (boxes[l]["L"] + 2), complex) for _ in range(2)]
              for b in boxes[l]["boxes"]} for l in boxes}
    LL = boxes[lmax]
    # P2M
    for i in range(ns):
        b = tr["leaf"][i]
        A, B = _T(pos[i], LL["centre"][b], k, Leff[i], LL["L"], 1, dsign)
        ya, yb = _translate(A, B, *xs(i))
        C[lmax][b][0] += ya
        C[lmax][b][1] += yb
    # M2M
    for l in range(lmax - 1, lmin - 1, -1):
        for b in boxes[l + 1]["boxes"]:
            p = b >> 3
            A, B = _T(boxes[l + 1]["centre"][b], boxes[l]["centre"][p], k, boxes[l + 1]["L"], boxes[l]["L"], 1, dsign)
            ya, yb = _translate(A, B, *C[l + 1][b])
            C[l][p][0] += ya
            C[l][p][1] += yb
    # M2L and L2L
    for l in range(lmin, lmax + 1):
        V = boxes[l]
        present = set(V["boxes"])
        for b in V["boxes"]:
            for src in i4(b, l):
                if src not in present:
                    continue
                A, B = _T(V["centre"][src], V["centre"][b], k, V["L"], V["L"], zsel, dsign)
                ya, yb = _translate(A, B, *C[l][src])
                Dl[l][b][0] += ya
                Dl[l][b][1] += yb
        if l > lmin:
            P = boxes[l - 1]
            for b in V["boxes"]:
                p = b >> 3
                A, B = _T(P["centre"][p], V["centre"][b], k, P["L"], V["L"], 1, dsign)
                ya, yb = _translate(A, B, *Dl[l - 1][p])
                Dl[l][b][0] += ya
                Dl[l][b][1] += yb
    # L2P + P2P + T + identity
    out = np.array(x, dtype=np.complex128)
    members = {}
    for i in range(ns):
        members.setdefault(tr["leaf"][i], []).append(i)
    for j in range(ns):
        b = tr["leaf"][j]
        A, B = _T(LL["centre"][b], pos[j], k, LL["L"], Leff[j], 1, dsign)
        va, vb = _translate(A, B, *Dl[lmax][b])
        near = set(adjacent(b, lmax)) | {b}
        for nb in sorted(near):
            for i in members.get(nb, []):
                if i == j:
                    continue
                A, B = _T(pos[i], pos[j], k, Leff[i], Leff[j], zsel, dsign)
                ya, yb = _translate(A, B, *xs(i))
                va = va + ya
                vb = vb + yb
        c = Leff[j] * (Leff[j] + 2)
        nidx = np.concatenate([[n] * (2 * n + 1) for n in range(1, Leff[j] + 1)]) - 1
        base = j * 2 * blk
        out[base:base + c] += an[j][nidx] * va
        out[base + blk:base + blk + c] += bn[j][nidx] * vb
    return out, tr
