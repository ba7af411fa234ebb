"""The coaxial translation tables on their own, through the test-support entry point
ifs_coaxial_block (one order k of a table along the z axis, P = A + B and Q = A - B in
the normalised basis):

  * at sphere orders (L <= 49) the Gauss-Legendre generator of the box tables equals
    the degree recurrence of the sphere-level tables for outgoing -> regular shifts.
    The direct matvec built on those recurrence tables (tables kept, cache 0) matches
    the oracle's Gaunt/3j evaluation to 1e-14 per target at the same orders and
    lengths (test_gpu_high_order.py; lib_math_vector_spherical_harmonics.f90:1354-1585).
  * at the box orders of the 2000-sphere cluster (L = 409 and 785), where the
    recurrence fails, the quadrature tables satisfy the composition of shifts along
    one axis: R(t1) R(t2) = R(t1 + t2) for the regular shift (M2M, L2L) and
    R(t1) S(t2) = S(t1 + t2) for the outgoing -> regular one (M2L, t1 < t2).  The
    relation holds order by order for P and for Q (the coaxial operator
    [[A, B], [B, A]] of one order decouples into them).  Its intermediate sum runs over
    all degrees; the tables stop at L, so only the rows n <= L - w with
    w = t1 + 8 t1^(1/3) + 30 are compared: beyond that width the entries of R(t1)
    (Legendre coefficients of exp(i t1 mu)) vanish to rounding.  Composition pins the
    structure of the tables (degrees, orders, kinds and their relative scale); the
    value of t is pinned by the sphere-order comparison, which runs the same
    quadrature code.

Errors are measured against the largest entry of the table (over the orders fetched,
order 0 among them), the scale at which a table acts on an input of O(1) entries in the
normalised basis.  The blocks of the highest orders are smaller than that by up to
1e-27 (order 49 of a table of length 38) and carry only its rounding.

Measured: quadrature = recurrence to <= 2.6e-14 at sphere orders (bar 1e-10).  The
compositions hold to <= 1.3e-11 at L = 409 and to 1.6e-11 (regular) and 4.4e-10
(outgoing, order 0) at L = 785, hence their bar of 1e-9.  That figure is the tables'
own rounding: the compared rows need no intermediate degree above L, the products
round at ~1e-15, and the outgoing kind's T(mu) at t ~ 2L is a sum of ~2L Legendre
terms of total size sum_p (2p+1) |h1_p(t)| = 2.1e3 at L = 785 (1.1e3 at L = 409),
which the quadrature's sums carry into every entry."""
import ctypes
import math

import numpy as np
import pytest

from itofoggysim_amd._lib import lib

pytestmark = pytest.mark.gpu


def block(kind, t, Lt, Ls, K, method, k):
    """(P, Q) of order k: rows n = q..Lt, columns nu = q..Ls, q = max(1, |k|)."""
    q = max(1, abs(k))
    nr, nc = Lt - q + 1, Ls - q + 1
    out = np.empty(2 * nr * nc, dtype=np.complex128)
    rc = lib().ifs_coaxial_block(kind, float(t), Lt, Ls, K, method, k, out.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0, f"ifs_coaxial_block({kind}, {t}, {Lt}, {Ls}, {K}, {method}, {k}) = {rc}"
    return out[: nr * nc].reshape(nr, nc), out[nr * nc:].reshape(nr, nc)


def compare(pairs):
    """pairs: {(order, part): (value, reference)} -> {(order, part): (error relative to the
    largest reference entry of all blocks, error relative to the block's own largest)}."""
    scale = max(float(np.max(np.abs(b))) for _, b in pairs.values())
    out = {}
    for key, (a, b) in pairs.items():
        d = float(np.max(np.abs(a - b)))
        out[key] = (d / scale, d / max(float(np.max(np.abs(b))), 1e-300))
    return out


@pytest.mark.parametrize("Lt,Ls,t", [(10, 10, 38.0), (20, 20, 120.0), (49, 49, 38.0), (49, 49, 343.0),
                                     (49, 49, 1257.0), (49, 20, 114.0), (20, 49, 457.0)])
def test_quadrature_equals_recurrence_sphere_orders(Lt, Ls, t):
    K = min(Lt, Ls)
    pairs = {}
    for k in sorted({0, 1, -1, 2, K // 3, -(K // 2), K - 1, -K, K}):
        for part, a, b in zip("PQ", block(3, t, Lt, Ls, K, 1, k), block(3, t, Lt, Ls, K, 0, k)):
            pairs[(k, part)] = (a, b)
    err = compare(pairs)
    for (k, part), (e, eb) in err.items():
        print(f"Lt={Lt} Ls={Ls} t={t} order {k} {part}: quadrature vs recurrence {e:.1e} of the table, "
              f"{eb:.1e} of the block")
    worst = max(e for e, _ in err.values())
    assert worst < 1e-10, f"quadrature vs recurrence {worst:.2e} of the table's largest entry"


@pytest.mark.parametrize("L,K,orders", [(409, 250, (0, 1, -2, 100, -200, 250)), (785, 2, (0, 1, -2))])
def test_quadrature_composition_box_orders(L, K, orders):
    t1, t2 = 40.0, 60.0      # regular: R(40) R(60) = R(100)
    s2 = 2.0 * L             # outgoing -> regular at two box widths (M2L): R(40) S(2L) = S(2L + 40)
    w = int(math.ceil(t1 + 8.0 * t1 ** (1.0 / 3.0) + 30.0))
    reg, out = {}, {}
    for k in orders:
        q = max(1, abs(k))
        rows = L - w - q + 1    # rows n = q..L - w
        assert rows > 0
        R1, R2, R12 = (block(1, t, L, L, K, 1, k) for t in (t1, t2, t1 + t2))
        S2, S12 = (block(3, t, L, L, K, 1, k) for t in (s2, s2 + t1))
        for part in (0, 1):
            reg[(k, "PQ"[part])] = ((R1[part] @ R2[part])[:rows], R12[part][:rows])
            out[(k, "PQ"[part])] = ((R1[part] @ S2[part])[:rows], S12[part][:rows])
    er, es = compare(reg), compare(out)
    for key in reg:
        print(f"L={L} order {key[0]} {key[1]} (rows n <= {L - w}): R(t1)R(t2) vs R(t1+t2) {er[key][0]:.1e} of the "
              f"table ({er[key][1]:.1e} of the block), R(t1)S(t2) vs S(t1+t2) {es[key][0]:.1e} "
              f"({es[key][1]:.1e})")
    assert max(e for e, _ in er.values()) < 1e-9
    assert max(e for e, _ in es.values()) < 1e-9
