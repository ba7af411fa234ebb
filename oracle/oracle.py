"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, the smoke check and the bench's CPU-baseline leg import this
module, and only as the checker.  The product package never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libifs_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.ifs_or_sph_j.argtypes = [ctypes.c_double, ctypes.c_int, _dp]
        L.ifs_or_sph_y.argtypes = [ctypes.c_double, ctypes.c_int, _dp]
        L.ifs_or_legendre.argtypes = [ctypes.c_double, ctypes.c_int, _dp]
        L.ifs_or_fact_ratio.argtypes = [ctypes.c_int, ctypes.c_int]
        L.ifs_or_fact_ratio.restype = ctypes.c_double
        L.ifs_or_wigner3j.argtypes = [ctypes.c_int] * 6
        L.ifs_or_wigner3j.restype = ctypes.c_double
        L.ifs_or_wigner3j_all.argtypes = [ctypes.c_int] * 4 + [_dp, _ip]
        L.ifs_or_xu_ab.argtypes = [ctypes.c_int] * 5 + [_dp, _dp]
        L.ifs_or_translation_ab.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int] * 3 + [ctypes.c_void_p] * 2
        L.ifs_or_mie_ab.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int] + [ctypes.c_void_p] * 2
        L.ifs_or_matvec_direct.argtypes = [ctypes.c_int, _dp, _ip, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_int]
        L.ifs_or_matvec_pairs.argtypes = [ctypes.c_int, _dp, _ip, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_double, ctypes.c_int, ctypes.c_int, _ip, _ip, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_int]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def fidx(n, m):
    """Flat index of (n, m) inside one a- or b-block."""
    return n * n - 1 + m + n


def sph_j(x, nmax):
    out = np.zeros(nmax + 1)
    assert lib().ifs_or_sph_j(float(x), nmax, _d(out)) == 0
    return out


def sph_y(x, nmax):
    out = np.zeros(nmax + 1)
    assert lib().ifs_or_sph_y(float(x), nmax, _d(out)) == 0
    return out


def legendre(x, nmax):
    """P_n^m(x) without Condon-Shortley phase; out[n*n + m + n]."""
    out = np.zeros((nmax + 1) ** 2)
    assert lib().ifs_or_legendre(float(x), nmax, _d(out)) == 0
    return out


def fact_ratio(n, m):
    return lib().ifs_or_fact_ratio(n, m)


def wigner3j(j1, j2, j3, m1, m2, m3):
    return lib().ifs_or_wigner3j(j1, j2, j3, m1, m2, m3)


def xu_ab(m, n, mu, nu, p):
    a = ctypes.c_double()
    b = ctypes.c_double()
    lib().ifs_or_xu_ab(m, n, mu, nu, p, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def translation_ab(x, theta, phi, nmax, numax, zsel):
    nr = nmax * (nmax + 2)
    nc = numax * (numax + 2)
    A = np.zeros((nr, nc), dtype=np.complex128)
    B = np.zeros((nr, nc), dtype=np.complex128)
    rc = lib().ifs_or_translation_ab(float(x), float(theta), float(phi), nmax, numax, zsel,
                                     A.ctypes.data, B.ctypes.data)
    assert rc == 0, rc
    return A, B


def mie_ab(x, m, nmax):
    a = np.zeros(nmax, dtype=np.complex128)
    b = np.zeros(nmax, dtype=np.complex128)
    rc = lib().ifs_or_mie_ab(float(x), float(np.real(m)), float(np.imag(m)), nmax, a.ctypes.data, b.ctypes.data)
    assert rc == 0, rc
    return a, b


def matvec_direct(pos, lmax, nglob, an, bn, k, x, zsel=3, ref_row_zeroing=False, nthreads=0):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    lmax = np.ascontiguousarray(lmax, dtype=np.int32)
    an = np.ascontiguousarray(an, dtype=np.complex128)
    bn = np.ascontiguousarray(bn, dtype=np.complex128)
    x = np.ascontiguousarray(x, dtype=np.complex128)
    ns = lmax.shape[0]
    assert an.shape == (ns, nglob) and bn.shape == (ns, nglob)
    assert x.shape == (ns * 2 * nglob * (nglob + 2),)
    b = np.zeros_like(x)
    rc = lib().ifs_or_matvec_direct(ns, _d(pos), _i(lmax), nglob, an.ctypes.data, bn.ctypes.data, float(k),
                                    zsel, int(bool(ref_row_zeroing)), x.ctypes.data, b.ctypes.data, nthreads)
    assert rc == 0, rc
    return b


def matvec_pairs(pos, lmax, nglob, an, bn, k, x, tgt, src, zsel=3, nthreads=0):
    """T A x restricted to the ordered pairs (tgt[i], src[i]) (no identity)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    lmax = np.ascontiguousarray(lmax, dtype=np.int32)
    an = np.ascontiguousarray(an, dtype=np.complex128)
    bn = np.ascontiguousarray(bn, dtype=np.complex128)
    x = np.ascontiguousarray(x, dtype=np.complex128)
    tgt = np.ascontiguousarray(tgt, dtype=np.int32)
    src = np.ascontiguousarray(src, dtype=np.int32)
    ns = lmax.shape[0]
    b = np.zeros_like(x)
    rc = lib().ifs_or_matvec_pairs(ns, _d(pos), _i(lmax), nglob, an.ctypes.data, bn.ctypes.data, float(k), zsel,
                                   len(tgt), _i(tgt), _i(src), x.ctypes.data, b.ctypes.data, nthreads)
    assert rc == 0, rc
    return b
