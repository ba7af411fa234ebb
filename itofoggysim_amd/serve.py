"""Serving: continuous batching of scattering requests on one GPU (C ABI
ifs_server_*).  Each request is a cluster plus incidence directions; its result is
the cross sections (C_sca, C_ext, C_abs) per direction."""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib


class Server:
    def __init__(self, slots=4, hbm_budget_gb=64.0):
        h = ctypes.c_void_p()
        check(lib().ifs_server_create(ctypes.byref(h), int(slots), float(hbm_budget_gb)), "ifs_server_create")
        self._h = h

    def submit(self, centres, radii, m_rel, k, directions, lmax=None, tol=1e-6, restart=30, n_global=None,
               z_selector=3):
        """Queues one request; directions = [(alpha, beta), ...]; returns its id."""
        centres = np.ascontiguousarray(centres, dtype=np.float64).reshape(-1, 3)
        radii = np.ascontiguousarray(radii, dtype=np.float64)
        ns = radii.shape[0]
        m = np.ascontiguousarray(np.broadcast_to(np.asarray(m_rel, dtype=np.complex128), (ns,)))
        if lmax is None:
            x = k * radii
            lmax = np.ceil(x + 4.05 * np.cbrt(x) + 2.0).astype(np.int32)
        lmax = np.ascontiguousarray(lmax, dtype=np.int32)
        cl = _lib.Cluster(ns, centres.ctypes.data, radii.ctypes.data, m.ctypes.data, lmax.ctypes.data)
        pr = _lib.Params()
        pr.k = float(k)
        pr.n_global = int(n_global if n_global is not None else lmax.max())
        pr.z_selector = int(z_selector)
        pr.use_fmm = 0
        d = np.asarray(directions, dtype=np.float64).reshape(-1, 2)
        al = np.ascontiguousarray(d[:, 0])
        be = np.ascontiguousarray(d[:, 1])
        rid = ctypes.c_int64()
        check(lib().ifs_server_submit(self._h, ctypes.byref(cl), ctypes.byref(pr), len(al), al.ctypes.data,
                                      be.ctypes.data, float(tol), int(restart), ctypes.byref(rid)),
              "ifs_server_submit")
        self._ndir = getattr(self, "_ndir", {})
        self._ndir[rid.value] = len(al)
        return rid.value

    def wait(self, rid):
        """(cross sections [ndir, 3], info dict) of a request."""
        nd = self._ndir.pop(rid)
        out = np.zeros(3 * nd)
        inf = _lib.RequestInfo()
        rc = lib().ifs_server_wait(self._h, int(rid), out.ctypes.data, ctypes.byref(inf))
        if rc not in (0, -5):
            check(rc, "ifs_server_wait")
        return out.reshape(nd, 3), dict(status=inf.status, iterations=inf.iterations, latency_s=inf.latency_s,
                                        queue_s=inf.queue_s, footprint_bytes=inf.footprint_bytes)

    def stats(self):
        out = np.zeros(6)
        check(lib().ifs_server_stats(self._h, out.ctypes.data), "ifs_server_stats")
        return dict(completed=int(out[0]), p50_s=out[1], p99_s=out[2], solves_per_s=out[3],
                    peak_concurrent=int(out[4]), peak_admitted_bytes=int(out[5]))

    def close(self):
        if getattr(self, "_h", None):
            lib().ifs_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
