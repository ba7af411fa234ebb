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
               z_selector=3, use_fmm=None, precision="fp64", leaf_size=32, fmm_d0=6.0):
        """Queues one request; directions = [(alpha, beta), ...]; returns its id.
        use_fmm None: the MLFMM above 200 spheres; precision 'fp64' or 'mixed'."""
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
        pr.use_fmm = int(ns > 200 if use_fmm is None else bool(use_fmm))
        pr.leaf_size = int(leaf_size)
        pr.fmm_d0 = float(fmm_d0)
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        pr.precision = 0 if precision == "fp64" else 1
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

    def set_verbose(self, verbose=True):
        """Progress on stderr: each request's GMRES steps and its completion."""
        check(lib().ifs_server_set_verbose(self._h, int(bool(verbose))), "ifs_server_set_verbose")

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


def plan_request(centres, radii, k, ndir, lmax=None, restart=30, n_global=None, use_fmm=None, leaf_size=32,
                 fmm_d0=6.0, hbm_budget_gb=200.0, slots=4):
    """The server's admission plan of one request, on the host (ifs_server_plan): bytes of
    vectors, MLFMM pools, workspace, kept tables, fixed buffers and their total, and the
    workspace_mb / cache_mb its context would get (cache_mb -1: no tables kept)."""
    centres = np.ascontiguousarray(centres, dtype=np.float64).reshape(-1, 3)
    radii = np.ascontiguousarray(radii, dtype=np.float64)
    ns = radii.shape[0]
    if lmax is None:
        x = k * radii
        lmax = np.ceil(x + 4.05 * np.cbrt(x) + 2.0).astype(np.int32)
    lmax = np.ascontiguousarray(lmax, dtype=np.int32)
    m = np.ones(ns, dtype=np.complex128)
    cl = _lib.Cluster(ns, centres.ctypes.data, radii.ctypes.data, m.ctypes.data, lmax.ctypes.data)
    pr = _lib.Params()
    pr.k = float(k)
    pr.n_global = int(n_global if n_global is not None else lmax.max())
    pr.use_fmm = int(ns > 200 if use_fmm is None else bool(use_fmm))
    pr.leaf_size = int(leaf_size)
    pr.fmm_d0 = float(fmm_d0)
    out = np.zeros(8)
    check(lib().ifs_server_plan(ctypes.byref(cl), ctypes.byref(pr), int(ndir), int(restart), float(hbm_budget_gb),
                                int(slots), out.ctypes.data), "ifs_server_plan")
    keys = ["vectors", "pools", "workspace", "tables", "fixed", "total", "workspace_mb", "cache_mb"]
    return {k_: (int(v) if k_ in ("workspace_mb", "cache_mb") else float(v)) for k_, v in zip(keys, out)}


def water_index(wavelength_um):
    """n(lambda) = 1.3199 + 3500 nm^2 / lambda^2, k = 1e-8 (lambda in um)."""
    return complex(1.3199 + 3500e-6 / wavelength_um ** 2, 1e-8)


def poisson_workload(n_requests, rate_hz, seed=0, n_min=10, n_max=20000, max_dirs=16, mixed_fraction=0.5):
    """Requests of the serving workload: arrival times of a Poisson process of the given
    rate; per request (seed = request id) a cluster of N log-uniform in [n_min, n_max]
    spheres with the 20-sphere benchmark's radius law (uniform 0.3-1.0 um) and number
    density (a cube of 10 (N/20)^(1/3) um), lambda uniform in [400, 1550] nm with the water
    index above, 1..max_dirs incidence directions, tagged plain FP64 or mixed."""
    from .cluster import make_cluster
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n_requests):
        t += rng.exponential(1.0 / rate_hz)
        n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        lam = float(rng.uniform(0.4, 1.55))
        m = water_index(lam)
        side = 10.0 * (n / 20.0) ** (1.0 / 3.0)
        c = make_cluster(n, ("uniform", 0.3, 1.0), side=side, wavelength=lam, m=m, seed=i)
        nd = int(rng.integers(1, max_dirs + 1))
        dirs = [(float(np.arccos(rng.uniform(-1, 1))), float(rng.uniform(0, 2 * np.pi))) for _ in range(nd)]
        out.append(dict(arrival_s=t, cluster=c, directions=dirs,
                        precision="mixed" if rng.uniform() < mixed_fraction else "fp64"))
    return out


def run_workload(server, requests, tol=1e-6, restart=30, clock=None):
    """Submits the requests at their arrival times and waits for all; returns
    (per-request results, server stats)."""
    import time
    clock = clock or time.perf_counter
    t0 = clock()
    ids = []
    for r in requests:
        delay = r["arrival_s"] - (clock() - t0)
        if delay > 0:
            time.sleep(delay)
        c = r["cluster"]
        ids.append(server.submit(c["centres"], c["radii"], c["m_rel"], c["k"], r["directions"], lmax=c["lmax"],
                                 tol=tol, restart=restart, precision=r["precision"]))
    res = [server.wait(i) for i in ids]
    return res, server.stats()
