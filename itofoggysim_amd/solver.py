"""Host-side mirror of the reference's multi-sphere solver interface.

``MultiSphereSolver`` follows lib_mie_ms_solver_interface
(FoggySim/src/lib/mie_theory/lib_mie_ms_solver_interface.f90): the constructor
builds the problem (:91-129), ``calculate_vector_b`` is the GMRES matvec callback
(:739-783), ``get_vector_b`` the right-hand side (:230-315) and ``solve`` the
GMRES run (lib_solver_GMRES.f90:116-316).  All compute happens in libifs.so.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib


def _prec(p):
    """'fp64' (IFS_PREC_FP64) or 'mixed' (IFS_PREC_MIXED: FP32 matrix-core box translations)."""
    if p not in ("fp64", "mixed"):
        raise ValueError("precision must be 'fp64' or 'mixed'")
    return 0 if p == "fp64" else 1


class MultiSphereSolver:
    def __init__(self, centres, radii, m_rel, k, n_global=None, lmax=None, z_selector=3, ref_row_zeroing=False,
                 use_fmm=False, leaf_size=32, fmm_d0=6.0, fmm_lmax_cap=0, workspace_mb=0,
                 d_convention="physical", cache_mb=0, precision="fp64"):
        centres = np.ascontiguousarray(centres, dtype=np.float64).reshape(-1, 3)
        radii = np.ascontiguousarray(radii, dtype=np.float64)
        ns = radii.shape[0]
        m_rel = np.ascontiguousarray(np.broadcast_to(np.asarray(m_rel, dtype=np.complex128), (ns,)))
        if lmax is None:
            x = k * radii
            lmax = np.ceil(x + 4.05 * np.cbrt(x) + 2.0).astype(np.int32)
        lmax = np.ascontiguousarray(lmax, dtype=np.int32)
        if n_global is None:
            n_global = int(lmax.max())
        self._keep = (centres, radii, m_rel, lmax)
        cl = _lib.Cluster(ns, centres.ctypes.data, radii.ctypes.data, m_rel.ctypes.data, lmax.ctypes.data)
        pr = _lib.Params()
        pr.k = float(k)
        pr.n_global = int(n_global)
        pr.z_selector = int(z_selector)
        pr.ref_row_zeroing = int(bool(ref_row_zeroing))
        pr.use_fmm = int(bool(use_fmm))
        pr.leaf_size = int(leaf_size)
        pr.fmm_d0 = float(fmm_d0)
        pr.fmm_lmax_cap = int(fmm_lmax_cap)
        pr.workspace_mb = int(workspace_mb)
        if d_convention not in ("physical", "reference"):
            raise ValueError("d_convention must be 'physical' or 'reference'")
        pr.d_convention = 0 if d_convention == "physical" else 1
        pr.cache_mb = int(cache_mb)
        pr.precision = _prec(precision)
        h = ctypes.c_void_p()
        check(lib().ifs_create(ctypes.byref(h), ctypes.byref(cl), ctypes.byref(pr)), "ifs_create")
        self._h = h
        self.n_spheres = ns
        self.n_global = int(n_global)
        self.lmax = lmax
        self.k = float(k)
        self.n = int(lib().ifs_vector_length(h))

    # -- reference interface ------------------------------------------------
    def calculate_vector_b(self, vector_x):
        """b = (I + T A) x on host arrays (lib_mie_ms_solver_calculate_vector_b)."""
        x = np.ascontiguousarray(vector_x, dtype=np.complex128)
        if x.shape != (self.n,):
            raise ValueError(f"vector_x has shape {x.shape}, expected ({self.n},)")
        b = np.empty_like(x)
        check(lib().ifs_matvec(self._h, x.ctypes.data, b.ctypes.data, self.n), "ifs_matvec")
        return b

    def calculate_vector_b_device(self, x_ptr, b_ptr):
        """Same on device pointers (e.g. torch tensors' data_ptr()), enqueued on the launch stream."""
        check(lib().ifs_matvec_device(self._h, ctypes.c_void_p(x_ptr), ctypes.c_void_p(b_ptr), self.n),
              "ifs_matvec_device")

    def calculate_vector_b_graph_device(self, x_ptr, b_ptr):
        """Same through a HIP graph of the matvec's launches (ifs_matvec_graph_device)."""
        check(lib().ifs_matvec_graph_device(self._h, ctypes.c_void_p(x_ptr), ctypes.c_void_p(b_ptr), self.n),
              "ifs_matvec_graph_device")

    def get_vector_b(self):
        """Right-hand side for an x-polarised plane wave along +z (lib_mie_ms_solver_get_vector_b_all)."""
        rhs = np.empty(self.n, dtype=np.complex128)
        check(lib().ifs_plane_wave_rhs(self._h, rhs.ctypes.data, self.n), "ifs_plane_wave_rhs")
        return rhs

    def plane_wave(self, alpha=0.0, beta=0.0):
        """Incident coefficients (p_mn, q_mn) of a unit plane wave along (alpha, beta)
        (get_p_q_j_j_plane_wave_core, lib_mie_illumination.f90:636-763)."""
        p = np.empty(self.n, dtype=np.complex128)
        check(lib().ifs_plane_wave(self._h, float(alpha), float(beta), p.ctypes.data, self.n), "ifs_plane_wave")
        return p

    def get_vector_b_dir(self, alpha=0.0, beta=0.0, te=False):
        """Right-hand side T p for a plane wave along (alpha, beta); te: the TE mode (p and
        q exchanged, lib_mie_illumination.f90:347-352)."""
        rhs = np.empty(self.n, dtype=np.complex128)
        check(lib().ifs_plane_wave_rhs_mode(self._h, float(alpha), float(beta), int(bool(te)), rhs.ctypes.data,
                                            self.n), "ifs_plane_wave_rhs_mode")
        return rhs

    def plane_wave_mode(self, alpha=0.0, beta=0.0, te=False):
        """Incident coefficients of the TM (te=False) or TE (te=True) wave along (alpha, beta)."""
        p = np.empty(self.n, dtype=np.complex128)
        check(lib().ifs_plane_wave_mode(self._h, float(alpha), float(beta), int(bool(te)), p.ctypes.data, self.n),
              "ifs_plane_wave_mode")
        return p

    def cross_sections(self, x, p):
        """(C_sca, C_ext, C_abs) of the cluster from solved coefficients x and the incident p."""
        x = np.ascontiguousarray(x, dtype=np.complex128)
        p = np.ascontiguousarray(p, dtype=np.complex128)
        out = np.zeros(3)
        check(lib().ifs_cross_sections(self._h, x.ctypes.data, p.ctypes.data, self.n, out.ctypes.data),
              "ifs_cross_sections")
        return tuple(out)

    def far_field(self, x, theta, phi):
        """(F_theta, F_phi) at the directions (theta, phi): E_s ~ e^{ikr}/(kr) F."""
        x = np.ascontiguousarray(x, dtype=np.complex128)
        th = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
        ph = np.ascontiguousarray(np.broadcast_to(np.atleast_1d(phi), th.shape), dtype=np.float64)
        out = np.empty(2 * th.size, dtype=np.complex128)
        check(lib().ifs_far_field(self._h, x.ctypes.data, self.n, th.size, th.ctypes.data, ph.ctypes.data,
                                  out.ctypes.data), "ifs_far_field")
        return out[0::2], out[1::2]

    @staticmethod
    def amplitude_matrix(phi, Fx, Fy):
        """S1..S4 from the far fields of the x- and y-polarised waves along +z."""
        ph = np.ascontiguousarray(np.atleast_1d(phi), dtype=np.float64)
        fx = np.ascontiguousarray(np.stack([Fx[0], Fx[1]], axis=1).ravel(), dtype=np.complex128)
        fy = np.ascontiguousarray(np.stack([Fy[0], Fy[1]], axis=1).ravel(), dtype=np.complex128)
        S = np.empty(4 * ph.size, dtype=np.complex128)
        check(lib().ifs_amplitude_matrix(ph.size, ph.ctypes.data, fx.ctypes.data, fy.ctypes.data, S.ctypes.data),
              "ifs_amplitude_matrix")
        S = S.reshape(-1, 4)
        return S[:, 0], S[:, 1], S[:, 2], S[:, 3]

    @staticmethod
    def mueller(S1, S2, S3, S4):
        """Mueller matrices [ndir, 4, 4]."""
        S = np.ascontiguousarray(np.stack([S1, S2, S3, S4], axis=1).ravel(), dtype=np.complex128)
        nd = S.size // 4
        M = np.empty(16 * nd)
        check(lib().ifs_mueller(nd, S.ctypes.data, M.ctypes.data), "ifs_mueller")
        return M.reshape(nd, 4, 4)

    def mie_coefficients(self):
        """(a_n^j, b_n^j) as arrays [n_spheres, N] (lib_mie_type_func_get_sphere_parameter)."""
        an = np.empty((self.n_spheres, self.n_global), dtype=np.complex128)
        bn = np.empty_like(an)
        check(lib().ifs_get_mie(self._h, an.ctypes.data, bn.ctypes.data), "ifs_get_mie")
        return an, bn

    def solve(self, rhs=None, restart=50, max_iter=1000, tol=1e-6, x0=None, raise_on_noconv=True, precision="fp64",
              verbose=False, graph=False):
        """Restarted GMRES on the device; returns (x, info dict) (lib_mie_ms_solver_run).
        precision 'mixed': FP32 box translations in the Arnoldi products, FP64 true residual
        per cycle (iterative refinement).  verbose: progress lines on stderr."""
        b = self.get_vector_b() if rhs is None else np.ascontiguousarray(rhs, dtype=np.complex128)
        x = np.zeros(self.n, dtype=np.complex128) if x0 is None else np.array(x0, dtype=np.complex128)
        op = _lib.SolveOpts(int(restart), int(max_iter), float(tol), int(x0 is not None), int(bool(verbose)),
                            _prec(precision), _lib.IFS_SOLVE_GRAPH if graph else 0)
        inf = _lib.SolveInfo()
        rc = lib().ifs_solve(self._h, b.ctypes.data, x.ctypes.data, self.n, ctypes.byref(op), ctypes.byref(inf))
        if rc not in (0, -5, -6) or (rc != 0 and raise_on_noconv):
            check(rc, "ifs_solve")
        return x, dict(iterations=inf.iterations, converged=bool(inf.converged), rel_residual=inf.rel_residual,
                       rel_residual_flat=inf.rel_residual_flat, seconds=inf.seconds)

    def calculate_vector_b_batch_device(self, x_ptr, b_ptr, nrhs, ld=None):
        """b_r = (I + T A) x_r for nrhs device vectors at stride ld (default n), one batched pass."""
        check(lib().ifs_matvec_batch_device(self._h, int(nrhs), ctypes.c_void_p(x_ptr), ctypes.c_void_p(b_ptr),
                                            self.n, int(ld or self.n)), "ifs_matvec_batch_device")

    def calculate_vector_b_batch(self, X):
        """b_r = (I + T A) x_r for the rows x_r of X [nrhs, n] in one batched pass."""
        X = np.ascontiguousarray(np.atleast_2d(X), dtype=np.complex128)
        if X.shape[1] != self.n:
            raise ValueError(f"X has shape {X.shape}, expected (nrhs, {self.n})")
        B = np.empty_like(X)
        check(lib().ifs_matvec_batch(self._h, X.shape[0], X.ctypes.data, B.ctypes.data, self.n), "ifs_matvec_batch")
        return B

    def solve_batch(self, RHS, restart=50, max_iter=1000, tol=1e-6, raise_on_noconv=True, precision="fp64",
                    verbose=False):
        """Lock-step GMRES for the rows of RHS [nrhs, n]; returns (X, [info dicts])."""
        Bm = np.ascontiguousarray(np.atleast_2d(RHS), dtype=np.complex128)
        nr = Bm.shape[0]
        X = np.zeros_like(Bm)
        op = _lib.SolveOpts(int(restart), int(max_iter), float(tol), 0, int(bool(verbose)), _prec(precision))
        infos = (_lib.SolveInfo * nr)()
        rc = lib().ifs_solve_batch(self._h, nr, Bm.ctypes.data, X.ctypes.data, self.n, ctypes.byref(op), infos)
        if rc not in (0, -5, -6) or (rc != 0 and raise_on_noconv):
            check(rc, "ifs_solve_batch")
        return X, [dict(iterations=i.iterations, converged=bool(i.converged), rel_residual=i.rel_residual,
                        rel_residual_flat=i.rel_residual_flat, seconds=i.seconds) for i in infos]

    def solve_block(self, RHS, restart=30, max_iter=1000, tol=1e-6, raise_on_noconv=True, precision="fp64",
                    verbose=False):
        """Block GMRES for the rows of RHS [nrhs, n] (one shared Krylov space); returns (X, [info dicts])."""
        Bm = np.ascontiguousarray(np.atleast_2d(RHS), dtype=np.complex128)
        nr = Bm.shape[0]
        X = np.zeros_like(Bm)
        op = _lib.SolveOpts(int(restart), int(max_iter), float(tol), 0, int(bool(verbose)), _prec(precision))
        infos = (_lib.SolveInfo * nr)()
        rc = lib().ifs_solve_block(self._h, nr, Bm.ctypes.data, X.ctypes.data, self.n, ctypes.byref(op), infos)
        if rc not in (0, -5, -6) or (rc != 0 and raise_on_noconv):
            check(rc, "ifs_solve_block")
        return X, [dict(iterations=i.iterations, converged=bool(i.converged), rel_residual=i.rel_residual,
                        rel_residual_flat=i.rel_residual_flat, seconds=i.seconds) for i in infos]

    def solve_device(self, rhs_ptr, x_ptr, restart=50, max_iter=1000, tol=1e-6, use_x0=False, verbose=False,
                     precision="fp64", graph=False):
        op = _lib.SolveOpts(int(restart), int(max_iter), float(tol), int(bool(use_x0)), int(bool(verbose)),
                            _prec(precision), _lib.IFS_SOLVE_GRAPH if graph else 0)
        inf = _lib.SolveInfo()
        rc = lib().ifs_solve_device(self._h, ctypes.c_void_p(rhs_ptr), ctypes.c_void_p(x_ptr), self.n,
                                    ctypes.byref(op), ctypes.byref(inf))
        if rc not in (0, -5, -6):
            check(rc, "ifs_solve_device")
        return dict(iterations=inf.iterations, converged=bool(inf.converged), rel_residual=inf.rel_residual,
                    rel_residual_flat=inf.rel_residual_flat, seconds=inf.seconds, status=rc)

    def set_precision(self, precision):
        """'fp64' or 'mixed' for this context's matvecs from now on."""
        check(lib().ifs_set_precision(self._h, _prec(precision)), "ifs_set_precision")

    # -- instrumentation ------------------------------------------------------
    def set_stream(self, stream_ptr):
        check(lib().ifs_set_stream(self._h, ctypes.c_void_p(stream_ptr)), "ifs_set_stream")

    def stream(self):
        return lib().ifs_get_stream(self._h)

    def profile(self, enable):
        check(lib().ifs_profile_enable(self._h, int(bool(enable))), "ifs_profile_enable")

    def profile_read(self, cap=64):
        """{kernel: (total ms, launches, algorithmic flops)} since profile(True)."""
        names = (ctypes.c_char_p * cap)()
        ms = (ctypes.c_double * cap)()
        cnt = (ctypes.c_int64 * cap)()
        fl = (ctypes.c_double * cap)()
        n = lib().ifs_profile_read(self._h, ctypes.cast(names, ctypes.c_void_p), ctypes.cast(ms, ctypes.c_void_p),
                                   ctypes.cast(cnt, ctypes.c_void_p), ctypes.cast(fl, ctypes.c_void_p), cap)
        if n < 0:
            check(n, "ifs_profile_read")
        return {names[i].decode(): (ms[i], cnt[i], fl[i]) for i in range(n)}

    def matvec_cost(self):
        f = ctypes.c_double()
        b = ctypes.c_double()
        check(lib().ifs_matvec_cost(self._h, ctypes.byref(f), ctypes.byref(b)), "ifs_matvec_cost")
        return f.value, b.value

    def set_class_tiles(self, fwd_rows=0, back_rows=0, autotune=True):
        """Rotation block heights of the box translations (64, 32, or 0 = autotuned)."""
        check(lib().ifs_set_class_tiles(self._h, int(fwd_rows), int(back_rows), int(bool(autotune))),
              "ifs_set_class_tiles")

    def class_tile_stats(self):
        """(64-wide, 32-wide remainder) coaxial-shift launches of the box translations so far."""
        t = (ctypes.c_int64 * 2)()
        check(lib().ifs_class_tile_stats(self._h, ctypes.cast(t, ctypes.c_void_p)), "ifs_class_tile_stats")
        return int(t[0]), int(t[1])

    def fmm_stats(self):
        counts = (ctypes.c_int64 * 6)()
        lv = (ctypes.c_int32 * 32)()
        check(lib().ifs_fmm_stats(self._h, ctypes.cast(counts, ctypes.c_void_p), ctypes.cast(lv, ctypes.c_void_p)),
              "ifs_fmm_stats")
        kinds = ["P2M", "M2M", "M2L", "L2L", "L2P", "P2P"]
        return {k: counts[i] for i, k in enumerate(kinds)}, {l: lv[l] for l in range(32) if lv[l]}

    def tree_export(self):
        u = np.empty(self.n_spheres, dtype=np.int64)
        lv = np.empty(2, dtype=np.int32)
        check(lib().ifs_tree_export(self._h, u.ctypes.data, lv.ctypes.data), "ifs_tree_export")
        return u, int(lv[0]), int(lv[1])

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().ifs_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ClusterBatch(MultiSphereSolver):
    """Many independent clusters in one context (ifs_create_multi): the operator is
    block diagonal, one matvec launch covers every cluster's sphere pairs, and
    ``solve`` runs GMRES per cluster in lock step (ifs_solve_multi).  ``clusters`` is a
    list of dicts with centres, radii, m_rel and lmax (cluster.make_cluster's form);
    one wavenumber k for all.  Vectors are the clusters' flat vectors one after
    another (``segment``)."""

    def __init__(self, clusters, k, n_global=None, z_selector=3, workspace_mb=0, cache_mb=0, precision="fp64"):
        keep, cls = [], []
        for c in clusters:
            centres = np.ascontiguousarray(c["centres"], dtype=np.float64).reshape(-1, 3)
            radii = np.ascontiguousarray(c["radii"], dtype=np.float64)
            ns = radii.shape[0]
            m_rel = np.ascontiguousarray(np.broadcast_to(np.asarray(c["m_rel"], dtype=np.complex128), (ns,)))
            lmax = c.get("lmax")
            if lmax is None:
                x = k * radii
                lmax = np.ceil(x + 4.05 * np.cbrt(x) + 2.0)
            lmax = np.ascontiguousarray(lmax, dtype=np.int32)
            keep.append((centres, radii, m_rel, lmax))
            cls.append(_lib.Cluster(ns, centres.ctypes.data, radii.ctypes.data, m_rel.ctypes.data, lmax.ctypes.data))
        all_l = np.concatenate([kp[3] for kp in keep])
        if n_global is None:
            n_global = int(all_l.max())
        self._keep = keep
        arr = (_lib.Cluster * len(cls))(*cls)
        pr = _lib.Params()
        pr.k = float(k)
        pr.n_global = int(n_global)
        pr.z_selector = int(z_selector)
        pr.workspace_mb = int(workspace_mb)
        pr.cache_mb = int(cache_mb)
        pr.precision = _prec(precision)
        h = ctypes.c_void_p()
        check(lib().ifs_create_multi(ctypes.byref(h), len(cls), arr, ctypes.byref(pr)), "ifs_create_multi")
        self._h = h
        self.n_spheres = int(all_l.shape[0])
        self.n_global = int(n_global)
        self.lmax = all_l
        self.k = float(k)
        self.n = int(lib().ifs_vector_length(h))
        self.n_clusters = int(lib().ifs_cluster_count(h))
        self.first_sphere = np.empty(self.n_clusters + 1, dtype=np.int64)
        self.first_entry = np.empty(self.n_clusters + 1, dtype=np.int64)
        check(lib().ifs_cluster_offsets(h, self.first_sphere.ctypes.data, self.first_entry.ctypes.data),
              "ifs_cluster_offsets")

    def segment(self, v, g):
        """Cluster g's entries of a flat vector of this context."""
        return v[self.first_entry[g]:self.first_entry[g + 1]]

    def solve(self, rhs=None, restart=50, max_iter=1000, tol=1e-6, x0=None, raise_on_noconv=True, precision="fp64",
              verbose=False, graph=False):
        """Every cluster by restarted GMRES in lock step; returns (x, [info per cluster])."""
        b = self.get_vector_b() if rhs is None else np.ascontiguousarray(rhs, dtype=np.complex128)
        x = np.zeros(self.n, dtype=np.complex128) if x0 is None else np.array(x0, dtype=np.complex128)
        op = _lib.SolveOpts(int(restart), int(max_iter), float(tol), int(x0 is not None), int(bool(verbose)),
                            _prec(precision), _lib.IFS_SOLVE_GRAPH if graph else 0)
        infos = (_lib.SolveInfo * self.n_clusters)()
        rc = lib().ifs_solve_multi(self._h, b.ctypes.data, x.ctypes.data, self.n, ctypes.byref(op), infos)
        if rc not in (0, -5, -6) or (rc != 0 and raise_on_noconv):
            check(rc, "ifs_solve_multi")
        return x, [dict(iterations=i.iterations, converged=bool(i.converged), rel_residual=i.rel_residual,
                        rel_residual_flat=i.rel_residual_flat, seconds=i.seconds) for i in infos]

    def cross_sections_cluster(self, x, p, g):
        """C_sca, C_ext, C_abs of cluster g alone (the other clusters' entries zeroed)."""
        xs = np.zeros_like(x)
        ps = np.zeros_like(p)
        sl = slice(self.first_entry[g], self.first_entry[g + 1])
        xs[sl] = x[sl]
        ps[sl] = p[sl]
        return self.cross_sections(xs, ps)
