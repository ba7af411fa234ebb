"""ctypes binding of libifs.so (the HIP library built from csrc/).

The library must be built in-tree (``make -C itofoggysim_amd``).  There is no
CPU fallback: if the library is missing every entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libifs.so")

IFS_OK = 0
ERRORS = {
    -1: "IFS_E_ARG (invalid argument)",
    -2: "IFS_E_NOMEM (device allocation failed)",
    -3: "IFS_E_HIP (HIP runtime error)",
    -4: "IFS_E_NODEVICE (no GPU visible)",
    -5: "IFS_E_NOCONV (not converged)",
    -6: "IFS_E_BREAKDOWN (solver breakdown)",
}


class IfsError(RuntimeError):
    def __init__(self, code, where):
        super().__init__(f"{where}: {ERRORS.get(code, code)}")
        self.code = code


class Cluster(ctypes.Structure):
    _fields_ = [("n_spheres", ctypes.c_int32), ("centres", ctypes.c_void_p), ("radii", ctypes.c_void_p),
                ("m_rel", ctypes.c_void_p), ("lmax", ctypes.c_void_p)]


class Params(ctypes.Structure):
    _fields_ = [("k", ctypes.c_double), ("n_global", ctypes.c_int32), ("z_selector", ctypes.c_int32),
                ("ref_row_zeroing", ctypes.c_int32), ("use_fmm", ctypes.c_int32), ("leaf_size", ctypes.c_int32),
                ("fmm_d0", ctypes.c_double), ("fmm_lmax_cap", ctypes.c_int32), ("workspace_mb", ctypes.c_int32),
                ("d_convention", ctypes.c_int32), ("cache_mb", ctypes.c_int32), ("precision", ctypes.c_int32)]


class SolveOpts(ctypes.Structure):
    _fields_ = [("restart", ctypes.c_int32), ("max_iter", ctypes.c_int32), ("tol", ctypes.c_double),
                ("use_x0", ctypes.c_int32), ("verbose", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("flags", ctypes.c_int32)]


IFS_SOLVE_GRAPH = 1


class RequestInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("iterations", ctypes.c_int32), ("latency_s", ctypes.c_double),
                ("queue_s", ctypes.c_double), ("footprint_bytes", ctypes.c_int64)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("converged", ctypes.c_int32), ("rel_residual", ctypes.c_double),
                ("seconds", ctypes.c_double), ("rel_residual_flat", ctypes.c_double)]


_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32

# name -> (restype, argtypes); every symbol include/ifs.h declares
SIGNATURES = {
    "ifs_create": (_i32, [ctypes.POINTER(_vp), ctypes.POINTER(Cluster), ctypes.POINTER(Params)]),
    "ifs_destroy": (_i32, [_vp]),
    "ifs_vector_length": (_i64, [_vp]),
    "ifs_matvec": (_i32, [_vp, _vp, _vp, _i64]),
    "ifs_matvec_device": (_i32, [_vp, _vp, _vp, _i64]),
    "ifs_set_stream": (_i32, [_vp, _vp]),
    "ifs_set_precision": (_i32, [_vp, _i32]),
    "ifs_get_stream": (_vp, [_vp]),
    "ifs_get_mie": (_i32, [_vp, _vp, _vp]),
    "ifs_plane_wave_rhs": (_i32, [_vp, _vp, _i64]),
    "ifs_solve": (_i32, [_vp, _vp, _vp, _i64, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveInfo)]),
    "ifs_solve_device": (_i32, [_vp, _vp, _vp, _i64, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveInfo)]),
    "ifs_tree_export": (_i32, [_vp, _vp, _vp]),
    "ifs_tree_scale": (_i32, [_vp, _i32, _vp, _vp]),
    "ifs_tree_universal_index": (_i32, [_vp, _i32, _i32, _vp]),
    "ifs_tree_neighbours": (_i32, [_i64, _i64, _i32, _vp]),
    "ifs_tree_centre": (_i32, [_i64, _i32, _vp]),
    "ifs_tree_domain_i4": (_i32, [_i64, _i32, _vp, _i32]),
    "ifs_tree_level_max": (_i32, [_vp, _i64, _i64]),
    "ifs_fmm_stats": (_i32, [_vp, _vp, _vp]),
    "ifs_profile_enable": (_i32, [_vp, _i32]),
    "ifs_profile_read": (_i32, [_vp, _vp, _vp, _vp, _vp, _i32]),
    "ifs_matvec_cost": (_i32, [_vp, _vp, _vp]),
    "ifs_coaxial_block": (_i32, [_i32, ctypes.c_double, _i32, _i32, _i32, _i32, _i32, _vp]),
    "ifs_plane_wave": (_i32, [_vp, ctypes.c_double, ctypes.c_double, _vp, _i64]),
    "ifs_plane_wave_rhs_dir": (_i32, [_vp, ctypes.c_double, ctypes.c_double, _vp, _i64]),
    "ifs_cross_sections": (_i32, [_vp, _vp, _vp, _i64, _vp]),
    "ifs_cross_sections_device": (_i32, [_vp, _vp, _vp, _i64, _vp]),
    "ifs_far_field": (_i32, [_vp, _vp, _i64, _i32, _vp, _vp, _vp]),
    "ifs_far_field_device": (_i32, [_vp, _vp, _i64, _i32, _vp, _vp, _vp]),
    "ifs_far_field_device_noalloc": (_i32, [_vp, _vp, _i64, _i32, _vp, _vp, _vp]),
    "ifs_amplitude_matrix": (_i32, [_i32, _vp, _vp, _vp, _vp]),
    "ifs_mueller": (_i32, [_i32, _vp, _vp]),
    "ifs_matvec_batch_device": (_i32, [_vp, _i32, _vp, _vp, _i64, _i64]),
    "ifs_matvec_batch": (_i32, [_vp, _i32, _vp, _vp, _i64]),
    "ifs_solve_batch_device": (_i32, [_vp, _i32, _vp, _vp, _i64, _i64, ctypes.POINTER(SolveOpts),
                                      ctypes.POINTER(SolveInfo)]),
    "ifs_solve_batch": (_i32, [_vp, _i32, _vp, _vp, _i64, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveInfo)]),
    "ifs_server_create": (_i32, [ctypes.POINTER(_vp), _i32, ctypes.c_double]),
    "ifs_server_submit": (_i32, [_vp, ctypes.POINTER(Cluster), ctypes.POINTER(Params), _i32, _vp, _vp,
                                 ctypes.c_double, _i32, ctypes.POINTER(_i64)]),
    "ifs_server_wait": (_i32, [_vp, _i64, _vp, ctypes.POINTER(RequestInfo)]),
    "ifs_server_stats": (_i32, [_vp, _vp]),
    "ifs_server_destroy": (_i32, [_vp]),
    "ifs_server_set_verbose": (_i32, [_vp, _i32]),
    "ifs_server_plan": (_i32, [ctypes.POINTER(Cluster), ctypes.POINTER(Params), _i32, _i32, ctypes.c_double, _i32,
                               _vp]),
    "ifs_tree_build_gpu": (_i32, [_vp, _i32, _i32, _vp, _vp, _vp]),
    "ifs_matvec_graph_device": (_i32, [_vp, _vp, _vp, _i64]),
    "ifs_class_tile_stats": (_i32, [_vp, _vp]),
    "ifs_set_class_tiles": (_i32, [_vp, _i32, _i32, _i32]),
    "ifs_plane_wave_mode": (_i32, [_vp, ctypes.c_double, ctypes.c_double, _i32, _vp, _i64]),
    "ifs_plane_wave_rhs_mode": (_i32, [_vp, ctypes.c_double, ctypes.c_double, _i32, _vp, _i64]),
    "ifs_create_multi": (_i32, [ctypes.POINTER(_vp), _i32, ctypes.POINTER(Cluster), ctypes.POINTER(Params)]),
    "ifs_cluster_count": (_i32, [_vp]),
    "ifs_cluster_offsets": (_i32, [_vp, _vp, _vp]),
    "ifs_solve_multi_device": (_i32, [_vp, _vp, _vp, _i64, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveInfo)]),
    "ifs_solve_block_device": (_i32, [_vp, _i32, _vp, _vp, _i64, _i64, ctypes.POINTER(SolveOpts),
                                      ctypes.POINTER(SolveInfo)]),
    "ifs_solve_block": (_i32, [_vp, _i32, _vp, _vp, _i64, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveInfo)]),
    "ifs_solve_multi": (_i32, [_vp, _vp, _vp, _i64, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveInfo)]),
    "ifs_tree_i4_gpu": (_i32, [_vp, _i64, _i32, _vp, _vp]),
}

_lib = None


def lib():
    """Loads libifs.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C itofoggysim_amd` "
                               "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, where):
    if rc != IFS_OK:
        raise IfsError(rc, where)
    return rc
