"""A Fortran caller of the C ABI (tests/fortran/ifs_caller.f90), compiled with amdflang:
the bind(C) types match the C layout, argument errors return IFS_E_ARG, and the
host-only octree calls agree with the same calls through ctypes.  (Its GPU mode, one
matvec through a calc_vector_b-shaped wrapper, runs as a program of its own on a GPU.)"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from itofoggysim_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)


@pytest.fixture(scope="module")
def caller(tmp_path_factory):
    if FLANG is None:
        pytest.skip("amdflang not available")
    d = tmp_path_factory.mktemp("fortran")
    exe = d / "ifs_caller"
    libdir = os.path.join(ROOT, "itofoggysim_amd")
    subprocess.run([FLANG, "-O1", "-o", str(exe), os.path.join(ROOT, "tests", "fortran", "ifs_caller.f90"),
                    "-L" + libdir, "-lifs", "-Wl,-rpath," + libdir], check=True, cwd=str(d))
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, cwd=str(d)).stdout
    lines = [l.strip() for l in out.splitlines() if l.strip()]

    def field(prefix):
        hit = [l[len(prefix):].split() for l in lines if l.startswith(prefix)]
        assert hit, (prefix, lines)
        return hit[0]
    return field


def test_fortran_struct_sizes(caller):
    sizes = [int(v) for v in caller("sizes")]
    want = [ctypes.sizeof(t) for t in (_lib.Cluster, _lib.Params, _lib.SolveOpts, _lib.SolveInfo)]
    assert sizes == want


def test_fortran_argument_errors(caller):
    assert caller("ifs_create(n_spheres = 0)") == ["-1"]
    assert caller("ifs_vector_length(null)") == ["-1"]
    assert caller("ifs_matvec(null)") == ["-1"]


def test_fortran_octree_calls_match_ctypes(caller):
    rc, box = caller("universal_index level 1")
    assert (int(rc), int(box)) == (0, 6)  # reference fixture, level 1
    L = _lib.lib()
    out = np.zeros(189, dtype=np.int64)
    n = L.ifs_tree_domain_i4(35, 2, out.ctypes.data, 189)
    assert int(caller("i4(35, level 2) count")[0]) == n
