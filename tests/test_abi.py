"""The C-ABI library loads and exports every symbol include/ifs.h declares (no
device work here)."""
import os
import re

from itofoggysim_amd import _lib

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "ifs.h")


def declared():
    txt = open(HDR).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|void\s*\*)\s+(ifs_\w+)\s*\(", txt, re.M)))


def test_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
        assert n in _lib.SIGNATURES, n


def test_errors_without_arguments():
    L = _lib.lib()
    assert L.ifs_create(None, None, None) == -1
    assert L.ifs_matvec(None, None, None, 0) == -1
    assert L.ifs_vector_length(None) == -1


def test_struct_layouts_match_header(tmp_path):
    """sizeof/offsetof of the ABI structs as a C compiler lays them out (the layout a
    Fortran bind(C) type or a ctypes Structure must reproduce) equal the ctypes
    definitions the Python wrapper uses."""
    import ctypes
    import subprocess
    src = tmp_path / "layout.c"
    fields = {"ifs_params": [f for f, _ in _lib.Params._fields_],
              "ifs_cluster": [f for f, _ in _lib.Cluster._fields_],
              "ifs_solve_opts": [f for f, _ in _lib.SolveOpts._fields_],
              "ifs_solve_info": [f for f, _ in _lib.SolveInfo._fields_]}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HDR}"', "int main(void) {"]
    for t, fs in fields.items():
        lines.append(f'printf("{t} %zu\\n", sizeof({t}));')
        for f in fs:
            lines.append(f'printf("{t}.{f} %zu\\n", offsetof({t}, {f}));')
    lines.append("return 0; }")
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                          text=True).stdout.splitlines())
    for t, cls in (("ifs_params", _lib.Params), ("ifs_cluster", _lib.Cluster), ("ifs_solve_opts", _lib.SolveOpts),
                   ("ifs_solve_info", _lib.SolveInfo)):
        assert int(out[t]) == ctypes.sizeof(cls), t
        for f in fields[t]:
            assert int(out[f"{t}.{f}"]) == getattr(cls, f).offset, (t, f)


def test_amplitude_and_mueller_host_functions():
    """ifs_amplitude_matrix / ifs_mueller are host-only: checked against the oracle here."""
    import numpy as np
    from itofoggysim_amd import MultiSphereSolver
    from oracle import farfield as FF
    rng = np.random.default_rng(0)
    nd = 7
    phi = rng.uniform(0, 6.3, nd)
    F = [rng.standard_normal(nd) + 1j * rng.standard_normal(nd) for _ in range(4)]
    S = MultiSphereSolver.amplitude_matrix(phi, (F[0], F[1]), (F[2], F[3]))
    So = FF.amplitude_matrix((F[0], F[1]), (F[2], F[3]), phi)
    for a, b in zip(S, So):
        assert np.allclose(a, b, rtol=1e-14, atol=1e-14)
    M = MultiSphereSolver.mueller(*S)
    assert np.allclose(M, FF.mueller(*So), rtol=1e-13, atol=1e-13)
