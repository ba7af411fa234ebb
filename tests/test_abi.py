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
