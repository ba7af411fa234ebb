"""ifs_far_field_device_noalloc (the far field without any allocation: its tables go
to the library's fixed per-GPU scratch array) against ifs_far_field_device, through
the C ABI, on two calls in a row, the second with a smaller input; and the size
limit."""
import ctypes
import math

import numpy as np
import pytest

from helpers import Hip
from itofoggysim_amd import MultiSphereSolver, config_cluster
from itofoggysim_amd._lib import lib

pytestmark = pytest.mark.gpu

SCRATCH_ENTRIES = 1 << 21  # IFS_FF_SCRATCH_ENTRIES


def _check(hip, s, ndir, seed):
    rng = np.random.default_rng(seed)
    x = hip.upload((rng.standard_normal(s.n) + 1j * rng.standard_normal(s.n)) * 1e-3)
    theta = hip.upload(rng.uniform(0.0, math.pi, ndir))
    phi = hip.upload(rng.uniform(0.0, 2 * math.pi, ndir))
    nan = np.full(2 * ndir, complex(np.nan, np.nan))
    ref = hip.upload(nan)
    out = hip.upload(nan)
    assert lib().ifs_far_field_device(s._h, x, s.n, ndir, theta, phi, ref) == 0
    assert lib().ifs_far_field_device_noalloc(s._h, x, s.n, ndir, theta, phi, out) == 0
    hip.sync(ctypes.c_void_p(s.stream()))  # both enqueue on the context stream
    want = hip.download(ref, nan)
    got = hip.download(out, nan)
    assert np.all(np.isfinite(got)) and np.all(np.isfinite(want))
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    print(f"ns={s.n_spheres} N={s.n_global} ndir={ndir}: noalloc vs allocating far field {err:.2e}")
    assert err < 1e-12


def test_far_field_noalloc_matches_allocating_entry_point():
    hip = Hip()
    c = config_cluster(1)
    big = MultiSphereSolver(c["centres"], c["radii"], c["m_rel"], c["k"], lmax=c["lmax"])
    per_dir = big.n_global * (big.n_global + 2)
    _check(hip, big, SCRATCH_ENTRIES // per_dir, 1)  # fills the scratch array
    # second call, smaller input on another context: the array still holds the first call's tables
    small = MultiSphereSolver(c["centres"][:5], c["radii"][:5], c["m_rel"][:5], c["k"],
                              lmax=np.minimum(c["lmax"][:5], 9))
    _check(hip, small, 37, 2)
    # beyond the limit: IFS_E_ARG, nothing runs
    n_over = SCRATCH_ENTRIES // per_dir + 1
    x = hip.alloc(16 * big.n)
    th = hip.alloc(8 * n_over)
    out = hip.alloc(32 * n_over)
    assert lib().ifs_far_field_device_noalloc(big._h, x, big.n, n_over, th, th, out) == -1
    hip.free_all()
    big.close()
    small.close()
