"""Mixed precision: FP32 matrix-core box translations (M2M, M2L, L2L) inside FP64
GMRES whose restart cycles recompute the true residual in FP64 (iterative refinement).
Every result is compared with the FP64 path on the same cluster: the matvec, the
solution coefficients, the cross sections and the far field."""
import math

import numpy as np
import pytest

from helpers import snm_flat
from itofoggysim_amd import MultiSphereSolver, config_cluster

pytestmark = pytest.mark.gpu


def _solver(c, **kw):
    return MultiSphereSolver(c["centres"], c["radii"], c["m_rel"], c["k"], lmax=c["lmax"], use_fmm=True,
                             leaf_size=32, fmm_d0=6.0, **kw)


@pytest.mark.parametrize("n", [200, 600])
def test_mixed_matvec_error(n):
    """FP32 box products: the matvec differs from the FP64 one by FP32 rounding only
    (normalised basis), far below the MLFMM truncation of d0 = 6."""
    c = config_cluster(2, n=n)
    s = _solver(c)
    S, mask = snm_flat(s.n_global, c["lmax"])
    rng = np.random.default_rng(n)
    xh = (rng.standard_normal(s.n) + 1j * rng.standard_normal(s.n)) * mask
    x = xh / S
    try:
        b64 = s.calculate_vector_b(x) * S
        s.set_precision("mixed")
        b32 = s.calculate_vector_b(x) * S
        s.set_precision("fp64")
        b64b = s.calculate_vector_b(x) * S
    finally:
        s.close()
    err = np.linalg.norm(b32 - b64) / np.linalg.norm(b64 - xh)
    print(f"n={n}: mixed vs FP64 matvec {err:.2e}")
    assert np.array_equal(b64, b64b)  # switching back restores the FP64 products bit for bit
    assert 1e-10 < err < 1e-5  # FP32 rounding is visible, and bounded


def test_mixed_solve_matches_fp64():
    """GMRES(50) to 1e-6 on a 600-sphere cluster of the benchmark generator, plain FP64
    and mixed: the mixed solve's FP64 true residual meets the tolerance, and its
    coefficients, cross sections and far field agree with the FP64 solution."""
    c = config_cluster(2, n=600)
    f64 = _solver(c)
    rhs = f64.get_vector_b()
    x64, i64 = f64.solve(rhs, restart=50, tol=1e-6, x0=rhs)
    x32, i32 = f64.solve(rhs, restart=50, tol=1e-6, x0=rhs, precision="mixed")
    print("fp64", i64, "\nmixed", i32)
    assert i64["converged"] and i32["converged"]
    assert i32["rel_residual"] < 1e-6  # the FP64 true residual
    S, _ = snm_flat(f64.n_global, c["lmax"])
    ecoef = np.linalg.norm((x32 - x64) * S) / np.linalg.norm(x64 * S)
    p = f64.plane_wave(0.0, 0.0)
    cs64 = np.array(f64.cross_sections(x64, p))
    cs32 = np.array(f64.cross_sections(x32, p))
    ecs = np.max(np.abs(cs32 - cs64) / np.abs(cs64))
    th = np.linspace(0.0, math.pi, 181)
    ph = np.full(181, 0.3)
    ft64, fp64 = f64.far_field(x64, th, ph)
    ft32, fp32 = f64.far_field(x32, th, ph)
    F64 = np.concatenate([ft64, fp64])
    eff = np.max(np.abs(np.concatenate([ft32, fp32]) - F64)) / np.max(np.abs(F64))
    print(f"mixed vs FP64: coefficients {ecoef:.2e}, C_sca/C_ext/C_abs {ecs:.2e}, far field {eff:.2e}")
    # both are solutions to 1e-6 of the same system: they differ by about the tolerance
    assert ecoef < 1e-5 and ecs < 1e-5 and eff < 1e-5
    f64.close()
