"""The 2000-sphere benchmark cluster at full size (lognormal radii, median 1 um,
sigma_g 1.5, [0.2, 3] um, volume fraction 1e-3, 550 nm; leaf <= 32, d0 = 6):

  * one target sphere's rows of the GPU direct matvec against the oracle's Gaunt/3j
    evaluation over all 1999 sources (lib_mie_ms_solver_interface.f90:806-998), 1e-10;
  * the MLFMM matvec against the direct one (normalised basis; FMM truncation d0 = 6);
  * GMRES(50) to 1e-6 with the MLFMM in FP64 and in mixed precision, and with the
    direct matvec in FP64; C_sca, C_ext, C_abs and the far field on the 181 x 36
    (theta, phi) grid of the MLFMM solutions against the direct FP64 solution, 1e-5
    (lib_mie_ss_helper_functions.f90:1380-1475)."""
import math

import numpy as np
import pytest

from helpers import per_sphere_rel, snm_flat
from itofoggysim_amd import MultiSphereSolver, config_cluster

pytestmark = pytest.mark.gpu


def _far_field_grid(s, x):
    th = np.repeat(np.linspace(0.0, math.pi, 181), 36)
    ph = np.tile(np.linspace(0.0, 2 * math.pi, 36, endpoint=False), 181)
    ft, fp = s.far_field(x, th, ph)
    return np.concatenate([ft, fp])


def test_config2_full_size(oracle, progress):
    c = config_cluster(2)
    lm = c["lmax"]
    # --- MLFMM: matvec, FP64 and mixed solves
    f = MultiSphereSolver(c["centres"], c["radii"], c["m_rel"], c["k"], lmax=lm, use_fmm=True, leaf_size=32,
                          fmm_d0=6.0)
    N = f.n_global
    S, mask = snm_flat(N, lm)
    rng = np.random.default_rng(2000)
    xh = (rng.standard_normal(f.n) + 1j * rng.standard_normal(f.n)) * mask
    x = xh / S
    bf = f.calculate_vector_b(x) * S
    progress("MLFMM context built, first matvec done (kept tables generated)")
    rhs = f.get_vector_b()
    p = f.plane_wave(0.0, 0.0)
    xf64, if64 = f.solve(rhs, restart=50, tol=1e-6, x0=rhs)
    progress(f"MLFMM FP64 GMRES(50): {if64['iterations']} products")
    xmix, imix = f.solve(rhs, restart=50, tol=1e-6, x0=rhs, precision="mixed")
    progress(f"MLFMM mixed GMRES(50): {imix['iterations']} products")
    print("MLFMM FP64 solve", if64, "\nMLFMM mixed solve", imix)
    assert if64["converged"] and imix["converged"]
    assert if64["rel_residual"] < 1e-6 and imix["rel_residual"] < 1e-6
    cs_f64 = np.array(f.cross_sections(xf64, p))
    cs_mix = np.array(f.cross_sections(xmix, p))
    ff_f64 = _far_field_grid(f, xf64)
    ff_mix = _far_field_grid(f, xmix)
    f.close()
    progress("MLFMM cross sections and far fields done")
    # --- direct: matvec, one target's rows against the oracle, FP64 solve
    d = MultiSphereSolver(c["centres"], c["radii"], c["m_rel"], c["k"], lmax=lm)
    bd = d.calculate_vector_b(x) * S
    progress("direct context built, first matvec done")
    err_fmm = np.linalg.norm(bf - bd) / np.linalg.norm(bd - xh)
    print(f"MLFMM vs direct matvec (normalised basis): {err_fmm:.2e}")
    assert err_fmm < 1e-8
    j = int(np.argmin(lm))
    an, bn = d.mie_coefficients()
    src = [i for i in range(len(lm)) if i != j]
    ref = (x + oracle.matvec_pairs(-c["centres"], lm, N, an, bn, c["k"], x, [j] * len(src), src)) * S
    e_or = per_sphere_rel(bd, ref, xh, len(lm), N)[j]
    progress(f"oracle rows of sphere {j} over {len(src)} sources done")
    print(f"direct rows of sphere {j} (L = {lm[j]}) vs oracle over {len(src)} sources: {e_or:.2e}")
    assert e_or < 1e-10
    xd, idd = d.solve(rhs, restart=50, tol=1e-6, x0=rhs)
    progress(f"direct FP64 GMRES(50): {idd['iterations']} products")
    print("direct FP64 solve", idd)
    assert idd["converged"]
    cs_d = np.array(d.cross_sections(xd, p))
    ff_d = _far_field_grid(d, xd)
    d.close()
    # --- results of the MLFMM solutions against the direct FP64 solution
    for name, cs, ff in (("MLFMM FP64", cs_f64, ff_f64), ("MLFMM mixed", cs_mix, ff_mix)):
        ecs = np.max(np.abs(cs - cs_d) / np.abs(cs_d))
        eff = np.max(np.abs(ff - ff_d)) / np.max(np.abs(ff_d))
        print(f"{name} vs direct FP64: C_sca/C_ext/C_abs {ecs:.2e}, far field (181 x 36) {eff:.2e}")
        assert ecs < 1e-5 and eff < 1e-5
