"""The 20000-sphere cluster of the mixed-precision workload at 1550 nm: radii uniform in
[0.3, 1] um at the 20-sphere cluster's number density (a 100 um cube), m = 1.318 +
1e-4 i, orders 8..13, 7.8 M unknowns; octree leaf <= 32, d0 = 6:

  * the FP64 MLFMM matvec's rows of two target spheres against the oracle's Gaunt/3j
    sum over all 19999 sources (lib_mie_ms_solver_interface.f90:806-998), evaluated on
    the host while the GPU solves: at d0 = 6 within the truncation's ~1e-6 (box orders
    L = kD + 1.8 d0^(2/3) (kD)^(1/3); measured 7.3e-7 and 4.8e-7), and at d0 = 10
    25-500x closer (1.4e-9 and 1.9e-8), so the difference is the truncation and shrinks
    with it (the MLFMM itself equals the oracle's MLFMM restatement to 1e-10 on small
    trees, test_gpu_fmm.py);
  * the mixed matvec (FP32 matrix-core box translations) against the FP64 one;
  * GMRES(30) to 1e-6 in FP64 and in mixed precision (FP64 true residual per restart
    cycle): coefficients, C_sca/C_ext/C_abs and the far field on the 181 x 36 grid of
    the mixed solution against the FP64 one (lib_mie_ss_helper_functions.f90:1380-1475)."""
import math
import threading

import numpy as np
import pytest

from helpers import per_sphere_rel, snm_flat
from itofoggysim_amd import MultiSphereSolver, config_cluster

pytestmark = pytest.mark.gpu


def test_config3_1550nm(oracle, progress):
    c = config_cluster(3, wavelength=1.55)
    lm = c["lmax"]
    ns = len(lm)
    s = MultiSphereSolver(c["centres"], c["radii"], c["m_rel"], c["k"], lmax=lm, use_fmm=True, leaf_size=32,
                          fmm_d0=6.0)
    try:
        N = s.n_global
        S, mask = snm_flat(N, lm)
        rng = np.random.default_rng(3)
        xh = (rng.standard_normal(s.n) + 1j * rng.standard_normal(s.n)) * mask
        x = xh / S
        an, bn = s.mie_coefficients()
        # oracle rows of the smallest-order sphere and of a median-order one, on the host
        order = np.argsort(lm, kind="stable")
        targets = [int(order[0]), int(order[ns // 2])]
        ref = {}

        def host_oracle():
            for j in targets:
                src = [i for i in range(ns) if i != j]
                ref[j] = oracle.matvec_pairs(-c["centres"], lm, N, an, bn, c["k"], x, [j] * len(src), src)

        th = threading.Thread(target=host_oracle)
        th.start()
        b64 = s.calculate_vector_b(x) * S
        s.set_precision("mixed")
        b32 = s.calculate_vector_b(x) * S
        s.set_precision("fp64")
        progress("context built; FP64 and mixed matvecs done (oracle rows running on the host)")
        emix = np.linalg.norm(b32 - b64) / np.linalg.norm(b64 - xh)
        print(f"mixed vs FP64 matvec (normalised basis): {emix:.2e}", flush=True)
        rhs = s.get_vector_b()
        x64, i64 = s.solve(rhs, restart=30, tol=1e-6, x0=rhs, max_iter=400)
        progress(f"FP64 GMRES(30): {i64['iterations']} products")
        print("FP64 GMRES(30)", i64, flush=True)
        x32, i32 = s.solve(rhs, restart=30, tol=1e-6, x0=rhs, max_iter=400, precision="mixed")
        progress(f"mixed GMRES(30): {i32['iterations']} products")
        print("mixed GMRES(30)", i32, flush=True)
        p = s.plane_wave(0.0, 0.0)
        cs64 = np.array(s.cross_sections(x64, p))
        cs32 = np.array(s.cross_sections(x32, p))
        thg = np.repeat(np.linspace(0.0, math.pi, 181), 36)
        phg = np.tile(np.linspace(0.0, 2 * math.pi, 36, endpoint=False), 181)
        f64 = np.concatenate(s.far_field(x64, thg, phg))
        f32 = np.concatenate(s.far_field(x32, thg, phg))
        s.close()
        s = MultiSphereSolver(c["centres"], c["radii"], c["m_rel"], c["k"], lmax=lm, use_fmm=True, leaf_size=32,
                              fmm_d0=10.0)
        b10 = s.calculate_vector_b(x) * S
        progress("d0 = 10 context built and matvec done; waiting for the oracle rows")
        th.join()
        progress("oracle rows done")
    finally:
        s.close()
    for j in targets:
        e6 = per_sphere_rel(b64, (x + ref[j]) * S, xh, ns, N)[j]
        e10 = per_sphere_rel(b10, (x + ref[j]) * S, xh, ns, N)[j]
        print(f"MLFMM rows of sphere {j} (L = {lm[j]}) vs oracle over {ns - 1} sources: d0 = 6 {e6:.2e}, "
              f"d0 = 10 {e10:.2e}")
        assert e6 < 1e-5 and e10 < 1e-7 and e10 < e6 / 10
    assert 1e-10 < emix < 1e-5
    assert i64["converged"] and i32["converged"]
    assert i64["rel_residual"] < 1e-6 and i32["rel_residual"] < 1e-6
    ecoef = np.linalg.norm((x32 - x64) * S) / np.linalg.norm(x64 * S)
    ecs = np.max(np.abs(cs32 - cs64) / np.abs(cs64))
    eff = np.max(np.abs(f32 - f64)) / np.max(np.abs(f64))
    print(f"mixed vs FP64 solution: coefficients {ecoef:.2e}, C_sca/C_ext/C_abs {ecs:.2e}, far field {eff:.2e}; "
          f"FP64 C_sca, C_ext, C_abs = {cs64}")
    assert ecoef < 1e-5 and ecs < 1e-5 and eff < 1e-5
