"""Synthetic sphere clusters for the benchmark configurations.

Draw order (numpy ``Generator(PCG64(seed))``): radii, then centres uniform in the
cube redrawn on overlap (surface gap >= 0.05 um), then per-sphere truncation
``L_j = ceil(x + 4 x^(1/3) + 2)`` with ``x = 2 pi n_medium r / lambda``.
Lengths are in micrometres.  Configuration c uses seeds ``1000 c + {0, 1, 2}``.
"""
import math

import numpy as np

WATER_550 = complex(1.333, 1e-9)


def _radii(rng, n, dist):
    kind = dist[0]
    if kind == "uniform":
        lo, hi = dist[1], dist[2]
        return rng.uniform(lo, hi, n)
    if kind == "lognormal":
        median, sigma_g, lo, hi = dist[1:5]
        out = np.empty(n)
        filled = 0
        mu, s = math.log(median), math.log(sigma_g)
        while filled < n:
            draw = rng.lognormal(mu, s, 2 * (n - filled) + 16)
            draw = draw[(draw >= lo) & (draw <= hi)]
            take = min(n - filled, draw.size)
            out[filled:filled + take] = draw[:take]
            filled += take
        return out
    raise ValueError(dist)


def _centres(rng, radii, side, gap=0.05, max_tries=100000):
    n = radii.size
    c = np.empty((n, 3))
    for i in range(n):
        for _ in range(max_tries):
            p = rng.uniform(0.0, side, 3)
            if i == 0:
                break
            d = np.sqrt(((c[:i] - p) ** 2).sum(axis=1))
            if np.all(d >= radii[:i] + radii[i] + gap):
                break
        else:
            raise RuntimeError("could not place sphere %d without overlap" % i)
        c[i] = p
    return c


def lmax_rule(radii, wavelength, n_medium=1.0):
    x = 2.0 * math.pi * n_medium * np.asarray(radii) / wavelength
    return np.ceil(x + 4.0 * np.cbrt(x) + 2.0).astype(np.int32)


def make_cluster(n, radius_dist, side=None, volume_fraction=None, wavelength=0.55, m=WATER_550, n_medium=1.0,
                 seed=0, gap=0.05):
    """Returns a dict describing one cluster (centres, radii, m_rel, lmax, k, wavelength, side)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    radii = _radii(rng, n, radius_dist)
    if side is None:
        vol = (4.0 / 3.0) * math.pi * float(np.sum(radii ** 3))
        side = (vol / volume_fraction) ** (1.0 / 3.0)
    centres = _centres(rng, radii, side, gap)
    lmax = lmax_rule(radii, wavelength, n_medium)
    return dict(centres=centres, radii=radii, m_rel=np.full(n, m, dtype=np.complex128), lmax=lmax,
                k=2.0 * math.pi * n_medium / wavelength, wavelength=wavelength, side=side, seed=seed)


WATER_1550 = complex(1.318, 1e-4)


def config_cluster(config, replica=0, n=None, wavelength=None):
    """The benchmark clusters: 1 = 20 spheres U[0.3,1] um in a 10 um cube;
    2 = 2000 spheres lognormal(1 um, 1.5) clipped to [0.2, 3] um at volume fraction 1e-3;
    3 = the generator of 1 scaled to 20000 spheres at the same number density (a cube of
    10 (N/20)^(1/3) um, 100 um for 20000), at 550 nm (m = 1.333 + 1e-9 i) or at
    ``wavelength=1.55`` (m = 1.318 + 1e-4 i).  ``n`` overrides the sphere count (same
    generator)."""
    seed = 1000 * config + replica
    if config == 1:
        return make_cluster(n or 20, ("uniform", 0.3, 1.0), side=10.0, seed=seed)
    if config == 2:
        return make_cluster(n or 2000, ("lognormal", 1.0, 1.5, 0.2, 3.0), volume_fraction=1e-3, seed=seed)
    if config == 3:
        n = n or 20000
        lam = wavelength or 0.55
        m = WATER_1550 if abs(lam - 1.55) < 1e-9 else WATER_550
        return make_cluster(n, ("uniform", 0.3, 1.0), side=10.0 * (n / 20.0) ** (1.0 / 3.0), wavelength=lam, m=m,
                            seed=seed)
    raise ValueError(config)
