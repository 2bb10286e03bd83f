"""CPU: the quantized-position geometry of the FP32_BULK pair walks
(esph_common.hpp qgeom, quantum and band set in esph_api.hip build_grid).

The walks decide the FP64 neighbour rule d^2 <= R^2 from float distances of
integer position differences, in units of the quantum, unless the float d^2
lies within `band` of R^2 / qu^2.  Emulating that float arithmetic in numpy
on pairs drawn near R (and at exactly R), the float distance must stay
within the band of the binary64 rule's d^2 / qu^2 -- so no pair is decided
wrongly -- for several domain extents and radii.  The quantum is a power of
two, so the same arithmetic in metres (each value times qu or qu^2) gives
the same decisions; that is checked too."""
import math

import numpy as np


def quantum_and_band(extent, R, D=3):
    """The quantum (smallest power of two with the extent within 2^30 quanta)
    and the band, as the library derives them."""
    e = math.ceil(math.log2(extent / 2.0 ** 30))
    u = 2.0 ** e
    eps = 2.0 ** -24
    dd = 1.01 * R
    dx = u + eps * dd
    band = 2.0 * math.sqrt(D) * dd * dx + D * dx * dx + (D + 1) * eps * dd * dd + eps * R * R
    return u, 2.0 * band


def test_band_covers_quantized_float_distance():
    rng = np.random.default_rng(3)
    for extent, R in ((3.2, 0.0234), (1.0, 0.05), (100.0, 0.01), (0.5, 0.2), (20.0, 0.3)):
        u, band_m = quantum_and_band(extent + 4.0 * R, R)  # the grid extends beyond the particles
        inv_u2 = 1.0 / (u * u)
        band = np.float32(band_m * inv_u2)
        R2q = np.float32(R * R * inv_u2)
        lo = np.array([-0.013, -0.02, 0.007])
        n = 200000
        xi = lo + rng.uniform(2.0 * R, extent, (n, 3))
        dirs = rng.normal(size=(n, 3))
        dirs /= np.linalg.norm(dirs, axis=1)[:, None]
        near = rng.uniform(size=n) < 0.2
        dist = np.where(near, R * (1.0 + rng.uniform(-3e-6, 3e-6, n)), R * rng.uniform(0.3, 1.01, n))
        xj = xi + dirs * dist[:, None]
        # a quarter of the near pairs exactly at R along an axis
        k = np.where(near)[0][::4]
        xj[k] = xi[k]
        xj[k, 0] = xi[k, 0] + R
        d2_rule = ((xi[:, 0] - xj[:, 0]) ** 2 + (xi[:, 1] - xj[:, 1]) ** 2) + (xi[:, 2] - xj[:, 2]) ** 2
        qi = np.rint((xi - lo) / u).astype(np.int64)
        qj = np.rint((xj - lo) / u).astype(np.int64)
        assert np.abs(qi).max() < 2 ** 30 and np.abs(qj).max() < 2 ** 30
        rv = (qi - qj).astype(np.float32)  # quanta
        d2q = (rv[:, 0] * rv[:, 0] + rv[:, 1] * rv[:, 1]) + rv[:, 2] * rv[:, 2]
        for d2 in (d2q, (rv.astype(np.float64) ** 2).sum(axis=1).astype(np.float32)):
            err = np.abs(d2.astype(np.float64) - d2_rule * inv_u2)
            assert err.max() < float(band), (extent, R, err.max(), float(band))
            decided = np.abs(d2 - R2q) > band
            assert np.array_equal(d2[decided] <= R2q, d2_rule[decided] <= R * R)
            assert not decided[k].any()  # pairs exactly at R go to the FP64 rule
        # in metres: every value scaled by a power of two, the same decisions
        rvm = rv * np.float32(u)
        d2m = (rvm[:, 0] * rvm[:, 0] + rvm[:, 1] * rvm[:, 1]) + rvm[:, 2] * rvm[:, 2]
        assert np.array_equal(d2m, d2q * np.float32(u * u))
        assert np.array_equal(d2m <= np.float32(R * R), d2q <= R2q)
