"""HEALPix keys of the configured workloads do not depend on the math library.

ang2pix's only library calls are cos(theta) and, within 0.01 rad of a pole,
sin(theta) (HCGrid/healpix.cu:208-224); the rest is IEEE arithmetic, which
the oracle and the device evaluate operation for operation without FMA
contraction.  The reference computed its keys with glibc's cos on an x86-64
host (HCGrid/gridding.cu:77-82); the device uses the HIP math library's.
Both are within 1-2 ulp of the correctly rounded value, so they are within 3
ulp of this container's glibc.  The oracle re-evaluates every sample's key
with each cos (and sin) result within +-3 ulp of this libm's: a sample that
is not flagged has the same key under any libm in that band -- glibc,
CUDA's or the device's.  Every sample of the configured workloads (250k and
20M single-channel rasters, the 2M-sample cube raster, the 64 batched jobs)
is unflagged, so their keys are the reference's whichever libm computes them.

The positive control bisects the latitude of one sample down to adjacent
doubles on either side of a pixel boundary: those samples must be flagged.
"""
import numpy as np
import pytest

import oracle as O
from hcgrid_amd import synth

ULPS = 3


def _boundary_pair(h, lon, lo, hi):
    klo = O.sample_keys(h, np.array([lon]), np.array([lo]))[0]
    assert O.sample_keys(h, np.array([lon]), np.array([hi]))[0] != klo
    while True:
        mid = 0.5 * (lo + hi)
        if mid in (lo, hi):
            return lo, hi
        if O.sample_keys(h, np.array([lon]), np.array([mid]))[0] == klo:
            lo = mid
        else:
            hi = mid


# (lon, lat interval, nside, ulps): the belt and southern belt, where one
# step of the latitude moves cos(theta) by about one ulp, and the polar cap
# with the sin branch (theta < 0.01 rad), where one step of the latitude
# moves sin(theta) by ~35 ulp, so the band there must be wider to reach it
@pytest.mark.parametrize("lon,lo,hi,nside,ulps", [(180.123456, 30.0, 30.01, 8192, ULPS),
                                                  (12.3, -45.0, -44.99, 4096, ULPS),
                                                  (77.7, 89.6, 89.61, 8192, 64)])
def test_positive_control_boundary_is_flagged(lon, lo, hi, nside, ulps):
    h = O.hpx(nside)
    a, b = _boundary_pair(h, lon, lo, hi)
    assert np.nextafter(a, b) == b
    cnt, flags = O.keys_libm_sensitive(h, np.array([lon, lon]), np.array([a, b]), ulps)
    assert cnt >= 1 and flags.max() == 1
    # with no slack at all nothing can change
    assert O.keys_libm_sensitive(h, np.array([lon, lon]), np.array([a, b]), 0)[0] == 0


def _check(h, lon, lat):
    cnt, _ = O.keys_libm_sensitive(h, lon, lat, ULPS)
    print(f"n={lon.size} nside={h.nside} libm-sensitive keys at +-{ULPS} ulp: {cnt}")
    assert cnt == 0


def test_raster_250k_and_20m_keys_libm_independent():
    h = O.hpx(8192)
    for wl in (synth.raster_250k(), synth.raster_20m()):
        _check(h, wl.lon, wl.lat)


def test_cube_raster_keys_libm_independent():
    lon, lat, _ = synth.cube_coords()
    _check(O.hpx(8192), lon, lat)


def test_batched_jobs_keys_libm_independent():
    # the 64-job mix (sigma 0.015/0.02 -> nside 8192, 0.03 -> 4096); the
    # served mix draws from the same distribution
    for sp in synth.batch_specs(64):
        nside = O.set_optimal_nside(np.radians(sp.hpx_res))
        _check(O.hpx(nside), sp.lon, sp.lat)
