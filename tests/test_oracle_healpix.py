"""Pins the oracle's HEALPix restatement (CPU only).

Checks: closed forms of the geometry, nside = 1 base pixels, exhaustive
pix2ang -> ang2pix round trips at small nside, ring monotonicity, and an
independent pixelisation (tests/healpix_indep.py) on random points.
"""
import math

import numpy as np
import pytest

import oracle as O
from healpix_indep import ang2pix_ring, ring_start_closed_form


@pytest.mark.parametrize("nside", [1, 2, 4, 8, 1024, 8192])
def test_geometry_closed_forms(nside):
    h = O.hpx(nside)
    assert h.npix == 12 * nside * nside
    assert h.ncap == 2 * nside * (nside - 1)
    assert h.nrings == 4 * nside - 1
    assert h.order == int(math.log2(nside))
    assert math.isclose(h.resolution, math.sqrt(4 * math.pi / h.npix), rel_tol=1e-15)
    rings = np.arange(1, 4 * nside, dtype=np.uint64)
    sp, npr = O.ring_info(h, rings)
    assert np.array_equal(sp, ring_start_closed_form(nside, rings))
    assert int(npr.sum()) == h.npix
    # pix2ring(startpix(r)) = r and pix2ring(startpix(r) + npr(r) - 1) = r
    assert np.array_equal(O.pix2ring(h, sp), rings)
    assert np.array_equal(O.pix2ring(h, sp + npr - 1), rings)


def test_nside1_base_pixels():
    h = O.hpx(1)
    z = np.array([2 / 3] * 4 + [0.0] * 4 + [-2 / 3] * 4)
    phi = np.array([math.pi / 4 + k * math.pi / 2 for k in range(4)] + [k * math.pi / 2 for k in range(4)] +
                   [math.pi / 4 + k * math.pi / 2 for k in range(4)])
    th, ph = O.pix2ang(h, np.arange(12, dtype=np.uint64))
    assert np.allclose(np.cos(th), z, atol=1e-15)
    assert np.allclose(ph, phi, atol=1e-15)
    assert np.array_equal(O.ang2pix(h, np.arccos(z), phi), np.arange(12, dtype=np.uint64))


@pytest.mark.parametrize("nside", [1, 2, 4, 8, 16, 32, 64, 128, 256])
def test_roundtrip_exhaustive(nside):
    h = O.hpx(nside)
    pix = np.arange(h.npix, dtype=np.uint64)
    th, ph = O.pix2ang(h, pix)
    assert np.array_equal(O.ang2pix(h, th, ph), pix)


@pytest.mark.parametrize("nside", [4096, 8192])
def test_roundtrip_random_pixels(nside):
    h = O.hpx(nside)
    rng = np.random.default_rng(nside)
    pix = rng.integers(0, h.npix, size=200_000).astype(np.uint64)
    th, ph = O.pix2ang(h, pix)
    assert np.array_equal(O.ang2pix(h, th, ph), pix)


@pytest.mark.parametrize("nside", [1, 2, 4, 16, 64, 1024, 4096, 8192])
def test_independent_pixelisation(nside):
    h = O.hpx(nside)
    rng = np.random.default_rng(100 + nside)
    n = 300_000
    th = np.arccos(rng.uniform(-1, 1, n))
    ph = rng.uniform(0, 2 * np.pi, n)
    assert np.array_equal(O.ang2pix(h, th, ph), ang2pix_ring(nside, th, ph))


def test_independent_pixelisation_raster_fields():
    # the scan geometries the workloads use, including polar caps
    h = O.hpx(8192)
    rng = np.random.default_rng(5)
    for dec_c in (-60.0, 0.0, 30.0, 45.0, 70.0, 80.0, 89.5):
        lat = dec_c + rng.uniform(-2, 2, 100_000)
        lat = np.clip(lat, -89.999, 89.999)
        lon = rng.uniform(170, 190, 100_000)
        keys = O.sample_keys(h, lon, lat)
        th = O.HALFPI - O.DEG2RAD * lat
        assert np.array_equal(keys, ang2pix_ring(8192, th, O.DEG2RAD * lon))


def test_ring_monotone_in_theta_and_pixel_monotone_in_phi():
    h = O.hpx(8192)
    th = np.linspace(1e-4, math.pi - 1e-4, 20001)
    rings = O.pix2ring(h, O.ang2pix(h, th, np.full_like(th, 1.0)))
    assert np.all(np.diff(rings.astype(np.int64)) >= 0)
    # along a ring-centre colatitude the pixel index grows with phi and stays in
    # the ring.  (phi on a polar facet boundary, tt integer, is a pixel corner:
    # there the reference's window pixel can leave the ring; the grid avoids it.)
    for r in (3, 700, 8192, 20000, 30000):
        sp, npr = O.ring_info(h, np.array([r], dtype=np.uint64))
        t, _ = O.pix2ang(h, sp)
        ph = np.linspace(0, 2 * math.pi, 10000, endpoint=False) + 1e-4 * math.sqrt(2)
        p = O.ang2pix(h, np.full_like(ph, t[0]), ph).astype(np.int64)
        assert np.all(np.diff(p) >= 0)
        assert p.min() >= int(sp[0]) and p.max() < int(sp[0] + npr[0])


def test_invalid_theta_gives_max_y():
    h = O.hpx(8192)
    out = O.ang2pix(h, np.array([-1e-9, math.pi + 1e-9, np.nan]), np.zeros(3))
    assert np.all(out == O.MAX_Y)


def test_set_optimal_nside():
    # one power of two above floor(log2(round(sqrt(pi/3)/res))) (HCGrid/helpers.cu:98-102)
    for res_deg in np.geomspace(1e-3, 5.0, 400):
        res = math.radians(res_deg)
        v = int(math.sqrt(math.pi / 3) / res + 0.5)
        expect = 1 << (v.bit_length() - 1 + 1) if v > 0 else 2
        assert O.set_optimal_nside(O.DEG2RAD * res_deg) == expect
    # the sigmas of the workloads: sigma/2 resolution
    assert O.set_optimal_nside(O.DEG2RAD * 0.0075) == 8192
    assert O.set_optimal_nside(O.DEG2RAD * 0.01) == 8192
    assert O.set_optimal_nside(O.DEG2RAD * 0.015) == 4096


def test_true_angular_distance_closed_form():
    L = O.lib()
    # along a meridian the distance is the latitude difference
    for b1, b2 in ((0.1, 0.3), (-0.5, 0.2), (1.2, 1.5)):
        assert math.isclose(L.orc_true_angular_distance(0.7, b1, 0.7, b2), abs(b1 - b2), rel_tol=1e-13)
    # on the equator the distance is the longitude difference
    assert math.isclose(L.orc_true_angular_distance(0.1, 0.0, 0.45, 0.0), 0.35, rel_tol=1e-13)
    # quarter circle
    assert math.isclose(L.orc_true_angular_distance(0.0, 0.0, 0.0, math.pi / 2), math.pi / 2, rel_tol=1e-13)
