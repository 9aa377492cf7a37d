"""Pins the oracle's pre-sort, ring LUT and gridding restatement (CPU only).

* presort     : equals numpy's stable argsort of the keys.
* ring LUT    : equals the definition lut[k] = first index with key >=
                startpix(firstring + k), startpix from the closed form.
* search      : searchLastPosLessThan semantics incl. the unsigned wrap.
* grid        : the loop restatement (LUT + binary search + break test)
                equals a pair-by-pair evaluation of the selection rule, bit
                for bit; and where the reference's window holds the whole
                disc (|dec| < 20 deg, or the polar whole-ring branch at
                |dec| >= 70 deg, first used ring far away) it equals a
                textbook long-double disc gridder.
* multi       : the multi-channel form equals the single-channel loop run
                once per channel, bit for bit (the reference's channel loop).
* bearing     : closed forms (north 0, east pi/2, south pi, west -pi/2) and
                the (east, north)-frame vector form in long double.
* kernels     : closed-form values of the three weight functions; GAUSS2D
                along its axes through the bearing.
* WCS         : closed forms of the SIN deprojection.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from hcgrid_amd import synth
from healpix_indep import ring_start_closed_form

from helpers import map_errors, normalise


def test_presort_is_stable_argsort():
    h = O.hpx(8192)
    wl = synth.raster_small(50, 400, 10, seed=2)
    lon, lat = wl.lon.copy(), wl.lat.copy()
    lon[::7] = lon[3]  # many equal keys
    lat[::7] = lat[3]
    keys = O.sample_keys(h, lon, lat)
    ks, perm = O.presort(h, lon, lat)
    ref = np.argsort(keys, kind="stable")
    assert np.array_equal(perm, ref.astype(np.uint32))
    assert np.array_equal(ks[:-1], keys[ref])
    assert ks[-1] == h.npix


def test_ring_lut_definition():
    h = O.hpx(8192)
    wl = synth.raster_250k()
    ks, perm = O.presort(h, wl.lon, wl.lat)
    n = wl.n
    fr, used, lut = O.ring_lut(h, ks, n)
    assert fr == int(O.pix2ring(h, ks[:1])[0])
    assert used == int(O.pix2ring(h, ks[n - 1:n])[0]) - fr + 1
    starts = ring_start_closed_form(8192, np.arange(fr, fr + used))
    expect = np.searchsorted(ks[:n], starts, side="left")
    expect[0] = 0
    assert np.array_equal(lut[:used], expect.astype(np.uint32))
    assert lut[used] == n


def test_search_last_less_semantics():
    L = O.lib()
    rng = np.random.default_rng(0)
    for _ in range(300):
        v = np.sort(rng.integers(0, 50, size=rng.integers(1, 40))).astype(np.uint64)
        left = int(rng.integers(0, len(v) + 1))
        right = int(rng.integers(0, len(v) + 1))
        key = int(rng.integers(0, 55))
        got = L.orc_search_last_less(v.ctypes.data_as(C.c_void_p), left, right, key)
        cand = [i for i in range(left, right) if v[i] < key]
        expect = cand[-1] if (right > left and v[left] < key) else right
        if right > left and v[left] < key:
            expect = max(cand)
        assert got == expect
    v = np.arange(10, dtype=np.uint64)
    # left = start_int - 1 wrapped to 0xFFFFFFFF: returns right (the first-ring skip)
    assert L.orc_search_last_less(v.ctypes.data_as(C.c_void_p), 0xFFFFFFFF, 5, 3) == 5


def _setup(wl, side=None, factor=1, sigma=0.02, radius=0.06, res=0.01, kernel_type=O.GAUSS1D, params=None):
    t = wl.target
    side = side or t.side
    xw, yw = O.wcs_sin(side, side, side / 2.0, side / 2.0, -t.cdelt, t.cdelt, t.crval1, t.crval2)
    r = O.pipeline(wl.lon, wl.lat, wl.values, xw, yw, side, sigma, radius, res, factor=factor,
                   kernel_type=kernel_type, kernel_params=params)
    return r, xw, yw


@pytest.mark.parametrize("dec,factor", [(30.0, 1), (30.0, 2), (30.0, 3), (-45.0, 1), (0.0, 1), (72.0, 1),
                                        (84.0, 1)])
def test_compat_equals_pairwise_bruteforce(dec, factor):
    sp = 0.0032 / max(math.cos(math.radians(dec)), 0.1)
    wl = synth.raster_small(30, 120, 16, dec_c=dec, spacing_ra=sp, seed=31)
    r, xw, yw = _setup(wl, factor=factor)
    dc, wc = O.grid_bruteforce(r["h"], r["g"], r["lon_s"], r["lat_s"], r["data_s"], None, r["keys_sorted"],
                               r["firstring"], r["usedrings"], xw, yw)
    assert r["hangs"] == 0
    assert np.array_equal(dc, r["datacube"]) and np.array_equal(wc, r["weightscube"])
    assert (wc > 0).sum() > 0


@pytest.mark.parametrize("kt,params", [(O.GAUSS2D, [0.02, 0.012, 0.4]), (O.TAPERED_SINC, [0.02, 1.55, 2.52])])
def test_compat_equals_bruteforce_other_kernels(kt, params):
    wl = synth.raster_small(30, 120, 16, seed=37)
    r, xw, yw = _setup(wl, kernel_type=kt, params=params)
    dc, wc = O.grid_bruteforce(r["h"], r["g"], r["lon_s"], r["lat_s"], r["data_s"], None, r["keys_sorted"],
                               r["firstring"], r["usedrings"], xw, yw)
    assert np.array_equal(dc, r["datacube"]) and np.array_equal(wc, r["weightscube"])


def test_first_used_ring_contributes_nothing():
    # a single scan line: every sample lies in the first used ring, so the
    # reference's start_int - 1 wrap leaves every cell empty
    wl = synth.raster_small(1, 300, 12, seed=3)
    r, xw, yw = _setup(wl)
    assert r["usedrings"] == 1
    assert np.all(r["weightscube"] == 0)


@pytest.mark.parametrize("dec", [0.0, -12.0, 15.0])
def test_compat_equals_textbook_disc_near_equator(dec):
    # At |dec| < 20 deg the +-disc longitude window (disc = R + resolution)
    # covers the whole R-disc, and with the first used ring far away the
    # reference's selection is exactly "every sample within R".
    wl = synth.raster_small(40, 160, 20, dec_c=dec, seed=41)
    lon = np.concatenate([[wl.lon[0]], wl.lon])
    lat = np.concatenate([[dec + 3.0], wl.lat])  # a far sample owns the first used ring
    val = np.concatenate([[0.0], wl.values.astype(np.float64)])
    t = wl.target
    xw, yw = O.wcs_sin(t.side, t.side, t.side / 2.0, t.side / 2.0, -t.cdelt, t.cdelt, t.crval1, t.crval2)
    r = O.pipeline(lon, lat, val, xw, yw, t.side, 0.02, 0.06, 0.01)
    p0 = 0.5 / 0.02 ** 2
    dc, wc = O.grid_disc_textbook(lon, lat, val, xw, yw, 0.06, p0)
    e = map_errors(normalise(r["datacube"], r["weightscube"]), normalise(dc, wc))
    assert e["nan_match"]
    # FP64 vs long-double summation of noise-valued terms: 1e-12 of the peak,
    # 1e-9 per cell.  One sample more or less would move a weight sum by
    # >= exp(-4.5) / sum(w) ~ 1e-4, so the 1e-12 weight check pins the selection.
    assert e["max_rel"] < 1e-12 and e["cell_rel"] < 1e-9
    assert np.allclose(r["weightscube"], wc, rtol=1e-12, atol=0)


def test_textbook_disc_closed_forms():
    # one sample per case at a known great-circle distance from one cell:
    # across the pole (lon + 180, d = 2 (90 - dec)), across longitude 0, along
    # a parallel at dec 60 (d = 2 asin(cos 60 sin(dlon / 2))), and just
    # outside the radius (no contribution, NaN map)
    p0 = 0.5 / 0.02 ** 2
    cases = [((10.0, 89.99), (190.0, 89.99), 0.02), ((359.99, 0.0), (0.01, 0.0), 0.02),
             ((100.0, 60.0), (100.05, 60.0), math.degrees(2 * math.asin(0.5 * math.sin(math.radians(0.025))))),
             ((50.0, -30.0), (50.0, -30.0605), None)]
    for (cx, cy), (sx, sy), d in cases:
        dc, wc = O.grid_disc_textbook(np.array([sx]), np.array([sy]), np.array([2.0]), np.array([cx]),
                                      np.array([cy]), 0.06, p0)
        if d is None:
            assert wc[0] == 0.0 and dc[0] == 0.0
        else:
            assert math.isclose(wc[0], math.exp(-d * d * p0), rel_tol=1e-12)
            assert math.isclose(dc[0], 2.0 * wc[0], rel_tol=1e-15)


def test_compat_window_truncates_east_west_at_high_dec():
    # At dec 60 the window is +-disc in longitude without 1/cos(dec), so it
    # holds only part of the disc: compat weights are strictly below the disc's.
    wl = synth.raster_small(40, 160, 12, dec_c=60.0, spacing_ra=0.0064, seed=43)
    r, xw, yw = _setup(wl)
    dc, wc = O.grid_disc_textbook(wl.lon, wl.lat, wl.values, xw, yw, 0.06, 0.5 / 0.02 ** 2)
    inner = r["weightscube"] > 0
    assert np.all(r["weightscube"][inner] < wc[inner])


def test_kernel_weight_closed_forms():
    L = O.lib()
    sig = 0.02
    g1 = O.grid_params(O.GAUSS1D, [0.5 / sig ** 2], 0.06, 0.001, 4)
    assert L.orc_kernel_weight(C.byref(g1), 0.0, 0.0) == 1.0
    assert math.isclose(L.orc_kernel_weight(C.byref(g1), sig, 0.0), math.exp(-0.5), rel_tol=1e-15)
    # circular GAUSS2D (sigma_maj = sigma_min) is exp(-d^2 / (2 sigma^2)) for any bearing
    g2 = O.grid_params(O.GAUSS2D, [sig, sig, 0.3], 0.06, 0.001, 4)
    for d, b in ((0.01, 0.0), (0.03, 1.1), (0.05, -2.0)):
        assert math.isclose(L.orc_kernel_weight(C.byref(g2), d, b), math.exp(-d * d / (2 * sig * sig)),
                            rel_tol=1e-13)
    g3 = O.grid_params(O.TAPERED_SINC, [sig, 1.55, 2.52], 0.06, 0.001, 4)
    assert L.orc_kernel_weight(C.byref(g3), 0.0, 0.0) == 1.0
    d = 0.01
    arg = math.pi * d / sig
    expect = math.sin(arg / 2.52) / (arg / 2.52) * math.exp((arg / 1.55) ** 2)  # the reference's sign
    assert math.isclose(L.orc_kernel_weight(C.byref(g3), d, 0.0), expect, rel_tol=1e-13)


# angular distance rho from the reference point at intermediate radius R
# (radians), each projection's R(theta) with theta = 90 deg - rho inverted
_RHO_OF_R = {"SIN": np.arcsin, "TAN": np.arctan, "ARC": lambda r: r,
             "ZEA": lambda r: 2 * np.arcsin(r / 2), "STG": lambda r: 2 * np.arctan(r / 2)}
_R_OF_RHO = {"SIN": np.sin, "TAN": np.tan, "ARC": lambda p: p,
             "ZEA": lambda p: 2 * np.sin(p / 2), "STG": lambda p: 2 * np.tan(p / 2)}


@pytest.mark.parametrize("proj", sorted(_RHO_OF_R))
def test_wcs_zenithal_closed_forms(proj):
    side, c, a0, d0 = 81, 0.4, 123.0, 21.0
    xw, yw = O.wcs_zenithal(proj, side, side, 41.0, 41.0, -c, c, a0, d0)
    k0 = 40 * side + 40
    assert abs(xw[k0] - a0) < 1e-12 and abs(yw[k0] - d0) < 1e-12
    # central column: ra = ra0 and dec = dec0 +- rho(R)
    X = np.arange(side)
    col = X * side + 40
    y = np.radians(c * (X + 1 - 41.0))
    assert np.allclose(yw[col], d0 + np.sign(y) * np.degrees(_RHO_OF_R[proj](np.abs(y))), atol=1e-11)
    assert np.allclose(xw[col], a0, atol=1e-11)
    # forward projection of every cell returns its pixel offsets: the
    # orthographic direction scaled by R(rho) / sin(rho)
    a, d = np.radians(xw - a0), np.radians(yw)
    dd = math.radians(d0)
    ox = np.cos(d) * np.sin(a)
    oy = np.sin(d) * math.cos(dd) - np.cos(d) * math.sin(dd) * np.cos(a)
    rho = np.arctan2(np.hypot(ox, oy), np.sin(d) * math.sin(dd) + np.cos(d) * math.cos(dd) * np.cos(a))
    s = np.where(rho > 0, _R_OF_RHO[proj](rho) / np.where(rho > 0, np.sin(rho), 1.0), 1.0)
    px = (np.arange(side * side) % side) + 1
    py = (np.arange(side * side) // side) + 1
    assert np.allclose(-np.degrees(s * ox), c * (px - 41.0), atol=1e-10)  # CDELT1 < 0
    assert np.allclose(np.degrees(s * oy), c * (py - 41.0), atol=1e-10)
    # SIN is the same function as the SIN-only entry
    if proj == "SIN":
        xs, ys = O.wcs_sin(side, side, 41.0, 41.0, -c, c, a0, d0)
        assert np.array_equal(xs, xw) and np.array_equal(ys, yw)


@pytest.mark.parametrize("proj,rmax", [("SIN", 1.0), ("ARC", math.pi), ("ZEA", 2.0), ("TAN", None), ("STG", None)])
def test_wcs_zenithal_domains(proj, rmax):
    # NaN exactly where R leaves the projection's domain; the pole target
    # with LONPOLE 0 keeps ra = ra0 + phi (Paper II eq. 2 at delta_p = 90)
    side, c = 41, 9.3  # no cell lies on a domain boundary
    xw, yw = O.wcs_zenithal(proj, side, side, 21.0, 21.0, -c, c, 30.0, 90.0, 0.0)
    px = (np.arange(side * side) % side) + 1 - 21.0
    py = (np.arange(side * side) // side) + 1 - 21.0
    r = np.radians(c * np.hypot(px, py))
    out = np.zeros_like(r, dtype=bool) if rmax is None else r > rmax
    assert np.array_equal(np.isnan(yw), out) and np.array_equal(np.isnan(xw), out)
    ok = ~out & (r > 0)
    assert np.allclose(yw[ok], 90.0 - np.degrees(_RHO_OF_R[proj](r[ok])), atol=1e-10)
    phi = np.degrees(np.arctan2(np.radians(-c * px), np.radians(-c * py)))
    dra = (xw[ok] - (30.0 + phi[ok] + 180.0)) % 360.0
    assert np.all(np.minimum(dra, 360.0 - dra) < 1e-9)


def test_wcs_sin_closed_forms():
    side, c = 101, 0.05
    xw, yw = O.wcs_sin(side, side, 51.0, 51.0, -c, c, 123.0, 41.0)
    X = np.arange(side)
    # reference pixel (x+1 = CRPIX) maps to CRVAL
    k0 = 50 * side + 50
    assert abs(xw[k0] - 123.0) < 1e-12 and abs(yw[k0] - 41.0) < 1e-12
    # central column: dec = dec0 + asin(y), ra = ra0
    col = X * side + 50
    y = c * (X + 1 - 51.0)
    assert np.allclose(yw[col], 41.0 + np.degrees(np.arcsin(np.radians(y))), atol=1e-11)
    assert np.allclose(xw[col], 123.0, atol=1e-11)
    # forward orthographic projection of every cell returns its pixel offsets
    a, d = np.radians(xw - 123.0), np.radians(yw)
    d0 = math.radians(41.0)
    fx = np.degrees(np.cos(d) * np.sin(a))
    fy = np.degrees(np.sin(d) * math.cos(d0) - np.cos(d) * math.sin(d0) * np.cos(a))
    px = (np.arange(side * side) % side) + 1
    py = (np.arange(side * side) // side) + 1
    assert np.allclose(-fx, c * (px - 51.0), atol=1e-11)  # CDELT1 < 0
    assert np.allclose(fy, c * (py - 51.0), atol=1e-11)


@pytest.mark.parametrize("dec,factor", [(30.0, 1), (72.0, 1), (84.0, 1), (30.0, 2), (30.0, 3), (-45.0, 2)])
def test_compat_multi_equals_compat_per_channel(dec, factor):
    # orc_grid_compat_multi (weights once, every channel) must be what the
    # reference's channel-by-channel pipeline gives: orc_grid_compat run once
    # per channel on that channel's values, bit for bit, weights included
    sp = 0.0032 / max(math.cos(math.radians(dec)), 0.1)
    wl = synth.raster_small(30, 120, 16, dec_c=dec, spacing_ra=sp, seed=47)
    r, xw, yw = _setup(wl, factor=factor)
    rng = np.random.default_rng(5)
    nchan = 5
    V = rng.normal(0.0, 1.0, (nchan, wl.n)).astype(np.float32)
    V[0] = wl.values  # channel 0: the single-channel data itself
    Vs = V[:, r["perm"]]
    side = int(r["g"].side)
    out, ws = O.grid_compat_multi(r["h"], r["g"], r["lon_s"], r["lat_s"], Vs, None, r["keys_sorted"],
                                  r["firstring"], r["usedrings"], r["lut"], xw, yw, 0, side)
    assert (ws > 0).sum() > 0
    for ch in range(nchan):
        dc, wc, hangs = O.grid_compat(r["h"], r["g"], r["lon_s"], r["lat_s"], Vs[ch].astype(np.float64), None,
                                      r["keys_sorted"], r["firstring"], r["usedrings"], r["lut"], xw, yw)
        assert hangs == 0
        assert np.array_equal(out[ch], dc), f"channel {ch}"
        assert np.array_equal(ws, wc)
    # a band of rows is the same band of the whole map
    b0, b1 = 5, 9
    ob, wb = O.grid_compat_multi(r["h"], r["g"], r["lon_s"], r["lat_s"], Vs, None, r["keys_sorted"],
                                 r["firstring"], r["usedrings"], r["lut"], xw, yw, b0, b1)
    assert np.array_equal(ob, out[:, b0 * side:b1 * side]) and np.array_equal(wb, ws[b0 * side:b1 * side])


def test_compat_multi_with_sample_weights():
    wl = synth.raster_small(30, 120, 16, seed=49)
    r, xw, yw = _setup(wl)
    w = np.random.default_rng(2).uniform(0.2, 2.0, wl.n)
    ws_s = w[r["perm"]]
    dc, wc, _ = O.grid_compat(r["h"], r["g"], r["lon_s"], r["lat_s"], r["data_s"], ws_s, r["keys_sorted"],
                              r["firstring"], r["usedrings"], r["lut"], xw, yw)
    out, ws = O.grid_compat_multi(r["h"], r["g"], r["lon_s"], r["lat_s"],
                                  r["data_s"].astype(np.float32)[None, :], ws_s, r["keys_sorted"], r["firstring"],
                                  r["usedrings"], r["lut"], xw, yw, 0, 16)
    assert np.array_equal(out[0], dc) and np.array_equal(ws, wc)


def _bearing(l1, b1, l2, b2):
    L = O.lib()
    L.orc_great_circle_bearing.restype = C.c_double
    L.orc_great_circle_bearing.argtypes = [C.c_double] * 4
    return L.orc_great_circle_bearing(l1, b1, l2, b2)


def test_great_circle_bearing_closed_forms():
    # bearing of (l2, b2) seen from (l1, b1), radians east of north
    r = math.radians
    assert _bearing(r(40), r(10), r(40), r(11)) == 0.0                    # due north
    assert math.isclose(_bearing(r(40), r(10), r(40), r(9)), math.pi, rel_tol=1e-15)  # due south
    assert math.isclose(_bearing(r(0), 0.0, r(1), 0.0), math.pi / 2, rel_tol=1e-15)   # due east on the equator
    assert math.isclose(_bearing(r(0), 0.0, r(-1), 0.0), -math.pi / 2, rel_tol=1e-15)  # due west
    # from a point towards the north pole, from the south pole towards anywhere
    assert abs(_bearing(r(123), r(-20), r(77), r(90))) < 1e-15
    # the back azimuth on the equator: swapping the points turns east into west
    assert math.isclose(_bearing(r(1), 0.0, r(0), 0.0), -math.pi / 2, rel_tol=1e-15)


def test_great_circle_bearing_against_vector_form():
    # an independent construction: the direction of q in the local
    # (east, north) frame of p, atan2(q.E, q.N), evaluated in long double
    rng = np.random.default_rng(3)
    n = 2000
    l1 = rng.uniform(0, 2 * np.pi, n)
    b1 = rng.uniform(-1.5, 1.5, n)
    d = rng.uniform(1e-4, 0.2, n)
    l2 = l1 + rng.normal(0, 1, n) * d
    b2 = np.clip(b1 + rng.normal(0, 1, n) * d, -1.55, 1.55)
    ld = np.longdouble
    L1, B1, L2, B2 = (a.astype(ld) for a in (l1, b1, l2, b2))
    qx, qy, qz = np.cos(B2) * np.cos(L2), np.cos(B2) * np.sin(L2), np.sin(B2)
    ex, ey = -np.sin(L1), np.cos(L1)
    nx, ny, nz = -np.sin(B1) * np.cos(L1), -np.sin(B1) * np.sin(L1), np.cos(B1)
    want = np.arctan2(qx * ex + qy * ey, qx * nx + qy * ny + qz * nz).astype(np.float64)
    got = np.array([_bearing(*v) for v in zip(l1, b1, l2, b2)])
    assert np.max(np.abs(got - want)) < 1e-9  # separations >= 1e-4 rad: ~1e-12 rad of FP64 rounding
    # arguments swapped (the bearing of p seen from q) would fail it
    swapped = np.array([_bearing(c, d_, a, b) for a, b, c, d_ in zip(l1, b1, l2, b2)])
    assert np.max(np.abs(swapped - want)) > 1.0


def test_gauss2d_weight_along_the_axes():
    # GAUSS2D (gridding.cu:242-249) with position angle 0: a sample due north
    # of the cell is weighted exp(-d^2 / (2 p0^2)), one due east on the
    # equator exp(-d^2 / (2 p1^2)) -- the pair of calls the compat loop makes
    L = O.lib()
    p0, p1 = 0.03, 0.015
    g2 = O.grid_params(O.GAUSS2D, [p0, p1, 0.0], 0.1, 0.001, 4)
    r = math.radians
    d = 0.02
    b_n = _bearing(r(10), r(5), r(10), r(5 + d))
    b_e = _bearing(r(10), 0.0, r(10 + d), 0.0)
    assert math.isclose(L.orc_kernel_weight(C.byref(g2), d, b_n), math.exp(-d * d / (2 * p0 * p0)), rel_tol=1e-13)
    assert math.isclose(L.orc_kernel_weight(C.byref(g2), d, b_e), math.exp(-d * d / (2 * p1 * p1)), rel_tol=1e-13)


@pytest.mark.parametrize("dec", [75.0, -80.0, 72.5])
def test_compat_polar_whole_rings_equal_textbook_disc(dec):
    # With the window anchor at colatitude <= 20 deg (or >= 160 deg) the
    # reference takes whole rings (gridding.cu:336-341), and its rings span
    # +-disc = R + resolution in colatitude: every sample within R of a cell
    # is a candidate.  With the first used ring owned by a far sample, the
    # compat selection is then exactly the disc, which the long-double
    # textbook gridder evaluates with no rings at all.
    sp = 0.0032 / math.cos(math.radians(dec))
    wl = synth.raster_small(40, 160, 20, dec_c=dec, spacing_ra=sp, seed=53)
    lon = np.concatenate([[wl.lon[0]], wl.lon])
    lat = np.concatenate([[dec + 3.0], wl.lat])  # north of the field: the first used ring
    val = np.concatenate([[0.0], wl.values.astype(np.float64)])
    t = wl.target
    xw, yw = O.wcs_sin(t.side, t.side, t.side / 2.0, t.side / 2.0, -t.cdelt, t.cdelt, t.crval1, t.crval2)
    theta = 90.0 - yw
    assert np.all((theta <= 20.0) | (theta >= 160.0))  # every cell's window is the whole-ring branch
    r = O.pipeline(lon, lat, val, xw, yw, t.side, 0.02, 0.06, 0.01)
    assert r["hangs"] == 0
    dc, wc = O.grid_disc_textbook(lon, lat, val, xw, yw, 0.06, 0.5 / 0.02 ** 2)
    assert (wc > 0).sum() > 0.5 * wc.size
    e = map_errors(normalise(r["datacube"], r["weightscube"]), normalise(dc, wc))
    assert e["nan_match"]
    assert e["max_rel"] < 1e-12 and e["cell_rel"] < 1e-9
    assert np.allclose(r["weightscube"], wc, rtol=1e-12, atol=0)
