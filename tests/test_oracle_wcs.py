"""The oracle's general celestial WCS (orc_wcs_celestial) pinned on the CPU:
the reference's set_WCS takes the target coordinates from wcslib's wcsp2s
(HCGrid/gmap.cpp:183-250), which is absent here, so the restatement is held
to what Calabretta & Greisen 2002 fix independently of it: closed forms of
each projection at delta_0 = 0, the forward projection and rotation of
every cell returning its pixel offsets, and, for the tilted cylindricals,
the reference point, the central meridian, the handedness and the isometry
of the native -> celestial rotation."""
import math

import numpy as np
import pytest

import oracle as O

SIDE = 61


def _offsets(m, crpix=31.0, side=SIDE):
    """Intermediate world coordinates (degrees) of every cell: m (p - crpix)."""
    px = (np.arange(side * side) % side) + 1 - crpix
    py = (np.arange(side * side) // side) + 1 - crpix
    m = np.asarray(m, dtype=np.float64).reshape(2, 2)
    return m[0, 0] * px + m[0, 1] * py, m[1, 0] * px + m[1, 1] * py


def _dlon(a, b):
    return (a - b + 180.0) % 360.0 - 180.0


def _rot(deg, cdelt1, cdelt2):
    r = math.radians(deg)
    return [cdelt1 * math.cos(r), -cdelt2 * math.sin(r), cdelt1 * math.sin(r), cdelt2 * math.cos(r)]


def _gc(a1, d1, a2, d2):
    """Great-circle distance (radians) by the haversine form."""
    a1, d1, a2, d2 = map(np.radians, (a1, d1, a2, d2))
    h = np.sin((d2 - d1) / 2) ** 2 + np.cos(d1) * np.cos(d2) * np.sin((a2 - a1) / 2) ** 2
    return 2 * np.arcsin(np.sqrt(np.clip(h, 0, 1)))


@pytest.mark.parametrize("proj", ["SIN", "TAN", "ARC", "ZEA", "STG"])
def test_zenithal_members_equal_the_zenithal_oracle(proj):
    c = 0.4
    got = O.wcs_celestial(proj, SIDE, SIDE, 31.0, 31.0, [-c, 0, 0, c], 123.0, 21.0, 180.0)
    ref = O.wcs_zenithal(proj, SIDE, SIDE, 31.0, 31.0, -c, c, 123.0, 21.0, 180.0)
    fin = np.isfinite(ref[1])
    assert np.array_equal(fin, np.isfinite(got[1]))
    assert np.abs(_dlon(got[0][fin], ref[0][fin])).max() < 1e-12
    assert np.abs(got[1][fin] - ref[1][fin]).max() < 1e-12


@pytest.mark.parametrize("xi,eta,d0", [(0.1, -0.05, 21.0), (0.3, 0.2, -48.0), (-0.25, 0.0, 70.0)])
def test_sin_slant_forward_projection(xi, eta, d0):
    # world -> native (the inverse of eq. 2 with (alpha_p, delta_p) = CRVAL,
    # phi_p = 180) -> the slanted orthographic projection (sec. 5.1.5)
    a0, c = 40.0, 0.5
    m = _rot(15.0, -c, c)
    xw, yw = O.wcs_celestial("SIN", SIDE, SIDE, 31.0, 31.0, m, a0, d0, 180.0, pv=(xi, eta))
    fin = np.isfinite(yw)
    assert fin.sum() > 0.9 * fin.size
    a, d = np.radians(xw[fin] - a0), np.radians(yw[fin])
    dp = math.radians(d0)
    # the native unit vector: (cos th sin(phi - phi_p), cos th cos(phi - phi_p),
    # sin th); with phi_p = 180, cos th sin(phi) = -n_x, cos th cos(phi) = -n_y
    n_x = -np.cos(d) * np.sin(a)
    n_y = np.sin(d) * math.cos(dp) - np.cos(d) * math.sin(dp) * np.cos(a)
    n_z = np.sin(d) * math.sin(dp) + np.cos(d) * math.cos(dp) * np.cos(a)
    X = -n_x + xi * (1 - n_z)
    Y = n_y + eta * (1 - n_z)
    x, y = _offsets(m)
    assert np.abs(np.degrees(X) - x[fin]).max() < 1e-10
    assert np.abs(np.degrees(Y) - y[fin]).max() < 1e-10


def test_sin_without_slant_is_sin():
    a = O.wcs_celestial("SIN", SIDE, SIDE, 31.0, 31.0, [-0.5, 0, 0, 0.5], 10.0, -30.0, 180.0, pv=(0.0, 0.0))
    b = O.wcs_zenithal("SIN", SIDE, SIDE, 31.0, 31.0, -0.5, 0.5, 10.0, -30.0, 180.0)
    fin = np.isfinite(b[1])
    assert np.array_equal(fin, np.isfinite(a[1]))
    assert np.abs(_dlon(a[0][fin], b[0][fin])).max() < 1e-12 and np.abs(a[1][fin] - b[1][fin]).max() < 1e-12


# (lon - lon_0, lat) of intermediate (x, y) degrees at delta_0 = 0, where the
# native and celestial frames coincide up to the shift alpha_0 (eq. 2 with
# delta_p = 90): each projection's textbook inverse, written independently
def _car(x, y):
    return x, np.where(np.abs(y) <= 90, y, np.nan)


def _mer(x, y):
    return x, np.degrees(np.arctan(np.sinh(np.radians(y))))  # the Gudermannian


def _cea(x, y, lam=0.7):
    s = lam * np.radians(y)
    return x, np.where(np.abs(s) <= 1, np.degrees(np.arcsin(np.clip(s, -1, 1))), np.nan)


def _sfl(x, y):
    lon = x / np.cos(np.radians(y))
    ok = (np.abs(y) <= 90) & (np.abs(lon) <= 180)
    return np.where(ok, lon, np.nan), np.where(ok, y, np.nan)


CLOSED = {"CAR": _car, "MER": _mer, "CEA": _cea, "SFL": _sfl}


@pytest.mark.parametrize("proj", sorted(CLOSED))
@pytest.mark.parametrize("rot", [0.0, 25.0])
def test_cylindricals_at_delta0_zero(proj, rot):
    a0, c = 200.0, 3.5  # reaches |y| > 90 and |x| > 180 at the grid's corners
    m = _rot(rot, -c, c)
    xw, yw = O.wcs_celestial(proj, SIDE, SIDE, 31.0, 31.0, m, a0, 0.0, pv=(0.7, 0.0))
    x, y = _offsets(m)
    dl, lat = CLOSED[proj](x, y)
    fin = np.isfinite(lat)
    assert np.array_equal(fin, np.isfinite(yw)), proj
    assert (~fin).sum() > 0 or proj == "MER"
    assert np.abs(yw[fin] - lat[fin]).max() < 1e-10
    keep = fin & (np.abs(lat) < 89.9)  # the longitude at a pole is arbitrary
    assert np.abs(_dlon(xw[keep], a0 + dl[keep])).max() < 1e-9


def test_ait_forward_projection():
    # Hammer-Aitoff forward: x = 2 g cos(d) sin(l/2), y = g sin(d), g =
    # sqrt(2 / (1 + cos(d) cos(l/2))), radians; NaN outside the ellipse
    a0, c = 0.0, 3.1
    m = _rot(10.0, -c, c)
    xw, yw = O.wcs_celestial("AIT", SIDE, SIDE, 31.0, 31.0, m, a0, 0.0)
    x, y = _offsets(m)
    inside = (np.radians(x) / 4) ** 2 + (np.radians(y) / 2) ** 2 <= 0.5
    assert np.array_equal(inside, np.isfinite(yw)) and (~inside).sum() > 0
    l, d = np.radians(_dlon(xw[inside], a0)), np.radians(yw[inside])
    g = np.sqrt(2 / (1 + np.cos(d) * np.cos(l / 2)))
    assert np.abs(np.degrees(2 * g * np.cos(d) * np.sin(l / 2)) - x[inside]).max() < 1e-9
    assert np.abs(np.degrees(g * np.sin(d)) - y[inside]).max() < 1e-9


@pytest.mark.parametrize("proj", ["CAR", "MER", "SFL", "AIT", "CEA"])
@pytest.mark.parametrize("d0,latpole,north_up", [(30.0, 90.0, True), (-40.0, 90.0, True), (75.0, 90.0, True),
                                                 (30.0, -90.0, False)])
def test_tilted_rotation_invariants(proj, d0, latpole, north_up):
    a0, c, side = 150.0, 0.9, 41
    m = [-c, 0.0, 0.0, c]
    xw, yw = O.wcs_celestial(proj, side, side, 21.0, 21.0, m, a0, d0, latpole=latpole, pv=(0.7, 0.0))
    # native coordinates: the same projection at delta_0 = 0, alpha_0 = 0
    nx, ny = O.wcs_celestial(proj, side, side, 21.0, 21.0, m, 0.0, 0.0, pv=(0.7, 0.0))
    k0 = 20 * side + 20
    assert abs(_dlon(xw[k0], a0)) < 1e-11 and abs(yw[k0] - d0) < 1e-11
    # the central column is the meridian alpha_0, north up or down: delta =
    # delta_0 +- the native latitude
    col = np.arange(side) * side + 20
    expect = d0 + (ny[col] if north_up else -ny[col])
    ok = np.isfinite(expect)
    ok[ok] &= np.abs(expect[ok]) < 89.0
    assert ok.sum() > 10
    assert np.abs(yw[col][ok] - expect[ok]).max() < 1e-9
    assert np.abs(_dlon(xw[col][ok], a0)).max() < 1e-8
    # handedness: east of the reference (x > 0, CDELT1 < 0 so smaller pixel
    # x) the longitude grows when north is up
    east = 20 * side + 19
    assert (_dlon(xw[east], a0) > 0) == north_up
    # the rotation is an isometry: pairwise distances equal the native ones
    fin = np.flatnonzero(np.isfinite(ny) & np.isfinite(yw))
    rng = np.random.default_rng(5)
    i, j = rng.choice(fin, 400), rng.choice(fin, 400)
    dw = _gc(xw[i], yw[i], xw[j], yw[j])
    dn = _gc(nx[i], ny[i], nx[j], ny[j])
    assert np.abs(dw - dn).max() < 1e-11


def test_gls_aips_closed_form():
    a0, d0, c = 80.0, 25.0, 2.0
    m = _rot(0.0, -c, c)
    xw, yw = O.wcs_celestial("GLS", SIDE, SIDE, 31.0, 31.0, m, a0, d0)
    x, y = _offsets(m)
    dec = d0 + y
    dl = x / np.cos(np.radians(dec))
    ok = (np.abs(dec) <= 90) & (np.abs(dl) <= 180)
    assert np.array_equal(ok, np.isfinite(yw)) and (~ok).sum() > 0
    assert np.abs(yw[ok] - dec[ok]).max() < 1e-11
    keep = ok & (np.abs(dec) < 89.9)
    assert np.abs(_dlon(xw[keep], a0 + dl[keep])).max() < 1e-9


def test_lonpole_without_a_native_pole():
    # eq. 8 has no root: sin(delta_0) > |cos(phi_p)|
    assert O.wcs_celestial("CAR", 8, 8, 4.0, 4.0, [-1, 0, 0, 1], 10.0, 30.0, 180.0) is None
    assert O.wcs_celestial("CAR", 8, 8, 4.0, 4.0, [-1, 0, 0, 1], 10.0, 30.0, 70.0) is None
    # phi_p = 90 at delta_0 = 0: delta_p is free and takes LATPOLE
    xw, yw = O.wcs_celestial("CAR", 9, 9, 5.0, 5.0, [-1, 0, 0, 1], 10.0, 0.0, 90.0, 90.0)
    k0 = 4 * 9 + 4
    assert abs(_dlon(xw[k0], 10.0)) < 1e-12 and abs(yw[k0]) < 1e-12
