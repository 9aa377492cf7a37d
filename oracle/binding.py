"""numpy binding of oracle/liboracle.so (the CPU restatement).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

MAX_Y = 1073741824
GAUSS1D, GAUSS2D, TAPERED_SINC = 1, 2, 3
PI = 3.1415926535897932384626433832
DEG2RAD = PI / 180.0
RAD2DEG = 180.0 / PI
HALFPI = PI / 2.0


class OrcHpx(C.Structure):
    _fields_ = [
        ("nside", C.c_uint64), ("order", C.c_uint64), ("nrings", C.c_uint64),
        ("npface", C.c_uint64), ("ncap", C.c_uint64), ("npix", C.c_uint64),
        ("fact1", C.c_double), ("fact2", C.c_double), ("omega", C.c_double),
        ("resolution", C.c_double),
    ]


class OrcGrid(C.Structure):
    _fields_ = [
        ("kernel_type", C.c_uint32), ("factor", C.c_uint32), ("side", C.c_uint32),
        ("bearing_needed", C.c_uint32), ("params", C.c_double * 3),
        ("sphere_radius", C.c_double), ("disc_size", C.c_double),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-B", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        vp = C.c_void_p
        L.orc_ilog2.restype = C.c_uint64
        L.orc_ilog2.argtypes = [C.c_uint64]
        L.orc_isqrt.restype = C.c_uint64
        L.orc_isqrt.argtypes = [C.c_uint64]
        L.orc_fmodulo.restype = C.c_double
        L.orc_fmodulo.argtypes = [C.c_double, C.c_double]
        L.orc_set_optimal_nside.restype = C.c_uint64
        L.orc_set_optimal_nside.argtypes = [C.c_double]
        L.orc_nside_to_order.restype = C.c_uint64
        L.orc_nside_to_order.argtypes = [C.c_uint64]
        L.orc_true_angular_distance.restype = C.c_double
        L.orc_true_angular_distance.argtypes = [C.c_double] * 4
        L.orc_hpx_init.restype = None
        L.orc_hpx_init.argtypes = [P(OrcHpx), C.c_uint64]
        L.orc_ang2pix.restype = C.c_uint64
        L.orc_ang2pix.argtypes = [P(OrcHpx), C.c_double, C.c_double]
        L.orc_pix2ring.restype = C.c_uint64
        L.orc_pix2ring.argtypes = [P(OrcHpx), C.c_uint64]
        for name in ("orc_ang2pix_many", "orc_pix2ang_many", "orc_pix2ring_many",
                     "orc_ring_info_many", "orc_sample_keys"):
            getattr(L, name).restype = None
        L.orc_ang2pix_many.argtypes = [P(OrcHpx), C.c_int64, vp, vp, vp]
        L.orc_pix2ang_many.argtypes = [P(OrcHpx), C.c_int64, vp, vp, vp]
        L.orc_pix2ring_many.argtypes = [P(OrcHpx), C.c_int64, vp, vp]
        L.orc_ring_info_many.argtypes = [P(OrcHpx), C.c_int64, vp, vp, vp]
        L.orc_sample_keys.argtypes = [P(OrcHpx), C.c_int64, vp, vp, vp]
        L.orc_sample_keys_libm_sensitive.restype = C.c_int64
        L.orc_sample_keys_libm_sensitive.argtypes = [P(OrcHpx), C.c_int64, vp, vp, C.c_int, vp]
        L.orc_great_circle_bearing.restype = C.c_double
        L.orc_great_circle_bearing.argtypes = [C.c_double] * 4
        L.orc_search_last_less.restype = C.c_uint32
        L.orc_search_last_less.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint64]
        L.orc_presort.restype = None
        L.orc_presort.argtypes = [P(OrcHpx), C.c_uint32, vp, vp, vp, vp]
        L.orc_gather.restype = None
        L.orc_gather.argtypes = [C.c_uint32, vp, vp, vp]
        L.orc_ring_lut.restype = C.c_uint32
        L.orc_ring_lut.argtypes = [P(OrcHpx), C.c_uint32, vp, P(C.c_uint64), vp, C.c_uint32]
        L.orc_grid_compat.restype = C.c_int64
        L.orc_grid_compat.argtypes = [P(OrcHpx), P(OrcGrid), C.c_uint32, vp, vp, vp, vp, vp,
                                      C.c_uint64, C.c_uint32, vp, vp, vp, vp, vp, C.c_int]
        L.orc_grid_compat_rows.restype = C.c_int64
        L.orc_grid_compat_rows.argtypes = [P(OrcHpx), P(OrcGrid), C.c_uint32, vp, vp, vp, vp, vp,
                                           C.c_uint64, C.c_uint32, vp, vp, vp, vp, vp,
                                           C.c_uint32, C.c_uint32, C.c_int]
        L.orc_grid_compat_multi.restype = C.c_int64
        L.orc_grid_compat_multi.argtypes = [P(OrcHpx), P(OrcGrid), C.c_uint32, vp, vp, C.c_uint32, vp, vp,
                                            vp, C.c_uint64, C.c_uint32, vp, vp, vp, C.c_uint32, C.c_uint32,
                                            vp, vp, C.c_int]
        L.orc_grid_bruteforce.restype = None
        L.orc_grid_bruteforce.argtypes = [P(OrcHpx), P(OrcGrid), C.c_uint32, vp, vp, vp, vp, vp,
                                          C.c_uint64, C.c_uint32, vp, vp, vp, vp, C.c_int]
        L.orc_grid_disc_textbook.restype = None
        L.orc_grid_disc_textbook.argtypes = [C.c_uint32, C.c_uint32, vp, vp, vp, vp, vp,
                                             C.c_double, C.c_double, vp, vp, C.c_int]
        L.orc_kernel_weight.restype = C.c_double
        L.orc_kernel_weight.argtypes = [P(OrcGrid), C.c_double, C.c_double]
        L.orc_wcs_sin.restype = None
        L.orc_wcs_zenithal.restype = None
        L.orc_wcs_zenithal.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32] + [C.c_double] * 7 + [vp, vp]
        L.orc_wcs_sin.argtypes = [C.c_uint32, C.c_uint32] + [C.c_double] * 6 + [vp, vp]
        L.orc_wcs_celestial.restype = C.c_int
        L.orc_wcs_celestial.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double, vp] + \
            [C.c_double] * 6 + [vp, vp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def hpx(nside: int) -> OrcHpx:
    h = OrcHpx()
    lib().orc_hpx_init(C.byref(h), nside)
    return h


def set_optimal_nside(res_rad: float) -> int:
    return int(lib().orc_set_optimal_nside(res_rad))


def ang2pix(h: OrcHpx, theta, phi) -> np.ndarray:
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    out = np.empty(theta.shape, dtype=np.uint64)
    lib().orc_ang2pix_many(C.byref(h), theta.size, _p(theta), _p(phi), _p(out))
    return out


def pix2ang(h: OrcHpx, pix):
    pix = np.ascontiguousarray(pix, dtype=np.uint64)
    th = np.empty(pix.shape)
    ph = np.empty(pix.shape)
    lib().orc_pix2ang_many(C.byref(h), pix.size, _p(pix), _p(th), _p(ph))
    return th, ph


def pix2ring(h: OrcHpx, pix) -> np.ndarray:
    pix = np.ascontiguousarray(pix, dtype=np.uint64)
    out = np.empty(pix.shape, dtype=np.uint64)
    lib().orc_pix2ring_many(C.byref(h), pix.size, _p(pix), _p(out))
    return out


def ring_info(h: OrcHpx, rings):
    rings = np.ascontiguousarray(rings, dtype=np.uint64)
    sp = np.empty(rings.shape, dtype=np.uint64)
    npr = np.empty(rings.shape, dtype=np.uint64)
    lib().orc_ring_info_many(C.byref(h), rings.size, _p(rings), _p(sp), _p(npr))
    return sp, npr


def sample_keys(h: OrcHpx, lon_deg, lat_deg) -> np.ndarray:
    lon = np.ascontiguousarray(lon_deg, dtype=np.float64)
    lat = np.ascontiguousarray(lat_deg, dtype=np.float64)
    out = np.empty(lon.shape, dtype=np.uint64)
    lib().orc_sample_keys(C.byref(h), lon.size, _p(lon), _p(lat), _p(out))
    return out


def keys_libm_sensitive(h: OrcHpx, lon_deg, lat_deg, ulps: int = 3):
    """(count, flags u8[n]): samples whose key differs for some cos/sin
    result within +-ulps of this libm's (the only library calls of ang2pix)."""
    lon = np.ascontiguousarray(lon_deg, dtype=np.float64)
    lat = np.ascontiguousarray(lat_deg, dtype=np.float64)
    flags = np.zeros(lon.shape, dtype=np.uint8)
    cnt = lib().orc_sample_keys_libm_sensitive(C.byref(h), lon.size, _p(lon), _p(lat), int(ulps), _p(flags))
    return int(cnt), flags


def presort(h: OrcHpx, lon_deg, lat_deg):
    """(keys_sorted[n+1] u64 with npix sentinel, perm[n] u32)."""
    lon = np.ascontiguousarray(lon_deg, dtype=np.float64)
    lat = np.ascontiguousarray(lat_deg, dtype=np.float64)
    n = lon.size
    keys = np.empty(n + 1, dtype=np.uint64)
    perm = np.empty(n, dtype=np.uint32)
    lib().orc_presort(C.byref(h), n, _p(lon), _p(lat), _p(keys), _p(perm))
    return keys, perm


def ring_lut(h: OrcHpx, keys_sorted: np.ndarray, n: int):
    """(firstring, usedrings, lut[usedrings+1])."""
    keys = np.ascontiguousarray(keys_sorted, dtype=np.uint64)
    cap = int(h.nrings) + 2
    lut = np.zeros(cap, dtype=np.uint32)
    fr = C.c_uint64(0)
    used = lib().orc_ring_lut(C.byref(h), n, _p(keys), C.byref(fr), _p(lut), cap)
    return int(fr.value), int(used), lut[: used + 1].copy()


def grid_params(kernel_type: int, params, sphere_radius: float, disc_size: float, side: int,
                factor: int = 1) -> OrcGrid:
    g = OrcGrid()
    g.kernel_type = kernel_type
    g.factor = factor
    g.side = side
    g.bearing_needed = 1 if kernel_type == GAUSS2D else 0
    for i, p in enumerate(list(params)[:3]):
        g.params[i] = float(p)
    g.sphere_radius = sphere_radius
    g.disc_size = disc_size
    return g


def prepare_grid_kernel(kernel_type: int, kernel_params, sphere_radius: float, hpx_max_res: float,
                        side: int, factor: int = 1):
    """Restates _prepare_grid_kernel (HCGrid/gmap.cpp:64-99) -> (OrcHpx, OrcGrid)."""
    params = [float(x) for x in kernel_params] + [0.0] * (3 - len(kernel_params))
    if kernel_type == GAUSS1D:
        params[0] = 0.5 / params[0] ** 2  # pow(x, 2.) is x*x correctly rounded
    nside = set_optimal_nside(DEG2RAD * hpx_max_res)
    h = hpx(nside)
    disc = DEG2RAD * sphere_radius + h.resolution
    return h, grid_params(kernel_type, params, sphere_radius, disc, side, factor)


def grid_compat(h, g, lon_s, lat_s, data_s, weights_s, keys_sorted, firstring, usedrings, lut,
                xwcs, ywcs, datacube=None, weightscube=None, nthreads=0):
    """Run the restated kernel; returns (datacube, weightscube, endless_loop_rings)."""
    n = int(np.asarray(lon_s).size)
    ncells = int(g.side) * int(g.side)
    lon_s = np.ascontiguousarray(lon_s, dtype=np.float64)
    lat_s = np.ascontiguousarray(lat_s, dtype=np.float64)
    data_s = np.ascontiguousarray(data_s, dtype=np.float64)
    w = None if weights_s is None else np.ascontiguousarray(weights_s, dtype=np.float64)
    keys = np.ascontiguousarray(keys_sorted, dtype=np.uint64)
    lut = np.ascontiguousarray(lut, dtype=np.uint32)
    xw = np.ascontiguousarray(xwcs, dtype=np.float64)
    yw = np.ascontiguousarray(ywcs, dtype=np.float64)
    dc = np.zeros(ncells) if datacube is None else np.array(datacube, dtype=np.float64)
    wc = np.zeros(ncells) if weightscube is None else np.array(weightscube, dtype=np.float64)
    hangs = lib().orc_grid_compat(C.byref(h), C.byref(g), n, _p(lon_s), _p(lat_s), _p(data_s),
                                  None if w is None else _p(w), _p(keys), firstring, usedrings, _p(lut),
                                  _p(xw), _p(yw), _p(dc), _p(wc), nthreads)
    return dc, wc, int(hangs)


def grid_compat_rows(h, g, lon_s, lat_s, data_s, weights_s, keys_sorted, firstring, usedrings, lut,
                     xwcs, ywcs, row_begin, row_end, nthreads=0):
    """Restated kernel over rows [row_begin, row_end) only (zeroed cubes);
    returns (datacube, weightscube) for those rows, shape (rows*side,)."""
    n = int(np.asarray(lon_s).size)
    side = int(g.side)
    lon_s = np.ascontiguousarray(lon_s, dtype=np.float64)
    lat_s = np.ascontiguousarray(lat_s, dtype=np.float64)
    data_s = np.ascontiguousarray(data_s, dtype=np.float64)
    w = None if weights_s is None else np.ascontiguousarray(weights_s, dtype=np.float64)
    keys = np.ascontiguousarray(keys_sorted, dtype=np.uint64)
    lut = np.ascontiguousarray(lut, dtype=np.uint32)
    xw = np.ascontiguousarray(xwcs, dtype=np.float64)
    yw = np.ascontiguousarray(ywcs, dtype=np.float64)
    dc = np.zeros(side * side)
    wc = np.zeros(side * side)
    lib().orc_grid_compat_rows(C.byref(h), C.byref(g), n, _p(lon_s), _p(lat_s), _p(data_s),
                               None if w is None else _p(w), _p(keys), firstring, usedrings, _p(lut),
                               _p(xw), _p(yw), _p(dc), _p(wc), row_begin, row_end, nthreads)
    sl = slice(row_begin * side, row_end * side)
    return dc[sl].copy(), wc[sl].copy()


def grid_compat_multi(h, g, lon_s, lat_s, values_cm_s, weights_s, keys_sorted, firstring, usedrings, lut,
                      xwcs, ywcs, row_begin, row_end, nthreads=0):
    """Multi-channel restated kernel over rows [row_begin, row_end).

    values_cm_s: (nchan, n) float32 channel-major, in sorted order.
    Returns (sums (nchan, rows*side), wsum (rows*side,))."""
    v = np.ascontiguousarray(values_cm_s, dtype=np.float32)
    nchan, n = v.shape
    side = int(g.side)
    ncr = (row_end - row_begin) * side
    lon_s = np.ascontiguousarray(lon_s, dtype=np.float64)
    lat_s = np.ascontiguousarray(lat_s, dtype=np.float64)
    w = None if weights_s is None else np.ascontiguousarray(weights_s, dtype=np.float64)
    keys = np.ascontiguousarray(keys_sorted, dtype=np.uint64)
    lut = np.ascontiguousarray(lut, dtype=np.uint32)
    xw = np.ascontiguousarray(xwcs, dtype=np.float64)
    yw = np.ascontiguousarray(ywcs, dtype=np.float64)
    out = np.zeros((nchan, ncr))
    ws = np.zeros(ncr)
    lib().orc_grid_compat_multi(C.byref(h), C.byref(g), n, _p(lon_s), _p(lat_s), nchan, _p(v),
                                None if w is None else _p(w), _p(keys), firstring, usedrings, _p(lut),
                                _p(xw), _p(yw), row_begin, row_end, _p(out), _p(ws), nthreads)
    return out, ws


def grid_bruteforce(h, g, lon_s, lat_s, data_s, weights_s, keys_sorted, firstring, usedrings,
                    xwcs, ywcs, nthreads=0):
    n = int(np.asarray(lon_s).size)
    ncells = int(g.side) * int(g.side)
    lon_s = np.ascontiguousarray(lon_s, dtype=np.float64)
    lat_s = np.ascontiguousarray(lat_s, dtype=np.float64)
    data_s = np.ascontiguousarray(data_s, dtype=np.float64)
    w = None if weights_s is None else np.ascontiguousarray(weights_s, dtype=np.float64)
    keys = np.ascontiguousarray(keys_sorted, dtype=np.uint64)
    xw = np.ascontiguousarray(xwcs, dtype=np.float64)
    yw = np.ascontiguousarray(ywcs, dtype=np.float64)
    dc = np.zeros(ncells)
    wc = np.zeros(ncells)
    lib().orc_grid_bruteforce(C.byref(h), C.byref(g), n, _p(lon_s), _p(lat_s), _p(data_s),
                              None if w is None else _p(w), _p(keys), firstring, usedrings,
                              _p(xw), _p(yw), _p(dc), _p(wc), nthreads)
    return dc, wc


def grid_disc_textbook(lon, lat, data, xwcs, ywcs, sphere_radius, p0, nthreads=0):
    lon = np.ascontiguousarray(lon, dtype=np.float64)
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    data = np.ascontiguousarray(data, dtype=np.float64)
    xw = np.ascontiguousarray(xwcs, dtype=np.float64)
    yw = np.ascontiguousarray(ywcs, dtype=np.float64)
    dc = np.zeros(xw.size)
    wc = np.zeros(xw.size)
    lib().orc_grid_disc_textbook(xw.size, lon.size, _p(lon), _p(lat), _p(data), _p(xw), _p(yw),
                                 sphere_radius, p0, _p(dc), _p(wc), nthreads)
    return dc, wc


WCS_PROJ = {"SIN": 0, "TAN": 1, "ARC": 2, "ZEA": 3, "STG": 4}


def wcs_zenithal(proj, naxis1, naxis2, crpix1, crpix2, cdelt1, cdelt2, crval1, crval2, lonpole=180.0):
    xw = np.empty(naxis1 * naxis2)
    yw = np.empty(naxis1 * naxis2)
    lib().orc_wcs_zenithal(WCS_PROJ[proj], naxis1, naxis2, crpix1, crpix2, cdelt1, cdelt2, crval1, crval2,
                           lonpole, _p(xw), _p(yw))
    return xw, yw


def wcs_sin(naxis1, naxis2, crpix1, crpix2, cdelt1, cdelt2, crval1, crval2):
    xw = np.empty(naxis1 * naxis2)
    yw = np.empty(naxis1 * naxis2)
    lib().orc_wcs_sin(naxis1, naxis2, crpix1, crpix2, cdelt1, cdelt2, crval1, crval2, _p(xw), _p(yw))
    return xw, yw


WCS_CELESTIAL = dict(WCS_PROJ, CAR=5, MER=6, CEA=7, SFL=8, GLS=9, AIT=10)


def wcs_celestial(proj, naxis1, naxis2, crpix1, crpix2, m, crval1, crval2, lonpole=None, latpole=90.0,
                  pv=(0.0, 0.0)):
    """General celestial WCS (orc_wcs_celestial): m is the 2x2 matrix of
    degrees per pixel (CD, or CDELT_i PC_ij); lonpole None takes the
    default; pv = (PV2_1, PV2_2).  Returns (lon, lat) or None when the
    LONPOLE admits no native pole."""
    xw = np.empty(naxis1 * naxis2)
    yw = np.empty(naxis1 * naxis2)
    mm = np.ascontiguousarray(np.asarray(m, dtype=np.float64).reshape(4))
    rc = lib().orc_wcs_celestial(WCS_CELESTIAL[proj], naxis1, naxis2, crpix1, crpix2, _p(mm), crval1, crval2,
                                 float("nan") if lonpole is None else lonpole, latpole, pv[0], pv[1],
                                 _p(xw), _p(yw))
    return None if rc else (xw, yw)


def pipeline(lon, lat, data, xwcs, ywcs, side, sigma_deg, radius_deg, hpxres_deg, weights=None,
             factor=1, kernel_type=GAUSS1D, kernel_params=None, nthreads=0, rows=None):
    """Whole reference path on the CPU: keys, stable sort, gathers, LUT, grid.

    Returns a dict with keys_sorted, perm, firstring, usedrings, lut, datacube,
    weightscube, hangs and (h, g).  rows=(r0, r1) grids only that band of map
    rows (datacube/weightscube then hold those rows; hangs is not counted)."""
    if kernel_params is None:
        kernel_params = [sigma_deg]
    h, g = prepare_grid_kernel(kernel_type, kernel_params, radius_deg, hpxres_deg, side, factor)
    lon = np.ascontiguousarray(lon, dtype=np.float64)
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    data = np.ascontiguousarray(data, dtype=np.float64)
    keys, perm = presort(h, lon, lat)
    n = lon.size
    valid = int(np.searchsorted(keys[:n], np.uint64(h.npix), side="left"))
    fr, used, lut = ring_lut(h, keys[:valid], valid) if valid else (0, 0, np.zeros(1, np.uint32))
    if valid and valid < n:
        lut[-1] = n  # invalid keys sort last; the last ring's range ends at n
    lon_s, lat_s, data_s = lon[perm], lat[perm], data[perm]
    w_s = None if weights is None else np.asarray(weights, dtype=np.float64)[perm]
    if used and rows is not None:
        dc, wc = grid_compat_rows(h, g, lon_s, lat_s, data_s, w_s, keys, fr, used, lut, xwcs, ywcs,
                                  rows[0], rows[1], nthreads=nthreads)
        hangs = 0
    elif used:
        dc, wc, hangs = grid_compat(h, g, lon_s, lat_s, data_s, w_s, keys, fr, used, lut, xwcs, ywcs,
                                    nthreads=nthreads)
    else:
        dc, wc, hangs = np.zeros(side * side), np.zeros(side * side), 0
    return dict(h=h, g=g, keys_sorted=keys, perm=perm, firstring=fr, usedrings=used, lut=lut,
                lon_s=lon_s, lat_s=lat_s, data_s=data_s, weights_s=w_s, datacube=dc,
                weightscube=wc, hangs=hangs, n_valid=valid)
