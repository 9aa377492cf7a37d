"""The C ABI library on the CPU: it loads, exports every symbol the header
declares, and its host-only set-up functions agree with the oracle.  No
device compute is launched here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from hcgrid_amd import capi, gridder

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    declared = capi.header_symbols()
    assert len(declared) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (hcg_\w+)", out))
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    for s in declared:
        getattr(lib, s)
    # every declared function has a ctypes prototype
    assert sorted(set(declared) - set(capi.PROTOTYPES)) == []


def test_strerror_and_codes():
    lib = capi.load()
    for code in (0, -1, -2, -3, -4, -5, -6):
        assert lib.hcg_strerror(code)
    assert lib.hcg_strerror(-99) == b"unknown error"


@pytest.mark.parametrize("nside", [1, 2, 64, 4096, 8192, 16384])
def test_healpix_init_matches_oracle(nside):
    hp = capi.Healpix()
    capi.call("hcg_healpix_init", nside, C.byref(hp))
    h = O.hpx(nside)
    for f in ("nside", "order", "nrings", "npface", "ncap", "npix", "fact1", "fact2", "omega", "resolution"):
        assert getattr(hp, f) == getattr(h, f), f


def test_healpix_init_rejects_bad_nside():
    hp = capi.Healpix()
    lib = capi.load()
    assert lib.hcg_healpix_init(3, C.byref(hp)) == capi.HCG_ERR_INVALID
    assert lib.hcg_healpix_init(0, C.byref(hp)) == capi.HCG_ERR_INVALID


def test_set_optimal_nside_matches_oracle():
    lib = capi.load()
    for res in np.geomspace(1e-5, 0.1, 300):
        assert lib.hcg_set_optimal_nside(res) == O.set_optimal_nside(res)


@pytest.mark.parametrize("kt,params", [(capi.GAUSS1D, [0.02]), (capi.GAUSS1D, [300 / 3600 / 2.3548200450309493]),
                                       (capi.GAUSS2D, [0.02, 0.01, 0.5]), (capi.TAPERED_SINC, [0.02, 1.55, 2.52])])
def test_prepare_grid_kernel_matches_oracle(kt, params):
    sigma = params[0]
    kern, hp = gridder.prepare_grid_kernel(kt, params, 3 * sigma, sigma / 2)
    h, g = O.prepare_grid_kernel(kt, params, 3 * sigma, sigma / 2, 16)
    assert hp.nside == h.nside
    assert kern.disc_size == g.disc_size
    assert kern.sphere_radius == g.sphere_radius
    assert list(kern.params) == list(g.params)
    assert kern.bearing_needed == (1 if kt == capi.GAUSS2D else 0)
    assert kern.factor == 1


def test_prepare_grid_kernel_rejects_unknown_type():
    lib = capi.load()
    params = (C.c_double * 3)(0.02, 0, 0)
    k, h = capi.Kernel(), capi.Healpix()
    assert lib.hcg_prepare_grid_kernel(7, params, 0.06, 0.01, C.byref(k), C.byref(h)) == capi.HCG_ERR_INVALID
    assert b"unknown kernel type" in lib.hcg_last_error_message()


def test_selection_mode_plumbing():
    kern, _ = gridder.prepare_grid_kernel(capi.GAUSS1D, [0.02], 0.06, 0.01)
    assert kern.selection == capi.SELECT_COMPAT
    kern, hp = gridder.prepare_grid_kernel(capi.GAUSS1D, [0.02], 0.06, 0.01, selection="disc")
    assert kern.selection == capi.SELECT_DISC
    with pytest.raises(capi.HcgError):
        gridder.prepare_grid_kernel(capi.GAUSS1D, [0.02], 0.06, 0.01, selection="box")
    lib = capi.load()
    kern.selection = 2
    assert lib.hcg_grid_plain(C.byref(hp), C.byref(kern), 4, 1, *([None] * 14)) == capi.HCG_ERR_INVALID
    assert b"selection" in lib.hcg_last_error_message()


def test_wcs_matches_oracle():
    for side, crval in ((90, (60.0, 30.0)), (400, (180.0, 30.0)), (64, (0.1, -45.0)), (50, (359.9, 80.0))):
        xw, yw = gridder.wcs_sin(side, side, side / 2, side / 2, -0.01, 0.01, *crval)
        xo, yo = O.wcs_sin(side, side, side / 2, side / 2, -0.01, 0.01, *crval)
        dx = np.abs((xw - xo + 180.0) % 360.0 - 180.0)
        assert dx.max() < 1e-10 and np.abs(yw - yo).max() < 1e-10


@pytest.mark.parametrize("proj", sorted(capi.PROJECTIONS))
def test_wcs_zenithal_matches_oracle(proj):
    # every projection, a polar target (default LONPOLE 0) and, with a wide
    # CDELT, cells outside SIN/ARC/ZEA's domains (NaN on both sides)
    for side, cdelt, crval in ((90, 0.01, (60.0, 30.0)), (64, 0.3, (0.1, -45.0)), (41, 9.3, (30.0, 90.0))):
        xw, yw = gridder.wcs(proj, side, side, side / 2, side / 2, -cdelt, cdelt, *crval)
        lonpole = 0.0 if crval[1] >= 90.0 else 180.0
        xo, yo = O.wcs_zenithal(proj, side, side, side / 2, side / 2, -cdelt, cdelt, *crval, lonpole)
        assert np.array_equal(np.isnan(xw), np.isnan(xo)) and np.array_equal(np.isnan(yw), np.isnan(yo))
        ok = ~np.isnan(yo) & (yo < 90.0 - 1e-9)  # the longitude of the pole itself is arbitrary
        dx = np.abs((xw[ok] - xo[ok] + 180.0) % 360.0 - 180.0)
        assert dx.max() < 1e-9 and np.abs(yw[ok] - yo[ok]).max() < 1e-10
    with pytest.raises(ValueError):
        gridder.wcs("CAR", 4, 4, 2, 2, -1, 1, 0, 0)
    lib = capi.load()
    buf = np.empty(4)
    assert lib.hcg_wcs_p2s(5, 2, 2, 1.0, 1.0, -1.0, 1.0, 0.0, 0.0, 180.0, buf.ctypes.data, buf.ctypes.data) == \
        capi.HCG_ERR_INVALID
    assert b"unknown projection" in lib.hcg_last_error_message()


def test_argument_errors_do_not_touch_the_device():
    lib = capi.load()
    hp = capi.Healpix()
    capi.call("hcg_healpix_init", 8192, C.byref(hp))
    assert lib.hcg_presort(None, 10, None, None, None, None, None, 0, None) == capi.HCG_ERR_INVALID
    assert lib.hcg_presort(C.byref(hp), 10, None, None, None, None, None, 0, None) == capi.HCG_ERR_INVALID
    big = capi.Healpix()
    capi.call("hcg_healpix_init", 16384, C.byref(big))
    one = C.c_void_p(1)
    assert lib.hcg_presort(C.byref(big), 10, one, one, one, one, one, 1 << 30, None) == capi.HCG_ERR_UNSUPPORTED
    assert lib.hcg_presort(C.byref(hp), 10, one, one, one, one, one, 1, None) == capi.HCG_ERR_WORKSPACE
    kern, hp2 = gridder.prepare_grid_kernel(capi.GAUSS1D, [0.02], 0.06, 0.01)
    kern.kernel_type = 9
    assert lib.hcg_grid_plain(C.byref(hp2), C.byref(kern), 4, 1, *([None] * 14)) == capi.HCG_ERR_INVALID
    assert lib.hcg_presort_workspace_size(20_000_000) >= 2 * 4 * 20_000_000


def test_product_code_never_uses_the_oracle():
    pat = re.compile(r"^\s*(import\s+oracle|from\s+oracle)|liboracle|hcg_oracle", re.M)
    for d in ("hcgrid_amd", "include"):
        for dirpath, _, files in os.walk(os.path.join(ROOT, d)):
            for f in files:
                if f.endswith((".py", ".hip", ".hpp", ".h", ".cpp", "Makefile")):
                    text = open(os.path.join(dirpath, f), errors="replace").read()
                    assert not pat.search(text), f"{dirpath}/{f} references the oracle"
