"""Target deprojection on the device against the host form, and gridding onto
each zenithal target against the oracle.

Bar: NaN cells coincide exactly; coordinates within 1e-11 deg (the device's
sin/cos/atan2/asin are the HIP math library's; the host's are libm's).  The
gridding is projection-agnostic (it takes the cell coordinates), so the map
bar is the single-channel path's, 1e-6.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from hcgrid_amd import capi, gridder, synth

from helpers import map_errors, normalise

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("proj", sorted(capi.PROJECTIONS))
def test_device_matches_host(gpu, proj):
    for nx, ny, cdelt, crval, lonpole in ((300, 200, 0.01, (60.0, 30.0), 180.0), (64, 64, 9.3, (0.1, -45.0), 180.0),
                                          (41, 41, 9.3, (30.0, 90.0), 0.0)):
        print(f"case: proj={proj} {nx}x{ny} cdelt={cdelt} crval={crval}", flush=True)
        xh, yh = gridder.wcs(proj, nx, ny, nx / 2, ny / 2, -cdelt, cdelt, *crval, lonpole)
        xd = torch.empty(nx * ny, dtype=torch.float64, device="cuda")
        yd = torch.empty_like(xd)
        capi.call("hcg_wcs_p2s_device", capi.PROJECTIONS[proj], nx, ny, nx / 2, ny / 2, -cdelt, cdelt, *crval,
                  lonpole, C.c_void_p(xd.data_ptr()), C.c_void_p(yd.data_ptr()), None)
        torch.cuda.synchronize()
        xd, yd = xd.cpu().numpy(), yd.cpu().numpy()
        assert np.array_equal(np.isnan(xd), np.isnan(xh)) and np.array_equal(np.isnan(yd), np.isnan(yh))
        ok = ~np.isnan(yh) & (yh < 90.0 - 1e-9)  # the longitude of the pole itself is arbitrary
        dx = np.abs((xd[ok] - xh[ok] + 180.0) % 360.0 - 180.0)
        assert dx.max() < 1e-11 and np.abs(yd[ok] - yh[ok]).max() < 1e-11


@pytest.mark.parametrize("proj", ["TAN", "ARC", "ZEA", "STG"])
def test_gridding_onto_zenithal_targets(gpu, proj):
    wl = synth.raster_small(40, 160, 24, seed=41)
    t = wl.target
    xw, yw = gridder.wcs(proj, t.side, t.side, t.crpix, t.crpix, -t.cdelt, t.cdelt, t.crval1, t.crval2)
    xo, yo = oracle.wcs_zenithal(proj, t.side, t.side, t.crpix, t.crpix, -t.cdelt, t.cdelt, t.crval1, t.crval2)
    assert np.abs(xw - xo).max() < 1e-10 and np.abs(yw - yo).max() < 1e-10
    print(f"case: proj={proj} n={wl.n} side={t.side}", flush=True)
    dev = gridder.grid_plain(wl.lon, wl.lat, wl.values, xw, yw, t.side, 0.02, 0.06, 0.01)
    ref = oracle.pipeline(wl.lon, wl.lat, wl.values, xw, yw, t.side, 0.02, 0.06, 0.01)
    e = map_errors(dev["map"], normalise(ref["datacube"], ref["weightscube"]))
    print(f"  errors {e}", flush=True)
    assert e["nan_match"] and e["max_rel"] <= 1e-6 and e["cell_rel"] <= 1e-6
