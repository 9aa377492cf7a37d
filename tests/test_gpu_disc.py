"""HCG_SELECT_DISC: every sample within the radius of a cell, and nothing else.

The oracle is the long-double brute force over every (cell, sample) pair
(orc_grid_disc_textbook: true angular distance, exp(-d^2 p0)), which shares
no window logic with the device.  Bars: the single-channel map and weights
within 1e-6 (as the compat path); the multi-channel cube to the cube bar
(peak-normalised max relative error <= 1e-4).  Cases cover the longitude-0
wrap, both caps, a disc around the pole's neighbourhood, the first used ring
(which compat skips) and the east-west extent that compat truncates.
"""
import numpy as np
import pytest

import oracle
from hcgrid_amd import batch, gridder, synth

from helpers import map_errors, normalise

pytestmark = pytest.mark.gpu

SIGMA, RADIUS, RES = 0.02, 0.06, 0.01


def _field(dec, ra=180.0, n_lines=40, n_per=120, side=24, seed=31):
    wl = synth.raster_small(n_lines, n_per, side, dec_c=dec, ra_c=ra,
                            spacing_ra=0.0032 / max(np.cos(np.radians(dec)), 0.05), seed=seed)
    wl.lon[:] = np.mod(wl.lon, 360.0)
    t = wl.target
    xw, yw = gridder.wcs_sin(t.side, t.side, t.crpix, t.crpix, -t.cdelt, t.cdelt, t.crval1, t.crval2)
    return wl, xw, yw


def _textbook(wl, data, xw, yw, sigma=SIGMA, radius=RADIUS):
    return oracle.grid_disc_textbook(wl.lon, wl.lat, data, xw, yw, radius, 0.5 / sigma ** 2)


@pytest.mark.parametrize("dec,ra", [(0.0, 180.0), (30.0, 180.0), (60.0, 180.0), (80.0, 180.0), (-85.0, 180.0),
                                    (89.3, 180.0), (30.0, 0.05), (80.0, 0.3)])
def test_disc_plain_matches_textbook(gpu, dec, ra):
    wl, xw, yw = _field(dec, ra)
    print(f"case: disc plain dec={dec} ra={ra} n={wl.n}", flush=True)
    dev = gridder.grid_plain(wl.lon, wl.lat, wl.values, xw, yw, wl.target.side, SIGMA, RADIUS, RES,
                             selection="disc")
    dc, wc = _textbook(wl, wl.values, xw, yw)
    e = map_errors(dev["map"], normalise(dc, wc))
    ew = map_errors(dev["weightscube"], wc)
    print(f"  map={e} weights={ew}", flush=True)
    assert e["nan_match"]
    assert e["max_rel"] <= 1e-6 and e["cell_rel"] <= 1e-6
    assert ew["max_rel"] <= 1e-6


def test_disc_around_the_pole(gpu):
    # cells within the radius of the north pole: whole rings near the pole
    wl, xw, yw = _field(89.85, n_lines=10, n_per=160)
    assert wl.lat.max() < 90.0 and yw.max() > 90.0 - RADIUS
    print(f"case: disc plain around the pole n={wl.n}", flush=True)
    dev = gridder.grid_plain(wl.lon, wl.lat, wl.values, xw, yw, wl.target.side, SIGMA, RADIUS, RES,
                             selection="disc")
    dc, wc = _textbook(wl, wl.values, xw, yw)
    e = map_errors(dev["map"], normalise(dc, wc))
    ew = map_errors(dev["weightscube"], wc)
    print(f"  map={e} weights={ew}", flush=True)
    assert e["nan_match"] and e["cells"] > 0
    assert e["max_rel"] <= 1e-6 and ew["max_rel"] <= 1e-6


def test_disc_is_a_superset_of_compat(gpu):
    # at dec 60 compat truncates the east-west extent (no 1/cos dec) and skips
    # the first used ring; disc keeps both, so its weights are never smaller
    # and are larger somewhere
    wl, xw, yw = _field(60.0)
    side = wl.target.side
    comp = gridder.grid_plain(wl.lon, wl.lat, wl.values, xw, yw, side, SIGMA, RADIUS, RES)
    disc = gridder.grid_plain(wl.lon, wl.lat, wl.values, xw, yw, side, SIGMA, RADIUS, RES, selection="disc")
    wcomp, wdisc = comp["weightscube"], disc["weightscube"]
    assert np.all(wdisc >= wcomp * (1 - 1e-12))
    assert np.any(wdisc > wcomp * (1 + 1e-6))


def test_disc_ignores_the_coarsening_factor(gpu):
    wl, xw, yw = _field(30.0)
    side = wl.target.side
    a = gridder.grid_plain(wl.lon, wl.lat, wl.values, xw, yw, side, SIGMA, RADIUS, RES, selection="disc")
    b = gridder.grid_plain(wl.lon, wl.lat, wl.values, xw, yw, side, SIGMA, RADIUS, RES, selection="disc",
                           factor=3)
    assert np.array_equal(a["map"], b["map"], equal_nan=True)


def _values(n, nchan, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((nchan, n)).astype(np.float32)


@pytest.mark.parametrize("dec,ra", [(30.0, 180.0), (80.0, 180.0), (30.0, 0.05)])
def test_disc_cube_matches_textbook(gpu, dec, ra):
    import torch
    wl, xw, yw = _field(dec, ra)
    nchan = 4
    Vt = torch.tensor(_values(wl.n, nchan, 5)).to(torch.bfloat16).cuda()
    Vr = Vt.float().cpu().numpy().astype(np.float64)
    print(f"case: disc cube dec={dec} ra={ra} n={wl.n} nchan={nchan}", flush=True)
    dev = gridder.grid_cube(wl.lon, wl.lat, Vt, xw, yw, wl.target.side, SIGMA, RADIUS, RES, selection="disc")
    for c in range(nchan):
        dc, wc = _textbook(wl, Vr[c], xw, yw)
        e = map_errors(dev["cube"][c].reshape(-1), normalise(dc, wc))
        print(f"  channel {c}: {e}", flush=True)
        assert e["nan_match"] and e["max_rel"] <= 1e-4


def test_disc_job_in_a_batch(gpu):
    # a disc job and a compat job on the same field in one persistent launch:
    # each equals its own one-job launch bit for bit, and the disc job the
    # textbook to the cube bar
    import torch
    wl, xw, yw = _field(45.0, n_lines=50, n_per=150, side=28)
    side, nchan = wl.target.side, 8
    V = torch.tensor(_values(wl.n, nchan, 9)).to(torch.bfloat16).cuda()
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")  # noqa: E731
    jobs, geos = [], []
    for sel in ("disc", "compat"):
        kern, hp = gridder.prepare_grid_kernel(1, [SIGMA], RADIUS, RES, selection=sel)
        g = batch.Geometry(kern, hp, wl.n, side)
        g.build(d(wl.lon), d(wl.lat), d(xw), d(yw))
        geos.append(g)
        jobs.append(dict(geometry=g, values=V, dtype="bf16",
                         cube=torch.empty((nchan, side * side), dtype=torch.float32, device="cuda"),
                         wsum=torch.empty(side * side, dtype=torch.float32, device="cuda")))
    b = batch.Batch(jobs)
    b.enqueue()
    torch.cuda.synchronize()
    for j in jobs:
        alone = dict(j, cube=torch.empty_like(j["cube"]), wsum=torch.empty_like(j["wsum"]))
        a = batch.Batch([alone])
        a.enqueue()
        torch.cuda.synchronize()
        assert np.array_equal(alone["cube"].cpu().numpy(), j["cube"].cpu().numpy(), equal_nan=True)
        assert torch.equal(alone["wsum"], j["wsum"])
    Vr = V.float().cpu().numpy().astype(np.float64)
    cube = jobs[0]["cube"].cpu().numpy()
    for c in range(nchan):
        dc, wc = _textbook(wl, Vr[c], xw, yw)
        e = map_errors(cube[c], normalise(dc, wc))
        assert e["nan_match"] and e["max_rel"] <= 1e-4, e
    assert not np.array_equal(jobs[0]["wsum"].cpu().numpy(), jobs[1]["wsum"].cpu().numpy())
