"""Multi-channel (spectral cube) path against the FP64 oracle.

The oracle applies FP64 weights of the reference's selection to the same
channel values (already rounded to bf16/f16) -- what the reference's
channel-by-channel pipeline computes for every channel.  The device carries
each weight as hi + lo in the channel type (relative error ~2^-17 for BF16)
and accumulates in FP32.  Bar: peak-normalised max relative error <= 1e-4
and RMS relative error <= 1e-4 per cube (the product target is 1e-3), NaN
cells identical, weight sums within 2^-14 relative.
"""
import numpy as np
import pytest

import oracle
from hcgrid_amd import gridder, synth

from helpers import map_errors

pytestmark = pytest.mark.gpu


def _values(wl, nchan, seed):
    rng = np.random.default_rng(seed)
    ch = np.arange(nchan)[:, None]
    mu, w = rng.uniform(0.2, 0.8) * nchan, max(nchan / 20.0, 1.0)
    profile = np.exp(-((ch - mu) ** 2) / (2 * w * w))
    return (wl.values[None, :].astype(np.float64) * (0.3 + profile) +
            rng.normal(0.0, 0.5, (nchan, wl.n))).astype(np.float32)


def _run_case(n_lines, n_per, side, nchan, dtype, layout, seed, dec=30.0, tol=None, shape=None):
    import torch
    wl = synth.raster_small(n_lines, n_per, side, dec_c=dec, seed=seed)
    t = wl.target
    xw, yw = gridder.wcs_sin(side, side, t.crpix, t.crpix, -t.cdelt, t.cdelt, t.crval1, t.crval2)
    V = _values(wl, nchan, seed)
    tdt = torch.bfloat16 if dtype == "bf16" else torch.float16
    Vt = torch.tensor(V).to(tdt).cuda()
    Vr = Vt.float().cpu().numpy().astype(np.float32)  # the rounded values both sides use
    print(f"case: n={wl.n} side={side} nchan={nchan} dtype={dtype} layout={layout} shape={shape}", flush=True)
    dev = gridder.grid_cube(wl.lon, wl.lat, Vt if layout == 0 else Vt.t().contiguous(), xw, yw, side, wl.sigma,
                            wl.radius, wl.hpx_res, dtype=dtype, layout=layout, shape=shape)
    ref = oracle.pipeline(wl.lon, wl.lat, Vr[0], xw, yw, side, wl.sigma, wl.radius, wl.hpx_res)
    assert np.array_equal(dev["perm"], ref["perm"])
    out, ws = oracle.grid_compat_multi(ref["h"], ref["g"], ref["lon_s"], ref["lat_s"], Vr[:, ref["perm"]], None,
                                       ref["keys_sorted"], ref["firstring"], ref["usedrings"], ref["lut"], xw, yw,
                                       0, side)
    with np.errstate(invalid="ignore", divide="ignore"):
        o = out / ws[None, :]
    e = map_errors(dev["cube"].reshape(-1), o.reshape(-1))
    wrel = np.abs(dev["wsum"] - ws)[ws > 0] / ws[ws > 0]
    print(f"  errors {e} wsum_rel_max {wrel.max() if wrel.size else 0:.2e}", flush=True)
    assert dev["nonfinite"] == 0
    assert e["nan_match"]
    tol = tol or 1e-4
    assert e["max_rel"] <= tol and e["rms_rel"] <= 1e-4
    assert wrel.size > 0 and wrel.max() <= 2.0 ** -14
    if nchan == 1:  # the FP64 tile path: unrounded weights
        assert e["max_rel"] <= 1e-6 and wrel.max() <= 1e-6
    elif nchan <= 8 and side >= 512 and wl.n * 10 <= side * side:
        # few-channel tiles (sparse maps): FP64 weight sums, FP32 channel sums
        assert wrel.max() <= 1e-6
    return dev, o


@pytest.mark.parametrize("nchan", [1, 48, 64, 130])
def test_cube_bf16_channel_counts(gpu, nchan):
    _run_case(40, 200, 24, nchan, "bf16", 0, seed=nchan)


@pytest.mark.parametrize("nchan,dtype", [(8, "bf16"), (8, "f16"), (6, "bf16"), (16, "bf16")])
def test_few_channel_tile_route(gpu, nchan, dtype):
    # 2..8 channels on a sparse 512^2 map (12000 samples): the tile loop with
    # one FP32 sum per channel (6 channels: the unused slots of the 8 stay out
    # of the cube); 16 channels on the same map take the contraction
    dev, _ = _run_case(40, 300, 512, nchan, dtype, 0, seed=78 + nchan)
    assert np.isnan(dev["cube"]).sum() > 0


@pytest.mark.parametrize("dtype", ["bf16", "f16"])
def test_one_channel_tile_route(gpu, dtype):
    # one channel on a 512^2 map (4096 tiles, one wave each): the FP64 tile
    # loop instead of the contraction; the samples cover part of the map, the
    # rest is NaN.  (Smaller maps split each tile over four ring streams:
    # test_cube_bf16_channel_counts[1].)
    dev, _ = _run_case(40, 300, 512, 1, dtype, 0, seed=77)
    assert np.isnan(dev["cube"]).sum() > 0


@pytest.mark.parametrize("shape", [(8, 1, 128), (16, 2, 256), (32, 4, 512), (64, 8, 128), (16, 8, 512),
                                   (8, 4, 256), (64, 1, 512)])
def test_cube_launch_shapes(gpu, shape):
    # every block shape, wave count and chunk; side 27 leaves partial blocks
    _run_case(30, 160, 27, 100, "bf16", 0, seed=23, shape=shape)


def test_cube_chunk_and_waves_change_no_bit(gpu):
    # a dense raster (rings of hundreds of samples, several chunks per block
    # at every chunk size): with the block shape fixed, the chunk size and
    # the wave count change nothing, since chunks end on k-block boundaries
    import torch
    wl = synth.raster_small(40, 2400, 24, spacing_ra=0.0005, seed=17)
    t = wl.target
    xw, yw = gridder.wcs_sin(24, 24, t.crpix, t.crpix, -t.cdelt, t.cdelt, t.crval1, t.crval2)
    Vt = torch.tensor(_values(wl, 96, 17)).to(torch.bfloat16).cuda()
    kern, hp = gridder.prepare_grid_kernel(1, [wl.sigma], wl.radius, wl.hpx_res)
    cg = gridder.CubeGridder(wl.n, 24, 96, kern, hp, dtype="bf16")
    d = lambda a: torch.as_tensor(a, device="cuda")  # noqa: E731
    xwd, ywd = d(xw), d(yw)
    cg.run(d(wl.lon), d(wl.lat), Vt, xwd, ywd)
    got = {}
    for shp in [(16, 1, 128), (16, 2, 256), (16, 4, 256), (16, 8, 512), (16, 4, 128)]:
        print(f"case: n={wl.n} side=24 nchan=96 shape={shp}", flush=True)
        cg.grid(xwd, ywd, shape=shp)
        torch.cuda.synchronize()
        got[shp] = (cg.cube.cpu().numpy().copy(), cg.wsum.cpu().numpy().copy())
    c0, w0 = got[(16, 8, 512)]
    assert np.isfinite(c0).sum() > 1000
    for shp, (c, w) in got.items():
        assert np.array_equal(c, c0, equal_nan=True), f"shape {shp} changes the cube"
        assert np.array_equal(w, w0), f"shape {shp} changes the weight sums"


def test_cube_f16(gpu):
    _run_case(40, 200, 24, 96, "f16", 0, seed=3)


def test_cube_sample_major_layout(gpu):
    _run_case(40, 200, 24, 80, "bf16", 1, seed=5)


def test_cube_many_channels_two_groups(gpu):
    # 600 channels: 10 waves of 64 -> two workgroups per row segment, last block partial
    _run_case(30, 150, 20, 600, "bf16", 0, seed=7)


@pytest.mark.parametrize("dec", [-78.0, -40.0, 75.0, 88.5])
def test_cube_declinations(gpu, dec):
    # |dec| >= 70 deg: the window anchors sit beyond NORTH_B / SOUTH_B,
    # whole-ring windows narrowed to the disc's longitude range
    _run_case(40, 120, 24, 32, "bf16", 0, seed=11, dec=dec)


def test_cube_denser_than_a_chunk(gpu):
    # dense sampling: a ring's union window exceeds one 512-sample chunk
    import torch
    wl = synth.raster_small(20, 2400, 16, spacing_ra=0.0002, seed=13)
    t = wl.target
    xw, yw = gridder.wcs_sin(16, 16, t.crpix, t.crpix, -t.cdelt, t.cdelt, t.crval1, t.crval2)
    V = _values(wl, 16, 13)
    Vt = torch.tensor(V).to(torch.bfloat16).cuda()
    Vr = Vt.float().cpu().numpy()
    dev = gridder.grid_cube(wl.lon, wl.lat, Vt, xw, yw, 16, wl.sigma, wl.radius, wl.hpx_res)
    ref = oracle.pipeline(wl.lon, wl.lat, Vr[0], xw, yw, 16, wl.sigma, wl.radius, wl.hpx_res)
    out, ws = oracle.grid_compat_multi(ref["h"], ref["g"], ref["lon_s"], ref["lat_s"], Vr[:, ref["perm"]], None,
                                       ref["keys_sorted"], ref["firstring"], ref["usedrings"], ref["lut"], xw, yw,
                                       0, 16)
    with np.errstate(invalid="ignore", divide="ignore"):
        o = out / ws[None, :]
    e = map_errors(dev["cube"].reshape(-1), o.reshape(-1))
    print(e, flush=True)
    assert e["nan_match"] and e["max_rel"] <= 1e-4


def test_cube_matches_plain_path_channel_by_channel(gpu):
    # the reference's semantics: every channel is a single-channel gridding
    import torch
    wl = synth.raster_small(40, 200, 24, seed=17)
    t = wl.target
    xw, yw = gridder.wcs_sin(24, 24, t.crpix, t.crpix, -t.cdelt, t.cdelt, t.crval1, t.crval2)
    V = _values(wl, 8, 17)
    Vt = torch.tensor(V).to(torch.bfloat16).cuda()
    Vr = Vt.float().cpu().numpy()
    cube = gridder.grid_cube(wl.lon, wl.lat, Vt, xw, yw, 24, wl.sigma, wl.radius, wl.hpx_res)["cube"]
    for ch in range(8):
        plain = gridder.grid_plain(wl.lon, wl.lat, Vr[ch], xw, yw, 24, wl.sigma, wl.radius, wl.hpx_res)["map"]
        e = map_errors(cube[ch], plain)
        assert e["nan_match"] and e["max_rel"] <= 1e-4


def test_cube_full_size_band(gpu):
    # 2M samples x 1024 BF16 channels onto 1024 x 1024: the whole cube on the
    # device, a band of 4 rows x every 8th channel against the FP64 oracle
    import torch
    wl = synth.raster_cube(device="cuda")
    t = wl.target
    side = t.side
    xw, yw = gridder.wcs_sin(side, side, t.crpix, t.crpix, -t.cdelt, t.cdelt, t.crval1, t.crval2)
    kern, hp = gridder.prepare_grid_kernel(1, [wl.sigma], wl.radius, wl.hpx_res)
    g = gridder.CubeGridder(wl.n, side, wl.nchan, kern, hp, dtype="bf16")
    dev = g.plain.device
    tt = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)  # noqa: E731
    print(f"case: full-size cube n={wl.n} nchan={wl.nchan} side={side}", flush=True)
    g.run(tt(wl.lon), tt(wl.lat), wl.values, tt(xw), tt(yw))
    torch.cuda.synchronize()
    assert int(g.nonfinite.item()) == 0
    r0, r1 = 510, 514
    chans = torch.arange(0, wl.nchan, 8, device=dev)
    band = g.cube.view(wl.nchan, side, side)[chans][:, r0:r1, :].cpu().numpy().reshape(len(chans), -1)
    wsum = g.wsum.view(side, side)[r0:r1].cpu().numpy().reshape(-1)
    Vr = wl.values[chans].float().cpu().numpy()
    ref = oracle.pipeline(wl.lon, wl.lat, Vr[0], xw, yw, side, wl.sigma, wl.radius, wl.hpx_res,
                          rows=(r0, r1))
    assert np.array_equal(g.plain.perm(), ref["perm"])
    out, ws = oracle.grid_compat_multi(ref["h"], ref["g"], ref["lon_s"], ref["lat_s"], Vr[:, ref["perm"]], None,
                                       ref["keys_sorted"], ref["firstring"], ref["usedrings"], ref["lut"], xw, yw,
                                       r0, r1)
    with np.errstate(invalid="ignore", divide="ignore"):
        o = out / ws[None, :]
    e = map_errors(band.reshape(-1), o.reshape(-1))
    wrel = np.abs(wsum - ws)[ws > 0] / ws[ws > 0]
    print(f"  errors {e} wsum_rel_max {wrel.max():.2e}", flush=True)
    assert e["nan_match"] and e["max_rel"] <= 1e-4 and e["rms_rel"] <= 1e-4
    assert wrel.max() <= 2.0 ** -14


@pytest.mark.parametrize("sigma,radius,shape", [(0.02, 0.06, (16, 0, 0)), (0.1, 0.3, (8, 4, 256)),
                                                (0.1, 0.3, (64, 2, 128))])
def test_cube_tile_plan_equals_search(gpu, sigma, radius, shape):
    # the tile plan holds 64 rings per tile; sigma 0.1 deg at nside 8192 spans
    # ~70-120 rings per tile, so planned and searched rings mix
    import torch
    wl = synth.raster_small(40, 200, 24, seed=31)
    t = wl.target
    xw, yw = gridder.wcs_sin(24, 24, t.crpix, t.crpix, -t.cdelt, t.cdelt, t.crval1, t.crval2)
    V = _values(wl, 40, 31)
    Vt = torch.tensor(V).to(torch.bfloat16).cuda()
    kern, hp = gridder.prepare_grid_kernel(1, [sigma], radius, 0.01)
    g = gridder.CubeGridder(wl.n, 24, 40, kern, hp)
    d = lambda a: torch.as_tensor(a, dtype=torch.float64, device="cuda")  # noqa: E731
    xd, yd = d(xw), d(yw)
    print(f"case: plan vs search sigma={sigma} shape={shape} nside={hp.nside}", flush=True)
    g.run(d(wl.lon), d(wl.lat), Vt, xd, yd, shape=shape)
    torch.cuda.synchronize()
    planned = g.cube.cpu().numpy().copy()
    g.grid(xd, yd, shape=shape, plan=False)
    torch.cuda.synchronize()
    assert np.array_equal(planned, g.cube.cpu().numpy(), equal_nan=True)
    Vr = Vt.float().cpu().numpy()
    ref = oracle.pipeline(wl.lon, wl.lat, Vr[0], xw, yw, 24, sigma, radius, 0.01)
    out, ws = oracle.grid_compat_multi(ref["h"], ref["g"], ref["lon_s"], ref["lat_s"], Vr[:, ref["perm"]], None,
                                       ref["keys_sorted"], ref["firstring"], ref["usedrings"], ref["lut"], xw, yw,
                                       0, 24)
    with np.errstate(invalid="ignore", divide="ignore"):
        o = out / ws[None, :]
    e = map_errors(planned.reshape(-1), o.reshape(-1))
    print(f"  errors {e}", flush=True)
    assert e["nan_match"] and e["max_rel"] <= 1e-4


@pytest.mark.parametrize("n,nchan,dtype,layout", [(1000, 5, "bf16", 0), (777, 33, "f16", 0), (1000, 70, "bf16", 1),
                                                  (64, 1, "f16", 0), (4100, 40, "f16", 1)])
def test_gather_channels_exact(gpu, n, nchan, dtype, layout):
    # the blocked layout, element for element: slot i of channel ch at
    # ((i // 8) * nchan + ch) * 8 + i % 8, zeros past n, and a count of the
    # non-finite values among the n samples (planted NaNs and infinities)
    import ctypes as C

    import torch

    from hcgrid_amd import capi
    rng = np.random.default_rng(n + nchan)
    perm = rng.permutation(n).astype(np.uint32)
    v = rng.standard_normal((nchan, n)).astype(np.float32)
    bad = [(0, 3), (nchan - 1, n - 1), (nchan // 2, n // 2)]
    for k, (ch, i) in enumerate(bad):
        v[ch, i] = np.nan if k % 2 == 0 else np.inf
    tdt = torch.bfloat16 if dtype == "bf16" else torch.float16
    vt = torch.tensor(v).to(tdt)
    src = (vt if layout == 0 else vt.t()).contiguous().cuda()
    n_pad = capi.load().hcg_channels_padded_len(n)
    out = torch.full((n_pad * nchan,), 7.0, dtype=tdt, device="cuda")  # every element must be written
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    pd = torch.tensor(perm.astype(np.int32), device="cuda")
    print(f"case: gather n={n} nchan={nchan} {dtype} layout={layout}", flush=True)
    capi.call("hcg_gather_channels", n, nchan, C.c_void_p(pd.data_ptr()), C.c_void_p(src.data_ptr()),
              capi.DTYPE_BF16 if dtype == "bf16" else capi.DTYPE_F16, layout, C.c_void_p(out.data_ptr()),
              C.c_void_p(cnt.data_ptr()), None)
    torch.cuda.synchronize()
    got = out.view(torch.int16).cpu().numpy()
    bits = vt.view(torch.int16).numpy()  # [nchan][n]
    want = np.zeros(n_pad * nchan, dtype=np.int16)
    i = np.arange(n)
    for ch in range(nchan):
        want[((i // 8) * nchan + ch) * 8 + i % 8] = bits[ch, perm]
    assert np.array_equal(got, want)
    assert int(cnt.item()) == len(bad)


@pytest.mark.parametrize("side,nchan", [(24, 1), (512, 1), (512, 8), (24, 48)])
def test_raw_sums_match_normalised(gpu, side, nchan):
    # normalize = 0 on every route (one channel on 4 ring streams and on one
    # wave per tile, few-channel tiles, the MFMA contraction): the raw cube
    # is the normalised cube times the weight sum, and 0 where no sample
    # contributes (NaN once normalised); the weight sum is the same either way
    import torch

    from hcgrid_amd import GAUSS1D
    wl = synth.raster_small(40, 300 if side >= 512 else 200, side, seed=90 + nchan)
    t = wl.target
    xw, yw = gridder.wcs_sin(side, side, t.crpix, t.crpix, -t.cdelt, t.cdelt, t.crval1, t.crval2)
    V = torch.tensor(_values(wl, nchan, 5)).to(torch.bfloat16).cuda()
    kern, hp = gridder.prepare_grid_kernel(GAUSS1D, [wl.sigma], wl.radius, wl.hpx_res)
    g = gridder.CubeGridder(wl.n, side, nchan, kern, hp, dtype="bf16")
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")  # noqa: E731
    print(f"case: raw sums side={side} nchan={nchan}", flush=True)
    out = []
    for norm in (True, False):
        g.run(d(wl.lon), d(wl.lat), V, d(xw), d(yw), normalize=norm)
        torch.cuda.synchronize()
        out.append((g.cube.float().cpu().numpy().reshape(nchan, -1).copy(), g.wsum.cpu().numpy().copy()))
    (cn, wn), (cr, wr) = out
    assert np.array_equal(wn, wr)
    has = wn > 0
    assert has.any()
    assert np.all(np.isnan(cn[:, ~has])) and np.all(cr[:, ~has] == 0)  # empty cells (the 512^2 maps)
    if side >= 512:
        assert (~has).any()
    scale = np.abs(cr[:, has]).max()
    np.testing.assert_allclose(cr[:, has], cn[:, has] * wn[has], rtol=1e-5, atol=1e-6 * scale)
