"""The configured workloads at full size on the device, against the oracle.

* 20M-sample raster onto a 2048 x 2048 SIN map (the single-channel drop-in
  at scale): HEALPix keys, sort order, firstring/usedrings and the ring LUT
  bit-exact over all 20M samples; map and weight sums within 1e-6 of the
  oracle on three bands of rows (16 across each edge of the scanned field,
  128 in the middle).
* 64 uneven batched jobs (N log-uniform in [5e4, 5e6], sides 128..2048,
  1..512 BF16/FP16 channels, sigma 0.015/0.02/0.03 deg, Dec -60..80 deg) in
  one captured graph (every geometry build + the persistent launch): every
  job's sort order bit-exact, a 4-row band x every 32nd channel within 1e-4
  (peak-normalised max and RMS) of the FP64 oracle (1e-6 for the one-channel
  jobs, which keep FP64 weights), NaN cells identical; and the 4 largest and
  the 2 most polar jobs bit-identical to their one-job launches.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from hcgrid_amd import batch, capi, gridder, synth

from helpers import map_errors, normalise

pytestmark = pytest.mark.gpu


def test_raster_20m_keys_lut_and_map_bands(gpu):
    import torch
    wl = synth.raster_20m()
    t = wl.target
    side, n = t.side, wl.n
    xw, yw = gridder.wcs_sin(side, side, t.crpix, t.crpix, -t.cdelt, t.cdelt, t.crval1, t.crval2)
    kern, hp = gridder.prepare_grid_kernel(1, [wl.sigma], wl.radius, wl.hpx_res)
    g = gridder.PlainGridder(n, side, kern, hp)
    d = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=g.device)  # noqa: E731
    print(f"case: raster_20m n={n} side={side} on the device", flush=True)
    g.run(d(wl.lon), d(wl.lat), d(wl.values, torch.float32), d(xw), d(yw))
    torch.cuda.synchronize()
    keys_dev = g.keys_sorted().astype(np.uint64)
    perm_dev = g.perm()
    meta = g.meta()
    lut_dev = g.lut()
    m_dev = g.buf.map.cpu().numpy()
    wc_dev = g.buf.weightscube.cpu().numpy()

    print("case: raster_20m oracle pre-sort and LUT", flush=True)
    h, gp = oracle.prepare_grid_kernel(oracle.GAUSS1D, [wl.sigma], wl.radius, wl.hpx_res, side)
    keys, perm = oracle.presort(h, wl.lon, wl.lat)
    # per-sample comparison first: how many samples got another pixel
    by_dev = np.empty(n, np.uint64)
    by_dev[perm_dev] = keys_dev[:n]
    by_ref = np.empty(n, np.uint64)
    by_ref[perm] = keys[:n]
    differ = int(np.count_nonzero(by_dev != by_ref))
    print(f"  samples whose device key differs from the oracle's: {differ} of {n}", flush=True)
    assert differ == 0
    assert np.array_equal(keys_dev, keys), "sorted keys differ"
    assert np.array_equal(perm_dev, perm), "sort order differs"
    fr, used, lut = oracle.ring_lut(h, keys, n)
    assert meta["n_valid"] == n and meta["firstring"] == fr and meta["usedrings"] == used
    assert np.array_equal(lut_dev, lut), "ring LUT differs"

    lon_s, lat_s = wl.lon[perm], wl.lat[perm]
    data_s = wl.values.astype(np.float64)[perm]
    # the map (20.48 deg) is wider than the scanned field (20 deg): rows 16-31
    # and 2016-2031 straddle its lower and upper edges (empty rows, partly
    # covered rows), 960-1087 lie in the middle
    for r0, r1 in ((16, 32), (960, 1088), (side - 32, side - 16)):
        print(f"case: raster_20m rows {r0}-{r1 - 1} against the oracle", flush=True)
        dc, wc = oracle.grid_compat_rows(h, gp, lon_s, lat_s, data_s, None, keys, fr, used, lut, xw, yw, r0, r1)
        sl = slice(r0 * side, r1 * side)
        e = map_errors(m_dev[sl], normalise(dc, wc))
        ew = map_errors(wc_dev[sl], wc)
        print(f"  errors map={e} weights={ew}", flush=True)
        assert e["nan_match"] and e["max_rel"] <= 1e-6 and e["cell_rel"] <= 1e-6
        assert ew["max_rel"] <= 1e-6
        assert e["cells"] > 0


def _oracle_band(sp, xw, yw, vals, chans, r0, r1):
    """Oracle pre-sort, LUT and the multi-channel compat loop over rows [r0, r1)."""
    side = sp.target.side
    V = vals[chans].float().cpu().numpy()
    h, g = oracle.prepare_grid_kernel(oracle.GAUSS1D, [sp.sigma], sp.radius, sp.hpx_res, side)
    keys, perm = oracle.presort(h, sp.lon, sp.lat)
    fr, used, lut = oracle.ring_lut(h, keys, sp.n)
    out, ws = oracle.grid_compat_multi(h, g, sp.lon[perm], sp.lat[perm], V[:, perm], None, keys, fr, used, lut,
                                       xw, yw, r0, r1)
    with np.errstate(invalid="ignore", divide="ignore"):
        return out / ws[None, :], ws, perm


def test_batched_64_job_mix(gpu):
    import torch
    specs = synth.batch_specs(64)
    jobs = []
    for sp in specs:
        t = sp.target
        kern, hp = gridder.prepare_grid_kernel(1, [sp.sigma], sp.radius, sp.hpx_res)
        d = lambda a: torch.as_tensor(a, dtype=torch.float64, device="cuda")  # noqa: E731
        # target coordinates made on the device (the WCS has its own tests);
        # the oracle reads the same values back
        m = t.side * t.side
        xwD = torch.empty(m, dtype=torch.float64, device="cuda")
        ywD = torch.empty(m, dtype=torch.float64, device="cuda")
        capi.call("hcg_wcs_sin_p2s_device", t.side, t.side, t.crpix, t.crpix, -t.cdelt, t.cdelt, t.crval1,
                  t.crval2, 180.0, C.c_void_p(xwD.data_ptr()), C.c_void_p(ywD.data_ptr()), None)
        torch.cuda.synchronize()
        xw, yw = xwD.cpu().numpy(), ywD.cpu().numpy()
        jobs.append(dict(spec=sp, xw=xw, yw=yw, kern=kern, hp=hp, geometry=batch.Geometry(kern, hp, sp.n, t.side),
                         lon=d(sp.lon), lat=d(sp.lat), xwD=xwD, ywD=ywD, values=synth.batch_values(sp),
                         cube=torch.empty((sp.nchan, t.side * t.side), dtype=torch.float32, device="cuda"),
                         wsum=torch.empty(t.side * t.side, dtype=torch.float32, device="cuda")))
    torch.cuda.synchronize()
    print(f"case: 64-job batch, {sum(sp.n for sp in specs)} samples, "
          f"{sum(sp.n * sp.nchan for sp in specs):.3e} updates", flush=True)
    b = batch.Batch([dict(geometry=j["geometry"], values=j["values"], cube=j["cube"], wsum=j["wsum"],
                          dtype=j["spec"].dtype) for j in jobs])
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()

    def seq(s):
        for j in jobs:
            j["geometry"].build(j["lon"], j["lat"], j["xwD"], j["ywD"], stream=s)
        b.enqueue(stream=s)

    graph = batch.Graph(seq, stream)
    graph.launch()
    stream.synchronize()
    assert b.nonfinite(stream) == [0] * len(jobs)
    print(f"  batch stats {b.stats()}", flush=True)

    worst = {"max_rel": 0.0, "rms_rel": 0.0}
    for j in jobs:
        sp = j["spec"]
        side = sp.target.side
        r0 = side // 2 - 2
        r1 = r0 + 4
        chans = list(range(0, sp.nchan, 32))
        print(f"case: batch job {sp.name} n={sp.n} side={side} nchan={sp.nchan} {sp.dtype} sigma={sp.sigma} "
              f"dec={sp.target.crval2}", flush=True)
        o, ws, perm = _oracle_band(sp, j["xw"], j["yw"], j["values"], chans, r0, r1)
        assert np.array_equal(j["geometry"].perm(), perm), "sort order differs"
        cube = j["cube"].view(sp.nchan, side, side)[chans][:, r0:r1, :].cpu().numpy().reshape(len(chans), -1)
        wsum = j["wsum"].view(side, side)[r0:r1].cpu().numpy().reshape(-1)
        e = map_errors(cube.reshape(-1), o.reshape(-1))
        wrel = np.abs(wsum - ws)[ws > 0] / ws[ws > 0]
        print(f"  errors {e} wsum_rel_max {wrel.max() if wrel.size else 0:.2e}", flush=True)
        assert e["nan_match"] and e["cells"] > 0
        tol = 1e-6 if sp.nchan == 1 else 1e-4
        assert e["max_rel"] <= tol and e["rms_rel"] <= tol
        assert wrel.max() <= (1e-6 if sp.nchan == 1 else 2.0 ** -14)
        worst = {k: max(worst[k], e[k]) for k in worst}
    print(f"  worst over the 64 jobs: {worst}", flush=True)

    # the batch changes no job's result: the largest and the most polar jobs alone
    by_size = sorted(range(len(jobs)), key=lambda i: -jobs[i]["spec"].n * jobs[i]["spec"].nchan)[:4]
    by_pole = sorted(range(len(jobs)), key=lambda i: -abs(jobs[i]["spec"].target.crval2))[:2]
    for i in dict.fromkeys(by_size + by_pole):
        j = jobs[i]
        sp = j["spec"]
        print(f"case: one-job launch of {sp.name} (n={sp.n} nchan={sp.nchan} dec={sp.target.crval2})", flush=True)
        cg = gridder.CubeGridder(sp.n, sp.target.side, sp.nchan, j["kern"], j["hp"], dtype=sp.dtype)
        cg.run(j["lon"], j["lat"], j["values"], j["xwD"], j["ywD"])
        torch.cuda.synchronize()
        assert torch.equal(torch.nan_to_num(cg.cube, nan=-7.0), torch.nan_to_num(j["cube"], nan=-7.0)), \
            f"{sp.name}: batch and one-job cubes differ"
        assert torch.equal(cg.wsum, j["wsum"])
        del cg
    graph.close()
    b.close()
