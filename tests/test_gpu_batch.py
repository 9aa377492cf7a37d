"""Batched multi-channel gridding (one persistent launch over uneven jobs),
resident geometries and graph replay.

Checks:
  * every job of a mixed batch (sample counts, map sides, channel counts,
    BF16/FP16, sigma, declination) against the FP64 oracle, to the cube
    bar: peak-normalised max relative error <= 1e-4, RMS <= 1e-4;
  * each job's cube equals the one-job launch's with the same tile shape,
    bit for bit (the batch does not change any job's result);
  * the result does not depend on the batch's waves/chunk;
  * a captured graph (geometry builds + batch) replays to the eager result,
    several times with eager work in between, and picks up new channel
    values written into the same buffers.
"""
import numpy as np
import pytest

import oracle
from hcgrid_amd import batch, gridder, synth

from helpers import map_errors

pytestmark = pytest.mark.gpu

# (n, side, nchan, dtype, sigma, dec, tile_cols)
JOBS = [
    (6000, 24, 70, "bf16", 0.02, 30.0, 16),
    (3000, 16, 1, "f16", 0.03, -40.0, 8),
    (9000, 32, 130, "bf16", 0.015, 62.0, 32),
    (2500, 20, 16, "f16", 0.02, 10.0, 64),
    (4000, 28, 48, "bf16", 0.03, 75.0, 0),
    (12000, 512, 1, "bf16", 0.02, 20.0, 0),  # one channel on a 512^2 map: the FP64 tile path
    (12000, 512, 8, "f16", 0.02, 45.0, 0),  # 8 channels on a sparse 512^2 map: few-channel tiles
    (12000, 512, 16, "bf16", 0.02, 50.0, 0),  # 16 channels there: the contraction
    (20000, 24, 40, "bf16", 0.02, 10.0, 0),  # ~35 samples per cell: the dense jobs' own 8-wave launch
]


def _build(torch, jobs, seed=0):
    out = []
    for i, (n, side, nchan, dtype, sigma, dec, tc) in enumerate(jobs):
        lon, lat, tgt = synth.batch_field(n, side, dec, cdelt=0.01, seed=seed + i)
        spec = synth.BatchSpec(f"t{i}", lon, lat, tgt, nchan, dtype, sigma, seed + i)
        xw, yw = gridder.wcs_sin(side, side, tgt.crpix, tgt.crpix, -tgt.cdelt, tgt.cdelt, tgt.crval1, tgt.crval2)
        kern, hp = gridder.prepare_grid_kernel(1, [sigma], 3 * sigma, sigma / 2)
        g = batch.Geometry(kern, hp, spec.n, side)
        d = lambda a: torch.as_tensor(a, dtype=torch.float64, device="cuda")  # noqa: E731
        lonD, latD, xwD, ywD = d(lon), d(lat), d(xw), d(yw)
        vals = synth.batch_values(spec)
        cube = torch.empty((nchan, side * side), dtype=torch.float32, device="cuda")
        wsum = torch.empty(side * side, dtype=torch.float32, device="cuda")
        out.append(dict(spec=spec, xw=xw, yw=yw, kern=kern, hp=hp, geometry=g, lon=lonD, lat=latD, xwD=xwD,
                        ywD=ywD, values=vals, cube=cube, wsum=wsum, dtype=dtype, tile_cols=tc))
    return out


def _batch_jobs(js):
    return [dict(geometry=j["geometry"], values=j["values"], cube=j["cube"], wsum=j["wsum"], dtype=j["dtype"],
                 tile_cols=j["tile_cols"]) for j in js]


def _single(torch, j):
    """The one-job path on the same geometry (hcg_grid_mfma_shaped)."""
    g = gridder.CubeGridder(j["spec"].n, j["spec"].target.side, j["spec"].nchan, j["kern"], j["hp"],
                            dtype=j["dtype"])
    g.run(j["lon"], j["lat"], j["values"], j["xwD"], j["ywD"], shape=(j["tile_cols"], 0, 0))
    torch.cuda.synchronize()
    return g.cube.cpu().numpy().copy(), g.wsum.cpu().numpy().copy()


def _oracle(j):
    s = j["spec"]
    side = s.target.side
    V = j["values"].float().cpu().numpy()
    ref = oracle.pipeline(s.lon, s.lat, V[0], j["xw"], j["yw"], side, s.sigma, s.radius, s.hpx_res)
    out, ws = oracle.grid_compat_multi(ref["h"], ref["g"], ref["lon_s"], ref["lat_s"], V[:, ref["perm"]], None,
                                       ref["keys_sorted"], ref["firstring"], ref["usedrings"], ref["lut"],
                                       j["xw"], j["yw"], 0, side)
    with np.errstate(invalid="ignore", divide="ignore"):
        return out / ws[None, :], ws, ref["perm"]


def _run_eager(torch, js, waves=0, chunk=0):
    for j in js:
        j["geometry"].build(j["lon"], j["lat"], j["xwD"], j["ywD"])
    b = batch.Batch(_batch_jobs(js), waves=waves, chunk=chunk)
    b.enqueue()
    torch.cuda.synchronize()
    res = [(j["cube"].cpu().numpy().copy(), j["wsum"].cpu().numpy().copy()) for j in js]
    return b, res


def test_batch_against_oracle_and_single(gpu):
    import torch
    js = _build(torch, JOBS)
    b, res = _run_eager(torch, js)
    st = b.stats()
    print(f"batch stats {st}", flush=True)
    assert st["njobs"] == len(JOBS) and st["nitems"] > 0
    assert b.nonfinite() == [0] * len(JOBS)
    for j, (cube, wsum) in zip(js, res):
        s = j["spec"]
        print(f"case: batch job {s.name} n={s.n} side={s.target.side} nchan={s.nchan} {j['dtype']} "
              f"sigma={s.sigma} dec={s.target.crval2}", flush=True)
        o, ws, perm = _oracle(j)
        assert np.array_equal(j["geometry"].perm(), perm)
        e = map_errors(cube.reshape(-1), o.reshape(-1))
        wrel = np.abs(wsum - ws)[ws > 0] / ws[ws > 0]
        print(f"  errors {e} wsum_rel_max {wrel.max() if wrel.size else 0:.2e}", flush=True)
        assert e["nan_match"] and e["max_rel"] <= 1e-4 and e["rms_rel"] <= 1e-4
        assert wrel.size > 0 and wrel.max() <= 2.0 ** -14
        if s.nchan == 1:
            # one channel takes the FP64 tile path (one wave per tile on the
            # 512^2 map, four ring streams per tile on the 16^2 one):
            # unrounded weights, so only the FP32 output rounding separates
            # it from the oracle
            assert e["max_rel"] <= 1e-6 and wrel.max() <= 1e-6
        sc, sw = _single(torch, j)
        assert np.array_equal(cube, sc, equal_nan=True), "batch and one-job cubes differ"
        assert np.array_equal(wsum, sw)


# (8, 128): the LDS of a chunk-128 workgroup holds five one-channel tiles,
# so three of its eight waves idle on those items
@pytest.mark.parametrize("waves,chunk", [(1, 128), (2, 256), (8, 512), (4, 128), (8, 128)])
def test_batch_shape_independent(gpu, waves, chunk):
    import torch
    js = _build(torch, JOBS[:3] + JOBS[5:], seed=40)
    _, ref = _run_eager(torch, js)
    print(f"case: batch waves={waves} chunk={chunk}", flush=True)
    _, res = _run_eager(torch, js, waves=waves, chunk=chunk)
    for (c0, w0), (c1, w1) in zip(ref, res):
        assert np.array_equal(c0, c1, equal_nan=True)
        assert np.array_equal(w0, w1)


def test_batch_graph_replay(gpu):
    import torch
    js = _build(torch, JOBS, seed=80)
    b, eager = _run_eager(torch, js)
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()

    def seq(s):
        for j in js:
            j["geometry"].build(j["lon"], j["lat"], j["xwD"], j["ywD"], stream=s)
        b.enqueue(stream=s)

    graph = batch.Graph(seq, stream)
    for rep in range(3):
        for j in js:
            j["cube"].fill_(-1.0)
            j["wsum"].fill_(-1.0)
        x = torch.randn(1 << 20, device="cuda")  # eager work between replays
        (x * 2).sum().item()
        torch.cuda.synchronize()
        print(f"case: graph replay {rep}", flush=True)
        graph.launch()
        stream.synchronize()
        for j, (c0, w0) in zip(js, eager):
            assert np.array_equal(j["cube"].cpu().numpy(), c0, equal_nan=True)
            assert np.array_equal(j["wsum"].cpu().numpy(), w0)
    # new channel values in the same buffers: the replay follows them
    for j in js:
        j["values"].mul_(2)
    torch.cuda.synchronize()
    graph.launch()
    stream.synchronize()
    for j, (c0, _) in zip(js, eager):
        got = j["cube"].cpu().numpy()
        fin = np.isfinite(c0)
        assert np.array_equal(np.isfinite(got), fin)
        np.testing.assert_allclose(got[fin], 2 * c0[fin], rtol=1e-5, atol=1e-5 * np.abs(c0[fin]).max())
    graph.close()


def test_batch_empty_and_zero_channel_jobs(gpu):
    import torch
    js = _build(torch, JOBS[:1], seed=120)
    j = js[0]
    j["geometry"].build(j["lon"], j["lat"], j["xwD"], j["ywD"])
    # a batch with no jobs and a batch whose only other job has no channels
    batch.Batch([]).enqueue()
    empty = dict(geometry=j["geometry"], values=None, cube=None, wsum=None)
    b = batch.Batch([empty] + _batch_jobs(js))
    b.enqueue()
    torch.cuda.synchronize()
    assert b.stats()["njobs"] == 2
    sc, _ = _single(torch, j)
    assert np.array_equal(j["cube"].cpu().numpy(), sc, equal_nan=True)
