"""Served gridding: continuous batching with HBM admission.

  * every request's cube in a mix (bursts of arrivals, two slots in flight,
    repeats of cached geometries) equals the same request served alone,
    bit for bit;
  * a budget too small for more than a couple of jobs forces small rounds
    and geometry evictions, and changes no result;
  * one served request against the FP64 oracle on the same (device-made)
    coordinates, values and target.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from hcgrid_amd import capi, serve, synth

from helpers import map_errors

pytestmark = pytest.mark.gpu

SMALL = dict(sides=(16, 24, 32), chans=(1, 16, 48, 80), sigmas=(0.015, 0.02, 0.03), n_range=(800, 6000),
             dec_range=(-50.0, 70.0))


def _requests(n, seed):
    return [r for _, r in serve.request_stream(n, rate=1e6, seed=seed, repeat_frac=0.3, recent=4, **SMALL)]


def _serve(reqs, **kw):
    srv = serve.GriddingServer(keep_results=True, **kw)
    for r in reqs:
        srv.submit(r)
    done = srv.serve(len(reqs), timeout_s=300)
    stats = dict(srv.stats)
    out = {r.rid: (r.cube.cpu().numpy().copy(), r.wsum.cpu().numpy().copy()) for r in done}
    srv.close()
    return out, stats


def _clone(r):
    return serve.Request(r.rid, r.n, r.side, r.nchan, r.dtype, r.sigma, r.dec, r.geo_seed, r.val_seed,
                         repeat=r.repeat)


def test_served_mix_equals_alone(gpu):
    reqs = _requests(14, seed=11)
    print(f"case: served mix of {len(reqs)} requests, {sum(r.repeat for r in reqs)} repeats", flush=True)
    mixed, stats = _serve([_clone(r) for r in reqs], slots=2, max_round=5)
    print(f"  server stats {stats}", flush=True)
    assert stats["rounds"] >= 2 and stats["geometry_hits"] >= 1
    for r in reqs:
        alone, _ = _serve([_clone(r)], slots=1)
        c0, w0 = alone[r.rid]
        c1, w1 = mixed[r.rid]
        assert np.array_equal(c0, c1, equal_nan=True), f"request {r.rid} differs alone and in the mix"
        assert np.array_equal(w0, w1)


def test_served_small_budget(gpu):
    reqs = _requests(8, seed=12)
    print("case: served with a small budget", flush=True)
    big, _ = _serve([_clone(r) for r in reqs], slots=2)
    sizes = []
    probe = serve.GriddingServer()
    for r in reqs:
        sizes.append(probe._job_bytes(r) + probe._geo_bytes(r))
    probe.close()
    # just above the largest job: the big ones run alone, geometries get evicted
    small, stats = _serve([_clone(r) for r in reqs], slots=2, budget_bytes=1.05 * max(sizes))
    print(f"  server stats {stats}", flush=True)
    assert stats["rounds"] >= 3
    for r in reqs:
        assert np.array_equal(big[r.rid][0], small[r.rid][0], equal_nan=True)


def test_served_request_against_oracle(gpu):
    import torch
    r = serve.Request(0, 5000, 24, 40, "bf16", 0.02, 35.0, geo_seed=77, val_seed=78)
    srv = serve.GriddingServer(keep_results=True, slots=1)
    srv.submit(r)
    srv.serve(1, timeout_s=120)
    cube = r.cube.cpu().numpy().copy()
    e = srv.geoms[r.geo_key]
    m = r.side * r.side
    xw = np.empty(m)
    yw = np.empty(m)
    torch.cuda.synchronize()
    capi.call("hcg_memcpy_d2h", C.c_void_p(xw.ctypes.data), C.c_void_p(e.geometry.view.xwcs), m * 8, None)
    capi.call("hcg_memcpy_d2h", C.c_void_p(yw.ctypes.data), C.c_void_p(e.geometry.view.ywcs), m * 8, None)
    torch.cuda.synchronize()
    lon, lat, tgt = synth.device_field(r.n, r.side, r.dec, r.geo_seed)
    V = synth.device_values(r.nchan, r.dtype, lon, lat, tgt, r.val_seed).float().cpu().numpy()
    lon, lat = lon.cpu().numpy(), lat.cpu().numpy()
    srv.close()
    print(f"case: served request n={r.n} side={r.side} nchan={r.nchan}", flush=True)
    ref = oracle.pipeline(lon, lat, V[0], xw, yw, r.side, r.sigma, 3 * r.sigma, r.sigma / 2)
    out, ws = oracle.grid_compat_multi(ref["h"], ref["g"], ref["lon_s"], ref["lat_s"], V[:, ref["perm"]], None,
                                       ref["keys_sorted"], ref["firstring"], ref["usedrings"], ref["lut"], xw, yw,
                                       0, r.side)
    with np.errstate(invalid="ignore", divide="ignore"):
        o = out / ws[None, :]
    err = map_errors(cube.reshape(-1), o.reshape(-1))
    print(f"  errors {err}", flush=True)
    assert err["nan_match"] and err["max_rel"] <= 1e-4 and err["rms_rel"] <= 1e-4
