"""Served gridding on the device: continuous batching with HBM admission on
the caller's data.

  * every request's cube in a mix (bursts of arrivals, two slots in flight,
    repeats of cached geometries) equals the same request served alone,
    bit for bit;
  * a budget too small for more than a couple of jobs forces small rounds
    and geometry evictions, and changes no result;
  * host (numpy) inputs give the same cube as the same data on the device;
  * an invalid request among valid ones fails alone;
  * one served request against the FP64 oracle on the same coordinates,
    values and target;
  * 60 requests of the configured mix (N up to 5M, maps up to 2048^2, up to
    512 channels) in one burst: a sample equals alone, and a 2048^2 request's
    band matches the oracle;
  * the device content hash (the geometry cache's key) equals its formula.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from hcgrid_amd import capi, serve

from helpers import map_errors

pytestmark = pytest.mark.gpu

SMALL = dict(sides=(16, 24, 32), chans=(1, 16, 48, 80), sigmas=(0.015, 0.02, 0.03), n_range=(800, 6000),
             dec_range=(-50.0, 70.0))


def _specs(n, seed, **dist):
    return [s for _, s in serve.request_stream(n, rate=1e6, seed=seed, repeat_frac=0.3, recent=4, **dist)]


def _serve(reqs, **kw):
    srv = serve.GriddingServer(keep_results=True, **kw)
    for r in reqs:
        srv.submit(r)
    done = srv.serve(len(reqs), timeout_s=300)
    stats = dict(srv.stats)
    out = {r.rid: (r.cube.cpu().numpy().copy(), r.wsum.cpu().numpy().copy()) for r in done}
    srv.close()
    return out, stats


def _alone(spec):
    out, _ = _serve([spec.make()], slots=1)
    return out[spec.rid]


def test_served_mix_equals_alone(gpu):
    specs = _specs(14, seed=11, **SMALL)
    print(f"case: served mix of {len(specs)} requests, {sum(s.repeat for s in specs)} repeats", flush=True)
    mixed, stats = _serve([s.make() for s in specs], slots=2, max_round=5)
    print(f"  server stats {stats}", flush=True)
    assert stats["rounds"] >= 2 and stats["geometry_hits"] >= 1
    for s in specs:
        c0, w0 = _alone(s)
        c1, w1 = mixed[s.rid]
        assert np.array_equal(c0, c1, equal_nan=True), f"request {s.rid} differs alone and in the mix"
        assert np.array_equal(w0, w1)


def test_served_small_budget(gpu):
    specs = _specs(8, seed=12, **SMALL)
    print("case: served with a small budget", flush=True)
    big, _ = _serve([s.make() for s in specs], slots=2)
    probe = serve.GriddingServer()
    sizes = []
    for s in specs:
        r = s.make()
        sizes.append(probe.be.job_bytes(r, probe.waves) + probe.be.geo_bytes(r))
    probe.close()
    # just above the largest job: the big ones run alone, geometries get evicted
    small, stats = _serve([s.make() for s in specs], slots=2, budget_bytes=1.05 * max(sizes))
    print(f"  server stats {stats}", flush=True)
    assert stats["rounds"] >= 3
    for s in specs:
        assert np.array_equal(big[s.rid][0], small[s.rid][0], equal_nan=True)


def test_served_host_inputs_equal_device_inputs(gpu):
    spec = _specs(1, seed=13, **SMALL)[0]
    r_dev = spec.make()
    r_dev.ready.synchronize()
    # the same bytes on the host; values as float32 (exact for BF16/FP16)
    r_host = serve.Request(spec.rid, r_dev.lon.cpu().numpy(), r_dev.lat.cpu().numpy(),
                           r_dev.values.float().cpu().numpy(), r_dev.target, spec.sigma, spec.dtype)
    print(f"case: host inputs n={r_host.n} nchan={r_host.nchan}", flush=True)
    dev, _ = _serve([r_dev], slots=1)
    host, stats = _serve([r_host], slots=1)
    assert stats["geometry_builds"] == 1
    assert np.array_equal(dev[spec.rid][0], host[spec.rid][0], equal_nan=True)
    assert np.array_equal(dev[spec.rid][1], host[spec.rid][1])


def test_served_invalid_request_fails_alone(gpu):
    specs = _specs(5, seed=14, **SMALL)
    reqs = [s.make() for s in specs]
    reqs[2].values = reqs[2].values[:, :-1]  # one sample short
    srv = serve.GriddingServer(keep_results=True, slots=2)
    for r in reqs:
        srv.submit(r)
    srv.serve(len(reqs), timeout_s=300)
    print(f"case: invalid request among {len(reqs)}: {reqs[2].error}", flush=True)
    assert isinstance(reqs[2].error, capi.HcgError) and reqs[2].cube is None
    got = {r.rid: r.cube.cpu().numpy().copy() for r in srv.finished}
    srv.close()
    assert len(got) == 4
    for s in specs:
        if s.rid != specs[2].rid:
            assert np.array_equal(got[s.rid], _alone(s)[0], equal_nan=True)


def _oracle_check(r, inputs, cube, geometry, rows=None, chans=None):
    """r: the served request; inputs: its (lon, lat, values), which the
    server released when the request completed."""
    import torch
    lon_d, lat_d, values = inputs
    side = r.target.side
    m = side * side
    xw = np.empty(m)
    yw = np.empty(m)
    torch.cuda.synchronize()
    capi.call("hcg_memcpy_d2h", C.c_void_p(xw.ctypes.data), C.c_void_p(geometry.view.xwcs), m * 8, None)
    capi.call("hcg_memcpy_d2h", C.c_void_p(yw.ctypes.data), C.c_void_p(geometry.view.ywcs), m * 8, None)
    torch.cuda.synchronize()
    chans = list(range(r.nchan)) if chans is None else chans
    r0, r1 = (0, side) if rows is None else rows
    V = values[chans].float().cpu().numpy()
    lon, lat = lon_d.cpu().numpy(), lat_d.cpu().numpy()
    h, g = oracle.prepare_grid_kernel(oracle.GAUSS1D, [r.sigma], 3 * r.sigma, r.sigma / 2, side)
    keys, perm = oracle.presort(h, lon, lat)
    assert np.array_equal(geometry.perm(), perm)
    fr, used, lut = oracle.ring_lut(h, keys, r.n)
    out, ws = oracle.grid_compat_multi(h, g, lon[perm], lat[perm], V[:, perm], None, keys, fr, used, lut, xw, yw,
                                       r0, r1)
    with np.errstate(invalid="ignore", divide="ignore"):
        o = out / ws[None, :]
    got = cube.reshape(r.nchan, side, side)[chans][:, r0:r1, :].reshape(len(chans), -1)
    return map_errors(got.reshape(-1), o.reshape(-1))


def test_served_request_against_oracle(gpu):
    from hcgrid_amd import synth
    lon, lat, tgt = synth.device_field(5000, 24, 35.0, seed=77)
    vals = synth.device_values(40, "bf16", lon, lat, tgt, seed=78)
    r = serve.Request(0, lon, lat, vals, tgt, 0.02, "bf16")
    srv = serve.GriddingServer(keep_results=True, slots=1)
    srv.submit(r)
    srv.serve(1, timeout_s=120)
    assert r.error is None
    assert r.lon is None and r.values is None  # released by the server on completion
    print(f"case: served request n={r.n} side={r.target.side} nchan={r.nchan}", flush=True)
    err = _oracle_check(r, (lon, lat, vals), r.cube.cpu().numpy(), srv.geoms[r.geo_key].geometry)
    srv.close()
    print(f"  errors {err}", flush=True)
    assert err["nan_match"] and err["max_rel"] <= 1e-4 and err["rms_rel"] <= 1e-4


def test_served_configured_mix_burst(gpu):
    # the served configuration's distribution (batched-job mix, 30% repeats),
    # 60 requests arriving at once
    specs = _specs(60, seed=4)
    reqs = [s.make() for s in specs]
    print(f"case: 60-request burst of the configured mix, {sum(s.repeat for s in specs)} repeats, "
          f"{sum(r.updates for r in reqs):.3e} updates", flush=True)
    side_max = max(r.target.side for r in reqs)
    keep = {r.rid: (r.lon, r.lat, r.values) for r in reqs if r.target.side == side_max}
    srv = serve.GriddingServer(keep_results=True, slots=2, budget_bytes=200e9)
    for r in reqs:
        srv.submit(r)
    srv.serve(len(reqs), timeout_s=600)
    print(f"  server stats {dict(srv.stats)}", flush=True)
    assert all(r.error is None for r in reqs) and len(srv.finished) == 60
    # a 2048^2 request against the oracle (4 middle rows, every 32nd channel)
    big = max((r for r in reqs if r.target.side == side_max), key=lambda r: r.updates)
    side = big.target.side
    print(f"case: burst request {big.rid} n={big.n} nchan={big.nchan} side {side} against the oracle", flush=True)
    err = _oracle_check(big, keep.pop(big.rid), big.cube.cpu().numpy(), srv.geoms[big.geo_key].geometry,
                        rows=(side // 2 - 2, side // 2 + 2), chans=list(range(0, big.nchan, 32)))
    print(f"  errors {err}", flush=True)
    assert err["nan_match"] and err["cells"] > 0
    assert err["max_rel"] <= (1e-6 if big.nchan == 1 else 1e-4) and err["rms_rel"] <= 1e-4
    # a sample of the burst alone: the largest, a repeat and a one-channel job
    mixed = {r.rid: r.cube for r in reqs}
    pick = [max(specs, key=lambda s: s.n * s.nchan)]
    pick += [s for s in specs if s.repeat][:1] + [s for s in specs if s.nchan == 1][:1]
    srv.close()
    del reqs, keep
    for s in pick:
        print(f"case: burst request {s.rid} alone (n={s.n} nchan={s.nchan} side={s.side})", flush=True)
        c0, _ = _alone(s)
        assert np.array_equal(c0, mixed[s.rid].cpu().numpy(), equal_nan=True)


def _hash_formula(b: bytes) -> tuple:
    """The hcg_content_hash formula in numpy (uint64 arithmetic wraps modulo 2^64)."""
    M = np.uint64(0xFFFFFFFFFFFFFFFF)

    def mix(z):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xbf58476d1ce4e5b9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94d049bb133111eb)
        return z ^ (z >> np.uint64(31))

    nb = len(b)
    pad = b + bytes((-nb) % 8)
    w = np.frombuffer(pad, dtype="<u8").astype(np.uint64)
    i = np.arange(w.size, dtype=np.uint64)
    A0, B0 = np.uint64(0x9e3779b97f4a7c15), np.uint64(0x243f6a8885a308d3)
    A1, B1 = np.uint64(0xc2b2ae3d27d4eb4f), np.uint64(0x13198a2e03707344)
    with np.errstate(over="ignore"):
        h0 = mix(np.array([nb], np.uint64) ^ B1)[0] + np.sum(mix(w ^ (i * A0 + B0)), dtype=np.uint64)
        h1 = mix(np.array([nb], np.uint64) ^ B0)[0] + np.sum(mix(w ^ (i * A1 + B1)), dtype=np.uint64)
    return int(h0 & M), int(h1 & M)


@pytest.mark.parametrize("nbytes", [0, 5, 8, 8 * 1000 + 3, 8 * 300001])
def test_content_hash_equals_formula(gpu, nbytes):
    import torch
    rng = np.random.default_rng(nbytes)
    host = rng.integers(0, 256, size=nbytes, dtype=np.uint8)
    d = torch.as_tensor(np.concatenate([host, np.zeros(8, np.uint8)]), device="cuda")
    out = torch.empty(2, dtype=torch.int64, device="cuda")
    print(f"case: content hash of {nbytes} bytes", flush=True)
    capi.call("hcg_content_hash", C.c_void_p(d.data_ptr()), nbytes, C.c_void_p(out.data_ptr()), None)
    torch.cuda.synchronize()
    got = tuple(int(x) & 0xFFFFFFFFFFFFFFFF for x in out.cpu().tolist())
    assert got == _hash_formula(host.tobytes())
    # one changed byte changes both words
    if nbytes:
        d[nbytes // 2] ^= 1
        capi.call("hcg_content_hash", C.c_void_p(d.data_ptr()), nbytes, C.c_void_p(out.data_ptr()), None)
        torch.cuda.synchronize()
        got2 = tuple(int(x) & 0xFFFFFFFFFFFFFFFF for x in out.cpu().tolist())
        assert got2[0] != got[0] and got2[1] != got[1]
