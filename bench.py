"""Benchmark of the gridding path on one MI355X.

One step = one pass of the hot path over a synthetic raster-scan workload
whose inputs are already resident in HBM: HEALPix keys + stable radix sort,
permutation gather, ring LUT, zeroed output cubes, the gridding kernel and
the output normalisation.  Default workload: 20M samples (1000 scan lines x
20000 samples over 20 x 20 deg) onto a 2048 x 2048 SIN map at 0.01 deg,
GAUSS1D sigma 0.02 deg, radius 0.06 deg, nside 8192.

Prints one JSON line (rank 0).  `value` = samples x channels / step time.
The roofline object covers the gridding kernel (the dominant launch), timed
with HIP events on the stream it is launched on.  The CPU baseline is the
oracle (the C/OpenMP restatement) run on the host cores over the same
inputs; the same oracle outputs are the checker behind every `error`
object (max and RMS relative error against FP64, NaN cells, bit-exact keys).

Further legs (each on by default):
  multichannel : 2M samples x 1024 BF16 channels onto 1024 x 1024 (MFMA
                 contraction), against the single-channel path applied
                 channel by channel to the same samples;
  batched      : 64 uneven jobs in one captured graph (geometry builds +
                 one persistent launch), autotuned;
  served       : Poisson arrivals at 20 jobs/s of the batched mix through
                 the continuous-batching server.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sample·channel updates/s and p50/p99 job latency at max rel err ≤1e-3 vs FP64"
UNIT = "sample·channel updates/s"
HBM_PEAK_GBS = 8000.0
# VALU issue capacity: 256 CUs x 4 SIMDs, one wave64 VALU instruction per
# SIMD every 4 cycles, at the 2.4 GHz peak engine clock
SIMDS, VALU_CYCLES, CLOCK_HZ = 1024, 4, 2.4e9


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="raster_20m", choices=["raster_20m", "raster_250k"])
    ap.add_argument("--cpu-baseline", dest="cpu_baseline", action="store_true", default=True,
                    help="time the oracle on the host cores and check every leg against it (default)")
    ap.add_argument("--no-cpu-baseline", dest="cpu_baseline", action="store_false")
    ap.add_argument("--cpu-reps", type=int, default=3, help="oracle runs per single-channel workload (median)")
    ap.add_argument("--cube", dest="cube", action="store_true", default=True,
                    help="also measure the multi-channel path (2M samples x 1024 BF16 channels)")
    ap.add_argument("--no-cube", dest="cube", action="store_false")
    ap.add_argument("--batch", dest="batch", action="store_true", default=True,
                    help="also measure 64 uneven multi-channel jobs in one persistent launch (graph replay)")
    ap.add_argument("--no-batch", dest="batch", action="store_false")
    ap.add_argument("--batch-jobs", type=int, default=64)
    ap.add_argument("--served", dest="served", action="store_true", default=True,
                    help="also run the served loop: Poisson arrivals at 20 jobs/s of the batched-job mix")
    ap.add_argument("--no-served", dest="served", action="store_false")
    ap.add_argument("--served-jobs", type=int, default=400,
                    help="requests in the served run (the full configuration is 10000, ~500 s of arrivals)")
    ap.add_argument("--calibrate-copy", dest="calibrate_copy", action="store_true",
                    help="first run one 1 GiB device-to-device copy (a reference for HBM byte counters)")
    return ap.parse_args()


# -- provenance of the counter-based figures ----------------------------------

def csrc_digest() -> str:
    """sha256 (12 hex) of the device sources: the kernels a counter reading
    belongs to."""
    d = os.path.join(ROOT, "hcgrid_amd", "csrc")
    h = hashlib.sha256()
    for name in sorted(os.listdir(d)):
        if name.endswith((".hip", ".hpp")):
            h.update(name.encode())
            with open(os.path.join(d, name), "rb") as f:
                h.update(f.read())
    return h.hexdigest()[:12]


def counter_fields(key: str, avg_ms: float) -> dict:
    """HBM traffic and VALU instructions per launch from the rocprofv3 --pmc
    runs recorded in profiles/traffic.json (FETCH_SIZE and WRITE_SIZE,
    calibrated on a device copy of known size; SQ_INSTS_VALU), with the
    commit and device-source digest they were taken at; `current` says
    whether the kernels are still those."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            rec = json.load(f).get("entries", {}).get(key)
    except (OSError, ValueError):
        rec = None
    if not rec:
        return {"traffic": None, "valu": None, "counters_from": None}
    src = {"commit": rec.get("commit"), "csrc_sha": rec.get("csrc_sha"),
           "current": rec.get("csrc_sha") == csrc_digest(), "profile": rec.get("profile")}
    valu = None
    insts = rec.get("valu_insts")
    if insts and avg_ms:
        issue_ms = insts * VALU_CYCLES / (SIMDS * CLOCK_HZ) * 1e3
        valu = {"insts": insts, "issue_ms": issue_ms, "frac": issue_ms / avg_ms}
    return {"traffic": rec.get("traffic"), "valu": valu, "counters_from": src}


def cpu_info() -> dict:
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"nproc": os.cpu_count(), "threads": threads, "cpu_model": model}


# -- error against the oracle ---------------------------------------------------

def map_error(m, o) -> dict:
    """Peak-normalised max and RMS relative error of m against the FP64
    oracle o, NaN agreement, and the per-cell relative error over cells with
    |o| >= 1% of the peak."""
    import numpy as np
    nan_m, nan_o = np.isnan(m), np.isnan(o)
    good = ~nan_m & ~nan_o
    out = {"nan_match": bool(np.array_equal(nan_m, nan_o)), "cells": int(good.sum())}
    if not good.any():
        out.update(max_rel=0.0, rms_rel=0.0, cell_rel=0.0)
        return out
    mg, og = m[good].astype(np.float64), o[good]
    peak = float(np.max(np.abs(og)))
    diff = np.abs(mg - og)
    out["max_rel"] = float(diff.max() / peak) if peak > 0 else float(diff.max())
    den = float(np.sqrt(np.sum(og * og)))
    out["rms_rel"] = float(np.sqrt(np.sum(diff * diff)) / den) if den > 0 else 0.0
    big = np.abs(og) >= 1e-2 * peak
    out["cell_rel"] = float(np.max(diff[big] / np.abs(og[big]))) if big.any() else 0.0
    return out


def _normalise(dc, wc):
    import numpy as np
    with np.errstate(invalid="ignore", divide="ignore"):
        return dc / wc


# -- the single-channel path ----------------------------------------------------

def grid_algorithmic_bytes(n: int, m: int) -> int:
    # per sample: sorted lon, lat (16 B), value (4 B) and key (4 B) read once;
    # per cell: target coordinates (16 B) read, FP32 data and weight written
    return n * 24 + m * 24


def run_plain_device(wl, xw, yw, steps, warmup, log):
    """Times `steps` passes of the single-channel path; returns the timings
    and the device results of the last pass."""
    import torch

    from hcgrid_amd import gridder

    dev = torch.device("cuda", 0)
    t = wl.target
    kern, hp = gridder.prepare_grid_kernel(1, [wl.sigma], wl.radius, wl.hpx_res)
    g = gridder.PlainGridder(wl.n, t.side, kern, hp)
    lon = torch.as_tensor(wl.lon, device=dev)
    lat = torch.as_tensor(wl.lat, device=dev)
    val = torch.as_tensor(wl.values, device=dev)
    xwd = torch.as_tensor(xw, device=dev)
    ywd = torch.as_tensor(yw, device=dev)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()

    def step(ev_pair=None):
        g.presort(lon, lat)
        g.gather(lon, lat, val)
        g.ring_lut()
        g.init_output()
        if ev_pair is not None:
            ev_pair[0].record(stream)
        g.grid(xwd, ywd)
        if ev_pair is not None:
            ev_pair[1].record(stream)
        g.normalize()

    for i in range(warmup):
        step()
        log(f"{wl.name}: warmup {i + 1}/{warmup} done")
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    start.record(stream)
    for i in range(steps):
        step(evs[i])
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms_step = start.elapsed_time(stop) / steps
    grid_ms = [a.elapsed_time(b) for a, b in evs]
    grid_avg = sum(grid_ms) / len(grid_ms)
    log(f"{wl.name}: timed {steps} steps: {ms_step:.3f} ms/step (events), wall {wall * 1e3 / steps:.3f} ms/step; "
        f"grid kernel avg {grid_avg:.3f} ms (min {min(grid_ms):.3f}, max {max(grid_ms):.3f})")
    res = dict(keys=g.keys_sorted(), perm=g.perm(), lut=g.lut(), meta=g.meta(),
               map=g.buf.map.cpu().numpy().copy(), hp=hp)
    del g, lon, lat, val, xwd, ywd
    torch.cuda.empty_cache()
    return ms_step, grid_avg, res


def cpu_plain(wl, xw, yw, reps, log):
    """The oracle over the whole workload, `reps` times: keys + stable sort,
    gathers, ring LUT, the compat grid of every cell (OpenMP) and the
    normalisation.  Returns (per-stage median seconds, oracle results of
    the last run)."""
    import numpy as np

    import oracle

    info = cpu_info()
    side = wl.target.side
    h, g = oracle.prepare_grid_kernel(oracle.GAUSS1D, [wl.sigma], wl.radius, wl.hpx_res, side)
    stages = {"keys_sort": [], "gather": [], "lut": [], "grid": [], "normalise": []}
    for rep in range(reps):
        t0 = time.perf_counter()
        keys, perm = oracle.presort(h, wl.lon, wl.lat)
        t1 = time.perf_counter()
        lon_s, lat_s = wl.lon[perm], wl.lat[perm]
        data_s = wl.values.astype(np.float64)[perm]
        t2 = time.perf_counter()
        fr, used, lut = oracle.ring_lut(h, keys, wl.n)
        t3 = time.perf_counter()
        dc, wc, _ = oracle.grid_compat(h, g, lon_s, lat_s, data_s, None, keys, fr, used, lut, xw, yw,
                                       nthreads=info["threads"])
        t4 = time.perf_counter()
        omap = _normalise(dc, wc)
        t5 = time.perf_counter()
        for k, v in zip(stages, (t1 - t0, t2 - t1, t3 - t2, t4 - t3, t5 - t4)):
            stages[k].append(v)
        log(f"cpu {wl.name} run {rep + 1}/{reps}: {t5 - t0:.2f} s (grid {t4 - t3:.2f} s, {info['threads']} threads)")
    med = {k: statistics.median(v) for k, v in stages.items()}
    totals = [sum(stages[k][i] for k in stages) for i in range(reps)]
    return med, statistics.median(totals), dict(keys=keys, perm=perm, lut=lut, fr=fr, used=used, map=omap)


def plain_error(dev, ref) -> dict:
    import numpy as np
    e = map_error(dev["map"], ref["map"])
    e["keys_exact"] = bool(np.array_equal(dev["keys"].astype(np.uint64), ref["keys"]))
    e["perm_exact"] = bool(np.array_equal(dev["perm"], ref["perm"]))
    e["lut_exact"] = bool(dev["meta"]["firstring"] == ref["fr"] and dev["meta"]["usedrings"] == ref["used"]
                          and np.array_equal(dev["lut"], ref["lut"]))
    e["against"] = "the oracle (FP64, the reference's compat selection) over the whole map"
    return e


def cpu_baseline_entry(wl, med, total, reps) -> dict:
    info = cpu_info()
    return {
        "value": wl.n / total,
        "unit": UNIT,
        "cores": info["threads"],
        "kind": "port",
        "sample": (f"{wl.name}: the whole workload ({wl.n} samples, {wl.target.side}^2 cells): keys + stable "
                   f"sort, gathers, ring LUT, compat grid of every cell, normalisation; median of {reps} runs"),
        "seconds": total,
        "stage_s": med,
        "reps": reps,
        "nproc": info["nproc"],
        "cpu_model": info["cpu_model"],
    }


# -- the multi-channel path -------------------------------------------------------

def cube_algorithmic_bytes(n: int, m: int, nchan: int) -> int:
    # channel values (16-bit) read once; per sample the sorted lon, lat (16 B)
    # and key (4 B); per cell the target coordinates (16 B) and the FP32
    # weight sum (4 B); the FP32 cube written once
    return n * nchan * 2 + n * 20 + m * 20 + m * nchan * 4


def bench_cube(args, log):
    """Multi-channel path on a 2M-sample x 1024-channel BF16 raster onto a
    1024 x 1024 map, against the single-channel path applied channel by
    channel to the same samples (what the reference's channel loop does: one
    pre-sort, then per channel a value gather, zeroed cubes, the gridding
    kernel and the normalisation)."""
    import torch

    from hcgrid_amd import gridder, synth

    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream()
    t0 = time.perf_counter()
    wl = synth.raster_cube(device="cuda")
    t = wl.target
    xw, yw = gridder.wcs_sin(t.side, t.side, t.crpix, t.crpix, -t.cdelt, t.cdelt, t.crval1, t.crval2)
    torch.cuda.synchronize()
    log(f"cube workload {wl.name}: n={wl.n} nchan={wl.nchan} side={t.side} generated in "
        f"{time.perf_counter() - t0:.1f}s")
    kern, hp = gridder.prepare_grid_kernel(1, [wl.sigma], wl.radius, wl.hpx_res)
    cg = gridder.CubeGridder(wl.n, t.side, wl.nchan, kern, hp, dtype="bf16")
    lon = torch.as_tensor(wl.lon, device=dev)
    lat = torch.as_tensor(wl.lat, device=dev)
    xwd = torch.as_tensor(xw, device=dev)
    ywd = torch.as_tensor(yw, device=dev)
    vals = wl.values

    def cube_step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        cg.plain.presort(lon, lat)
        cg.gather_records(lon, lat)
        cg.plain.ring_lut()
        cg.gather_channels(vals)
        if ev is not None:
            ev[1].record(stream)
        cg.grid(xwd, ywd)
        if ev is not None:
            ev[2].record(stream)

    for _ in range(args.warmup):
        cube_step()
    torch.cuda.synchronize()
    evs = [tuple(torch.cuda.Event(enable_timing=True) for _ in range(3)) for _ in range(args.steps)]
    for i in range(args.steps):
        cube_step(evs[i])
    torch.cuda.synchronize()
    ms_step = sum(a.elapsed_time(c) for a, b, c in evs) / args.steps
    ms_grid = sum(b.elapsed_time(c) for a, b, c in evs) / args.steps
    updates = wl.n * wl.nchan
    log(f"cube: {ms_step:.3f} ms/step, k_grid_mfma avg {ms_grid:.3f} ms")

    # device results of the checked band: rows 504-519, every 32nd channel
    r0, r1 = t.side // 2 - 8, t.side // 2 + 8
    chans = list(range(0, wl.nchan, 32))
    band = cg.cube.view(wl.nchan, t.side, t.side)[chans][:, r0:r1, :].cpu().numpy().reshape(-1)
    dev_perm = cg.plain.perm()
    vsub = wl.values[chans].float().cpu().numpy()
    nonfinite = int(cg.nonfinite.item())

    # the single-channel path, channel by channel, on the same samples
    pg = gridder.PlainGridder(wl.n, t.side, kern, hp)
    ch0 = vals[0].float().contiguous()
    pg.presort(lon, lat)
    pg.ring_lut()

    def channel_step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        pg.gather(lon, lat, ch0)
        pg.init_output()
        if ev is not None:
            ev[1].record(stream)
        pg.grid(xwd, ywd)
        if ev is not None:
            ev[2].record(stream)
        pg.normalize()
        if ev is not None:
            ev[3].record(stream)

    for _ in range(args.warmup):
        channel_step()
    torch.cuda.synchronize()
    evp = [tuple(torch.cuda.Event(enable_timing=True) for _ in range(4)) for _ in range(args.steps)]
    for i in range(args.steps):
        channel_step(evp[i])
    torch.cuda.synchronize()
    ch_ms = sum(a.elapsed_time(d) for a, b, c, d in evp) / args.steps
    ch_grid = sum(b.elapsed_time(c) for a, b, c, d in evp) / args.steps
    # one pre-sort + LUT, then the per-channel work for every channel
    ps = [tuple(torch.cuda.Event(enable_timing=True) for _ in range(2)) for _ in range(args.steps)]
    for i in range(args.steps):
        ps[i][0].record(stream)
        pg.presort(lon, lat)
        pg.ring_lut()
        ps[i][1].record(stream)
    torch.cuda.synchronize()
    presort_ms = sum(a.elapsed_time(b) for a, b in ps) / args.steps
    chan_total_ms = presort_ms + wl.nchan * ch_ms
    chan_rate = updates / (chan_total_ms * 1e-3)
    cube_rate = updates / (ms_step * 1e-3)
    log(f"channel by channel: pre-sort {presort_ms:.3f} ms + {wl.nchan} x {ch_ms:.3f} ms "
        f"(k_grid_tiles {ch_grid:.3f} ms) = {chan_total_ms:.1f} ms; ratio {cube_rate / chan_rate:.2f}x")
    del cg, pg
    torch.cuda.empty_cache()

    m = t.side * t.side
    abytes = cube_algorithmic_bytes(wl.n, m, wl.nchan)
    achieved = abytes / (ms_grid * 1e-3) / 1e9
    key = f"k_grid_mfma/{wl.name}"
    out = {
        "workload": (f"{wl.name}: {wl.n} samples (raster scan, 500 lines x 4000) x {wl.nchan} BF16 channels, "
                     f"{t.side}x{t.side} SIN map at 0.01 deg, GAUSS1D sigma {wl.sigma} deg, radius {wl.radius} deg, "
                     f"nside {hp.nside}; FP64 weights carried as BF16 hi + lo, FP32 accumulation"),
        "value": cube_rate,
        "unit": UNIT,
        "ms_per_step": ms_step,
        "dtype": "bf16",
        "nonfinite_values": nonfinite,
        "channel_by_channel": {
            "value": chan_rate,
            "ms_total": chan_total_ms,
            "presort_ms": presort_ms,
            "per_channel_ms": ch_ms,
            "per_channel_grid_ms": ch_grid,
        },
        "speedup_vs_channel_by_channel": cube_rate / chan_rate,
        "roofline": {
            "kernel": "k_grid_mfma",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "algorithmic_bytes": abytes,
            "avg_ms": ms_grid,
            **counter_fields(key, ms_grid),
        },
    }
    if args.cpu_baseline:
        cb, err = cpu_cube(wl, xw, yw, r0, r1, chans, vsub, band, dev_perm, log)
        out["cpu_baseline"] = cb
        out["error"] = err
    return out


def cpu_cube(wl, xw, yw, r0, r1, chans, vsub, band, dev_perm, log):
    """Oracle on the cube workload: the full pre-sort, gathers and LUT, then
    the multi-channel compat loop over a band of map rows twice, with no
    channels (selection and FP64 weights alone) and with the checked
    channel subset.  The weight cost per row and the cost per channel and
    row give the whole cube's time, scaled linearly in rows and channels.
    The subset's output is the checker of the device band."""
    import numpy as np

    import oracle

    info = cpu_info()
    side = wl.target.side
    rows = r1 - r0
    h, g = oracle.prepare_grid_kernel(oracle.GAUSS1D, [wl.sigma], wl.radius, wl.hpx_res, side)
    t0 = time.perf_counter()
    keys, perm = oracle.presort(h, wl.lon, wl.lat)
    t1 = time.perf_counter()
    lon_s, lat_s = wl.lon[perm], wl.lat[perm]
    vs = np.ascontiguousarray(vsub[:, perm])
    t2 = time.perf_counter()
    fr, used, lut = oracle.ring_lut(h, keys, wl.n)
    t3 = time.perf_counter()
    oracle.grid_compat_multi(h, g, lon_s, lat_s, vs[:0], None, keys, fr, used, lut, xw, yw, r0, r1,
                             nthreads=info["threads"])
    t4 = time.perf_counter()
    out, ws = oracle.grid_compat_multi(h, g, lon_s, lat_s, vs, None, keys, fr, used, lut, xw, yw, r0, r1,
                                       nthreads=info["threads"])
    t5 = time.perf_counter()
    per_row_w = (t4 - t3) / rows
    per_row_ch = max(0.0, (t5 - t4) - (t4 - t3)) / (rows * len(chans))
    gather_full = (t2 - t1) * wl.nchan / len(chans)
    total = (t1 - t0) + gather_full + (t3 - t2) + side * (per_row_w + per_row_ch * wl.nchan)
    log(f"cpu cube: sort {t1 - t0:.2f}s, band of {rows} rows: weights {t4 - t3:.2f}s, +{len(chans)} channels "
        f"{(t5 - t4) - (t4 - t3):.2f}s -> whole cube {total:.1f}s scaled")
    with np.errstate(invalid="ignore", divide="ignore"):
        o = out / ws[None, :]
    err = map_error(band, o.reshape(-1))
    err["perm_exact"] = bool(np.array_equal(dev_perm, perm))
    err["against"] = (f"the oracle's FP64 weights applied to the same BF16 values: rows {r0}-{r1 - 1} x every "
                      f"32nd channel ({len(chans)} channels)")
    cb = {
        "value": wl.n * wl.nchan / total,
        "unit": UNIT,
        "cores": info["threads"],
        "kind": "port",
        "scaled": True,
        "sample": (f"{wl.name}: full pre-sort and LUT; the compat loop over rows {r0}-{r1 - 1} with 0 and "
                   f"{len(chans)} channels, scaled linearly to {side} rows x {wl.nchan} channels"),
        "seconds": total,
        "nproc": info["nproc"],
        "cpu_model": info["cpu_model"],
    }
    return cb, err


# -- batched jobs -------------------------------------------------------------------

def bench_batch(args, log):
    """64 independent jobs (N log-uniform in [50k, 5M], map side in
    {128..2048}, channels in {1, 16, 64, 256, 512}, BF16/FP16, sigma in
    {0.015, 0.02, 0.03} deg, Dec in [-60, 80]) gridded by one graph: every
    job's geometry build (keys, sort, records, LUT) and one persistent
    launch over all jobs' tiles.  Block shapes per job class and the launch
    shape are picked by the autotuner first (untimed).  Compared with the
    single-channel path applied channel by channel to the same jobs."""
    import torch

    from hcgrid_amd import autotune, batch, gridder, synth

    t0 = time.perf_counter()
    specs = synth.batch_specs(njobs=args.batch_jobs)
    jobs, keys = [], []
    for sp in specs:
        t = sp.target
        xw, yw = gridder.wcs_sin(t.side, t.side, t.crpix, t.crpix, -t.cdelt, t.cdelt, t.crval1, t.crval2)
        kern, hp = gridder.prepare_grid_kernel(1, [sp.sigma], sp.radius, sp.hpx_res)
        d = lambda a: torch.as_tensor(a, dtype=torch.float64, device="cuda")  # noqa: E731
        g = batch.Geometry(kern, hp, sp.n, t.side)
        jobs.append(dict(geometry=g, values=synth.batch_values(sp), dtype=sp.dtype,
                         cube=torch.empty((sp.nchan, t.side * t.side), dtype=torch.float32, device="cuda"),
                         wsum=torch.empty(t.side * t.side, dtype=torch.float32, device="cuda"),
                         lon=d(sp.lon), lat=d(sp.lat), xw=d(xw), yw=d(yw), xw_host=xw, yw_host=yw, kern=kern,
                         hp=hp, spec=sp))
        keys.append(autotune.job_class(sp.n, t.side, sp.nchan, sp.sigma, sp.dtype))
    for j in jobs:
        j["geometry"].build(j["lon"], j["lat"], j["xw"], j["yw"])
    torch.cuda.synchronize()
    updates = sum(sp.n * sp.nchan for sp in specs)
    log(f"batch: {len(specs)} jobs, {sum(sp.n for sp in specs)} samples, {updates:.3e} sample-channel updates, "
        f"set up in {time.perf_counter() - t0:.1f}s")

    t1 = time.perf_counter()
    tuner = autotune.Autotuner()
    bjobs = [dict(geometry=j["geometry"], values=j["values"], cube=j["cube"], wsum=j["wsum"], dtype=j["dtype"])
             for j in jobs]
    bjobs = tuner.tune_jobs(bjobs, keys)
    waves, chunk = tuner.launch_for(bjobs, "batch")
    log(f"batch: autotuned {len(tuner.tiles)} job classes and the launch shape in {time.perf_counter() - t1:.1f}s "
        f"(waves {waves}, chunk {chunk})")
    b = batch.Batch(bjobs, waves=waves, chunk=chunk)
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()

    def seq(s):
        for j in jobs:
            j["geometry"].build(j["lon"], j["lat"], j["xw"], j["yw"], stream=s)
        b.enqueue(stream=s)

    graph = batch.Graph(seq, stream)
    for _ in range(max(1, args.warmup)):
        graph.launch()
    stream.synchronize()
    steps = max(1, min(args.steps, 5))
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for e0, e1 in evs:
        e0.record(stream)
        graph.launch()
        e1.record(stream)
    stream.synchronize()
    ms = sum(e0.elapsed_time(e1) for e0, e1 in evs) / steps
    nonfinite = sum(b.nonfinite(stream))
    st = b.stats()
    log(f"batch: {ms:.2f} ms per graph replay ({st['nitems']} work items on {st['blocks']} workgroups); "
        f"non-finite values {nonfinite}")

    # the persistent gridding launches alone: HIP events recorded by the
    # library around them in eager enqueues of the same batch
    b.timing(True)
    prep, grid = [], []
    for _ in range(3):
        b.enqueue(stream=stream)
        p_ms, g_ms = b.last_times()
        prep.append(p_ms)
        grid.append(g_ms)
    b.timing(False)
    grid_ms = sum(grid) / len(grid)
    prep_ms = sum(prep) / len(prep)
    abytes = sum(cube_algorithmic_bytes(j["spec"].n, j["spec"].target.side ** 2, j["spec"].nchan) for j in jobs)
    log(f"batch: eager enqueue: prep {prep_ms:.2f} ms, persistent gridding launches {grid_ms:.2f} ms")

    # device results of the checked bands (4 rows x every 32nd channel per job)
    checked = []
    for j in jobs:
        sp = j["spec"]
        side = sp.target.side
        r0 = side // 2 - 2
        chans = list(range(0, sp.nchan, 32))
        band = j["cube"].view(sp.nchan, side, side)[chans][:, r0:r0 + 4, :].cpu().numpy().reshape(-1)
        checked.append(dict(r0=r0, r1=r0 + 4, chans=chans, band=band, perm=j["geometry"].perm(),
                            vsub=j["values"][chans].float().cpu().numpy()))

    # the single-channel path on the same jobs, channel by channel
    tot_ms = 0.0
    for j in jobs:
        sp = j["spec"]
        pg = gridder.PlainGridder(sp.n, sp.target.side, j["kern"], j["hp"])
        ch0 = j["values"][0].float().contiguous()
        e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        pg.presort(j["lon"], j["lat"])
        pg.ring_lut()
        pg.gather(j["lon"], j["lat"], ch0)
        pg.init_output()
        pg.grid(j["xw"], j["yw"])
        pg.normalize()
        e[0].record()
        pg.presort(j["lon"], j["lat"])
        pg.ring_lut()
        e[1].record()
        for _ in range(2):
            pg.gather(j["lon"], j["lat"], ch0)
            pg.init_output()
            pg.grid(j["xw"], j["yw"])
            pg.normalize()
        e[2].record()
        torch.cuda.synchronize()
        tot_ms += e[0].elapsed_time(e[1]) + sp.nchan * e[1].elapsed_time(e[2]) / 2
        del pg
    chan_rate = updates / (tot_ms * 1e-3)
    rate = updates / (ms * 1e-3)
    log(f"batch: channel-by-channel single-channel path on the same jobs: {tot_ms:.0f} ms; "
        f"ratio {rate / chan_rate:.2f}x")
    graph.close()
    b.close()
    achieved = abytes / (grid_ms * 1e-3) / 1e9
    out = {
        "workload": (f"{len(specs)} jobs: N log-uniform in [5e4, 5e6] (sum {sum(sp.n for sp in specs)}), side in "
                     "{128..2048} at 0.01 deg, channels in {1,16,64,256,512} BF16/FP16, sigma in {0.015,0.02,0.03} "
                     "deg, Dec in [-60, 80]; one graph = every job's geometry build + one persistent launch"),
        "value": rate,
        "unit": UNIT,
        "ms_per_batch": ms,
        "updates": updates,
        "work_items": st["nitems"],
        "nonfinite_values": nonfinite,
        "launch": {"waves": waves, "chunk": chunk, "blocks": st["blocks"]},
        "autotuned_classes": len(tuner.tiles),
        "channel_by_channel": {"value": chan_rate, "ms_total": tot_ms},
        "speedup_vs_channel_by_channel": rate / chan_rate,
        "roofline": {
            "kernel": "k_grid_mfma_batch (+ the dense jobs' launch)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "algorithmic_bytes": abytes,
            "avg_ms": grid_ms,
            "prep_ms": prep_ms,
            "traffic": None,
            "note": "persistent gridding launches of eager enqueues, timed by the library's HIP events on the "
                    "enqueue stream; the channel gathers and tile plans run before, in the prep launch",
        },
    }
    if args.cpu_baseline:
        out["cpu_baseline"], out["error"] = cpu_batch(jobs, checked, log)
    return out, tuner


def cpu_batch(jobs, checked, log):
    """Oracle on every job: the full pre-sort and LUT, then the multi-channel
    compat loop over 4 rows with no channels and with every 32nd channel;
    each job's time scaled linearly to all rows and channels, as for the
    cube.  The subsets' outputs check the device bands."""
    import numpy as np

    import oracle

    info = cpu_info()
    total = 0.0
    worst = {"max_rel": 0.0, "rms_rel": 0.0, "cell_rel": 0.0}
    nan_ok = perm_ok = True
    per_job = []
    for j, c in zip(jobs, checked):
        sp = j["spec"]
        side = sp.target.side
        h, g = oracle.prepare_grid_kernel(oracle.GAUSS1D, [sp.sigma], sp.radius, sp.hpx_res, side)
        t0 = time.perf_counter()
        keys, perm = oracle.presort(h, sp.lon, sp.lat)
        t1 = time.perf_counter()
        lon_s, lat_s = sp.lon[perm], sp.lat[perm]
        vs = np.ascontiguousarray(c["vsub"][:, perm])
        t2 = time.perf_counter()
        fr, used, lut = oracle.ring_lut(h, keys, sp.n)
        t3 = time.perf_counter()
        oracle.grid_compat_multi(h, g, lon_s, lat_s, vs[:0], None, keys, fr, used, lut, j["xw_host"], j["yw_host"],
                                 c["r0"], c["r1"], nthreads=info["threads"])
        t4 = time.perf_counter()
        out, ws = oracle.grid_compat_multi(h, g, lon_s, lat_s, vs, None, keys, fr, used, lut, j["xw_host"],
                                           j["yw_host"], c["r0"], c["r1"], nthreads=info["threads"])
        t5 = time.perf_counter()
        rows, nsub = c["r1"] - c["r0"], len(c["chans"])
        per_row_w = (t4 - t3) / rows
        per_row_ch = max(0.0, (t5 - t4) - (t4 - t3)) / (rows * nsub)
        total += ((t1 - t0) + (t2 - t1) * sp.nchan / nsub + (t3 - t2) + side * (per_row_w + per_row_ch * sp.nchan))
        with np.errstate(invalid="ignore", divide="ignore"):
            o = out / ws[None, :]
        e = map_error(c["band"], o.reshape(-1))
        nan_ok &= e["nan_match"]
        perm_ok &= bool(np.array_equal(c["perm"], perm))
        for k in worst:
            worst[k] = max(worst[k], e[k])
        per_job.append(round(e["max_rel"], 9))
    updates = sum(j["spec"].n * j["spec"].nchan for j in jobs)
    log(f"cpu batch: {len(jobs)} jobs, whole batch {total:.1f}s scaled; worst max_rel {worst['max_rel']:.2e}")
    cb = {
        "value": updates / total,
        "unit": UNIT,
        "cores": info["threads"],
        "kind": "port",
        "scaled": True,
        "sample": ("every job: full pre-sort and LUT; the compat loop over its 4 middle rows with 0 and every 32nd "
                   "channel, scaled linearly to all rows and channels; summed over the jobs"),
        "seconds": total,
        "nproc": info["nproc"],
        "cpu_model": info["cpu_model"],
    }
    err = dict(worst, nan_match=nan_ok, perm_exact=perm_ok, jobs=len(jobs), max_rel_per_job=per_job,
               against="the oracle's FP64 weights on the same 16-bit values, 4 middle rows x every 32nd channel "
                       "of every job; worst job shown")
    return cb, err


# -- served ---------------------------------------------------------------------------

def bench_served(args, log, tuner=None):
    """Served gridding: Poisson arrivals at 20 jobs/s drawn from the batched
    mix (30% repeats of one of the last 32 distinct maps with new channel
    data), HBM admission budget 200 GB, two slots; p50/p99 of completion -
    arrival and throughput over the run.  The load generator makes each
    request's coordinates and values on the device before submitting it
    (no host-to-device transfer in the latency); results stay in HBM.
    tuner: the batched leg's autotuner (block shape per job class), when
    that leg ran; the persistent launch runs at its tuned waves/chunk."""
    from hcgrid_amd import serve

    waves, chunk = 4, 256
    if tuner is not None and "batch" in tuner.launch:
        waves, chunk = (int(x) for x in tuner.launch["batch"])
    warm = serve.request_stream(12, rate=20.0, seed=99)
    srv = serve.GriddingServer(budget_bytes=200e9, slots=2, waves=waves, chunk=chunk, tuner=tuner)
    serve.run_load(srv, warm)
    srv.close()
    stream = serve.request_stream(args.served_jobs, rate=20.0, seed=4, repeat_frac=0.3, recent=32)
    srv = serve.GriddingServer(budget_bytes=200e9, slots=2, waves=waves, chunk=chunk, tuner=tuner)
    t0 = time.perf_counter()
    done = serve.run_load(srv, stream,
                          progress=lambda k, n: log(f"served: {k}/{n} requests done, "
                                                    f"{time.perf_counter() - t0:.0f}s"))
    wall = time.perf_counter() - t0
    summ = serve.latency_summary(done)
    stats = dict(srv.stats)
    srv.close()
    log(f"served: {summ['jobs']} jobs in {wall:.1f}s, p50 {summ['p50_ms']:.1f} ms, p99 {summ['p99_ms']:.1f} ms, "
        f"{summ['value']:.3e} updates/s, rounds {stats.get('rounds')}, geometry builds "
        f"{stats.get('geometry_builds')}, hits {stats.get('geometry_hits')}, evictions {stats.get('evictions', 0)}")
    summ.update({
        "workload": (f"{args.served_jobs} requests, Poisson 20 jobs/s from the batched mix, 30% repeats of recent "
                     "maps with new channel data, HBM budget 200 GB, 2 slots; the load generator makes each "
                     "request's inputs in HBM before submitting it, cubes left in HBM"),
        "unit": UNIT,
        "launch": {"waves": waves, "chunk": chunk,
                   "tile_cols": "autotuned per job class" if tuner is not None else "default"},
        "server": stats,
    })
    return summ


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the gridding path shards by map bands or jobs, not across GPUs here.", file=sys.stderr)
        sys.exit(2)

    import torch

    from hcgrid_amd import gridder, synth

    log = lambda s: print(s, file=sys.stderr, flush=True)  # noqa: E731
    if not torch.cuda.is_available():
        raise SystemExit("bench.py: no GPU visible")
    dev = torch.device("cuda", 0)

    if args.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 1 GiB read + 1 GiB written")
        del src, dst

    t0 = time.perf_counter()
    wl = synth.raster_20m() if args.workload == "raster_20m" else synth.raster_250k()
    t = wl.target
    xw, yw = gridder.wcs_sin(t.side, t.side, t.crpix, t.crpix, -t.cdelt, t.cdelt, t.crval1, t.crval2)
    log(f"workload {wl.name}: n={wl.n} side={t.side} generated in {time.perf_counter() - t0:.1f}s")
    ms_step, grid_avg, dres = run_plain_device(wl, xw, yw, args.steps, args.warmup, log)
    hp = dres["hp"]

    m = t.side * t.side
    abytes = grid_algorithmic_bytes(wl.n, m)
    achieved = abytes / (grid_avg * 1e-3) / 1e9
    counters = counter_fields(f"k_grid_tiles/{wl.name}", grid_avg)
    roof = {
        "kernel": "k_grid_tiles",
        "bound": "valu",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "algorithmic_bytes": abytes,
        "avg_ms": grid_avg,
        **counters,
        "primary_frac": counters["valu"]["frac"] if counters["valu"] else None,
        "note": ("FP64-VALU-issue-bound: frac is the HBM fraction of the algorithmic bytes (low by construction); "
                 "primary_frac = valu.frac, the share of the SIMDs' VALU issue slots its VALU instructions fill"),
    }
    result = {
        "metric": METRIC,
        "value": wl.n * 1 / (ms_step * 1e-3),
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{wl.name}: {wl.n} samples (raster scan), {t.side}x{t.side} SIN map at {t.cdelt} deg, "
                         f"GAUSS1D sigma {wl.sigma} deg, radius {wl.radius} deg, nside {hp.nside}, 1 channel, "
                         "FP64 coordinates, FP32 values, FP64 accumulation"),
            "samples": wl.n,
            "channels": 1,
            "cells": m,
            "usedrings": dres["meta"]["usedrings"],
        },
        "roofline": roof,
    }

    if args.cpu_baseline:
        med, total, ref = cpu_plain(wl, xw, yw, args.cpu_reps, log)
        result["error"] = plain_error(dres, ref)
        log(f"error {wl.name}: {result['error']}")
        cb = cpu_baseline_entry(wl, med, total, args.cpu_reps)
        # the small configuration as well, on the device once and on the CPU in full
        if wl.name != "raster_250k":
            small = synth.raster_250k()
            st = small.target
            sxw, syw = gridder.wcs_sin(st.side, st.side, st.crpix, st.crpix, -st.cdelt, st.cdelt, st.crval1,
                                       st.crval2)
            s_ms, _, sres = run_plain_device(small, sxw, syw, max(args.steps, 5), 1, log)
            smed, stotal, sref = cpu_plain(small, sxw, syw, args.cpu_reps, log)
            cb["by_config"] = {small.name: dict(cpu_baseline_entry(small, smed, stotal, args.cpu_reps),
                                                gpu_value=small.n / (s_ms * 1e-3),
                                                error=plain_error(sres, sref))}
        result["cpu_baseline"] = cb
    torch.cuda.empty_cache()
    if args.cube:
        result["multichannel"] = bench_cube(args, log)
        torch.cuda.empty_cache()
    tuner = None
    if args.batch:
        result["batched"], tuner = bench_batch(args, log)
        torch.cuda.empty_cache()
    if args.served:
        result["served"] = bench_served(args, log, tuner)
        torch.cuda.empty_cache()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
