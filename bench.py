"""Benchmark of the single-channel gridding path on one MI355X.

One step = one pass of the hot path over a synthetic raster-scan workload
whose inputs are already resident in HBM: HEALPix keys + stable radix sort,
permutation gather, ring LUT, zeroed output cubes, the gridding kernel and
the output normalisation.  Default workload: 20M samples (1000 scan lines x
20000 samples over 20 x 20 deg) onto a 2048 x 2048 SIN map at 0.01 deg,
GAUSS1D sigma 0.02 deg, radius 0.06 deg, nside 8192.

Prints one JSON line (rank 0).  `value` = samples x channels / step time.
The roofline object covers the gridding kernel (the dominant launch), timed
with HIP events on the stream it is launched on; the CPU baseline is the
oracle (the C/OpenMP restatement) on a bounded sample of the same workload.

The `multichannel` object (on by default, --no-cube to skip) measures the
MFMA path on a 2M-sample x 1024-channel BF16 raster onto a 1024 x 1024 map
and the single-channel path applied channel by channel to the same samples.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sample·channel updates/s and p50/p99 job latency at max rel err ≤1e-3 vs FP64"
HBM_PEAK_GBS = 8000.0
# VALU issue capacity: 256 CUs x 4 SIMDs, one wave64 VALU instruction per
# SIMD every 4 cycles, at the 2.4 GHz peak engine clock
SIMDS, VALU_CYCLES, CLOCK_HZ = 1024, 4, 2.4e9
FP64_VECTOR_PEAK_TFLOPS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="raster_20m", choices=["raster_20m", "raster_250k"])
    ap.add_argument("--cpu-baseline", dest="cpu_baseline", action="store_true", default=True)
    ap.add_argument("--no-cpu-baseline", dest="cpu_baseline", action="store_false")
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


def measured_traffic(key: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, calibrated on a device copy of known size), kept in
    profiles/traffic.json; None where no measurement is recorded."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            v = json.load(f).get(key)
        return int(v) if v is not None else None
    except (OSError, ValueError):
        return None


def valu_issue(key: str, avg_ms: float):
    """VALU issue share of a launch: the wave instructions rocprofv3 counted
    for it (SQ_INSTS_VALU, kept in profiles/traffic.json under valu_insts)
    at 4 cycles each over 1024 SIMDs, against its measured time."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            insts = json.load(f).get("valu_insts", {}).get(key)
    except (OSError, ValueError):
        insts = None
    if insts is None or not avg_ms:
        return None
    issue_ms = insts * VALU_CYCLES / (SIMDS * CLOCK_HZ) * 1e3
    return {"insts": insts, "issue_ms": issue_ms, "frac": issue_ms / avg_ms}


def grid_algorithmic_bytes(n: int, m: int) -> int:
    # per sorted sample: 32 B record + 4 B key; per cell: xwcs, ywcs (16 B)
    # and the in-place accumulators read + written (32 B)
    return n * 36 + m * 48


def cpu_baseline(wl, xw, yw, log):
    """Oracle timing: full key computation + stable sort + LUT over all
    samples, gridding of a band of rows, scaled to the whole map."""
    import numpy as np

    import oracle

    side = wl.target.side
    nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    h, g = oracle.prepare_grid_kernel(oracle.GAUSS1D, [wl.sigma], wl.radius, wl.hpx_res, side)
    t0 = time.perf_counter()
    keys, perm = oracle.presort(h, wl.lon, wl.lat)
    t1 = time.perf_counter()
    lon_s, lat_s = wl.lon[perm], wl.lat[perm]
    data_s = wl.values.astype(np.float64)[perm]
    t2 = time.perf_counter()
    fr, used, lut = oracle.ring_lut(h, keys, wl.n)
    t3 = time.perf_counter()
    rows = max(16, side // 16)
    r0 = side // 2 - rows // 2
    tg = time.perf_counter()
    oracle.grid_compat_rows(h, g, lon_s, lat_s, data_s, None, keys, fr, used, lut, xw, yw, r0, r0 + rows,
                            nthreads=nthreads)
    t4 = time.perf_counter()
    grid_full = (t4 - tg) * side / rows
    total = (t1 - t0) + (t2 - t1) + (t3 - t2) + grid_full
    log(f"cpu_baseline: keys+sort {t1 - t0:.2f}s gather {t2 - t1:.2f}s lut {t3 - t2:.3f}s "
        f"grid {rows} rows {t4 - tg:.2f}s -> full map {grid_full:.1f}s ({nthreads} threads)")
    return {
        "value": wl.n / total,
        "unit": "sample·channel updates/s",
        "cores": nthreads,
        "kind": "port",
        "sample": (f"{wl.name}: keys, stable sort, gathers and ring LUT over all {wl.n} samples; "
                   f"gridding of rows {r0}-{r0 + rows - 1} ({rows} of {side}), time scaled x{side / rows:g} "
                   f"to the full map"),
    }


def cube_algorithmic_bytes(n: int, m: int, nchan: int) -> int:
    # channel values (bf16) read once; per sample a 32 B record + 4 B key;
    # per cell the target coordinates (16 B) and the weight sum (4 B); the
    # FP32 cube written once
    return n * nchan * 2 + n * 36 + m * 20 + m * nchan * 4


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
    step_ms = [a.elapsed_time(c) for a, b, c in evs]
    grid_ms = [b.elapsed_time(c) for a, b, c in evs]
    ms_step = sum(step_ms) / len(step_ms)
    ms_grid = sum(grid_ms) / len(grid_ms)
    updates = wl.n * wl.nchan
    log(f"cube: {ms_step:.3f} ms/step, k_grid_mfma avg {ms_grid:.3f} ms")

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

    m = t.side * t.side
    abytes = cube_algorithmic_bytes(wl.n, m, wl.nchan)
    achieved = abytes / (ms_grid * 1e-3) / 1e9
    return {
        "workload": (f"{wl.name}: {wl.n} samples (raster scan, 500 lines x 4000) x {wl.nchan} BF16 channels, "
                     f"{t.side}x{t.side} SIN map at 0.01 deg, GAUSS1D sigma {wl.sigma} deg, radius {wl.radius} deg, "
                     f"nside {hp.nside}; FP64 weights rounded to BF16, FP32 accumulation"),
        "value": cube_rate,
        "unit": "sample·channel updates/s",
        "ms_per_step": ms_step,
        "dtype": "bf16",
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
            "traffic": measured_traffic(f"k_grid_mfma/{wl.name}"),
            "algorithmic_bytes": abytes,
            "avg_ms": ms_grid,
            "valu": valu_issue(f"k_grid_mfma/{wl.name}", ms_grid),
        },
    }


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
                         lon=d(sp.lon), lat=d(sp.lat), xw=d(xw), yw=d(yw), kern=kern, hp=hp, spec=sp))
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
    return {
        "workload": (f"{len(specs)} jobs: N log-uniform in [5e4, 5e6] (sum {sum(sp.n for sp in specs)}), side in "
                     "{128..2048} at 0.01 deg, channels in {1,16,64,256,512} BF16/FP16, sigma in {0.015,0.02,0.03} "
                     "deg, Dec in [-60, 80]; one graph = every job's geometry build + one persistent launch"),
        "value": rate,
        "unit": "sample·channel updates/s",
        "ms_per_batch": ms,
        "updates": updates,
        "work_items": st["nitems"],
        "launch": {"waves": waves, "chunk": chunk, "blocks": st["blocks"]},
        "autotuned_classes": len(tuner.tiles),
        "channel_by_channel": {"value": chan_rate, "ms_total": tot_ms},
        "speedup_vs_channel_by_channel": rate / chan_rate,
    }


def bench_served(args, log):
    """Served gridding: Poisson arrivals at 20 jobs/s drawn from the batched
    mix (30% repeats of one of the last 32 distinct maps with new channel
    data), HBM admission budget 200 GB, two slots; p50/p99 of completion -
    arrival and throughput over the run.  Inputs are made on the device at
    admission; results stay in HBM."""
    from hcgrid_amd import serve

    warm = serve.request_stream(12, rate=20.0, seed=99)
    srv = serve.GriddingServer(budget_bytes=200e9, slots=2)
    serve.run_load(srv, warm)
    srv.close()
    stream = serve.request_stream(args.served_jobs, rate=20.0, seed=4, repeat_frac=0.3, recent=32)
    srv = serve.GriddingServer(budget_bytes=200e9, slots=2)
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
                     "maps with new channel data, HBM budget 200 GB, 2 slots; inputs made on the device at "
                     "admission, cubes left in HBM"),
        "unit": "sample·channel updates/s",
        "server": stats,
    })
    return summ


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the gridding path shards by map bands or jobs, not across GPUs here.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
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

    for i in range(args.warmup):
        step()
        log(f"warmup {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()

    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    start.record(stream)
    for i in range(args.steps):
        step(evs[i])
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms_total = start.elapsed_time(stop)
    ms_step = ms_total / args.steps
    grid_ms = [a.elapsed_time(b) for a, b in evs]
    grid_avg = sum(grid_ms) / len(grid_ms)
    log(f"timed {args.steps} steps: {ms_step:.3f} ms/step (events), wall {wall * 1e3 / args.steps:.3f} ms/step; "
        f"grid kernel avg {grid_avg:.3f} ms (min {min(grid_ms):.3f}, max {max(grid_ms):.3f})")

    meta = g.meta()
    m = t.side * t.side
    abytes = grid_algorithmic_bytes(wl.n, m)
    achieved = abytes / (grid_avg * 1e-3) / 1e9
    result = {
        "metric": METRIC,
        "value": wl.n * 1 / (ms_step * 1e-3),
        "unit": "sample·channel updates/s",
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
            "usedrings": meta["usedrings"],
        },
        "roofline": {
            "kernel": "k_grid_tiles",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(f"k_grid_tiles/{wl.name}"),
            "algorithmic_bytes": abytes,
            "avg_ms": grid_avg,
            "valu": valu_issue(f"k_grid_tiles/{wl.name}", grid_avg),
            "note": "FP64-VALU-bound kernel: the HBM fraction is low by construction; valu.frac is its share "
                    "of the SIMDs' VALU issue slots",
        },
    }

    if args.cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(wl, xw, yw, log)
    if args.cube:
        del g, lon, lat, val, xwd, ywd
        torch.cuda.empty_cache()
        result["multichannel"] = bench_cube(args, log)
        torch.cuda.empty_cache()
    if args.batch:
        result["batched"] = bench_batch(args, log)
        torch.cuda.empty_cache()
    if args.served:
        result["served"] = bench_served(args, log)
        torch.cuda.empty_cache()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
