"""Served gridding: continuous batching with HBM-footprint admission.

A request is one multi-channel gridding job on the caller's data: sample
coordinates (lon, lat in degrees), channel values [nchan][n] and a target
map.  The data may be device tensors (torch, with an optional ready event)
or host arrays (numpy), which the server copies to the device on the
request's slot stream, so that copy is part of the request's latency.

The server keeps a small number of slots (streams).  Whenever a slot is
free it forms a round from every pending request that fits the HBM budget
(FIFO, no overtaking), enqueues the round on that slot's stream (input
copies, geometry builds for new coordinates, one persistent batched launch)
and goes back to admitting, so rounds on different slots overlap on the
device while new requests keep arriving.  A round's memory returns to the
budget when its completion event fires.

Geometries (sorted samples, ring LUT, target coordinates) stay resident in
an LRU cache keyed by the content of the coordinates (a 128-bit hash made on
the device for device inputs, on the host for host inputs), the target and
the kernel: a repeat request with new channel data skips the pre-sort.
Cached geometries count against the budget; a geometry is pinned from the
moment a request is admitted against it until that request completes, and
only unpinned ones are evicted, least recently used first.

A request that cannot be served -- invalid arguments, larger than the whole
budget, or a failure while its work is set up or launched -- completes with
`error` set and `cube` None; the other requests of its round go on.

When a request completes (served or failed) the server drops its
references to the request's inputs (lon, lat, values and the ready event
become None), so a long run holds only the caller's own references and,
with keep_results, the cubes.

The reference has no server (HCGrid/HCGrid.cpp:38-197 grids one map per
process run); this is the request loop around the batched path, and
HCGrid+'s overlap of host work with device work (HCGrid+/gridding-omp.cu:
616-626) generalised to independent jobs.

The device side goes through a backend object (TorchBackend here); the
admission, caching and failure logic only sees its interface, so it can be
exercised on the CPU with a stand-in.
"""
from __future__ import annotations

import collections
import dataclasses
import hashlib
import math
import queue
import threading
import time
import traceback

import numpy as np

from . import autotune, capi, synth

DTYPES = ("bf16", "f16")


@dataclasses.dataclass(eq=False)
class Request:
    rid: int
    lon: object            # float64 [n] degrees: torch tensor on the GPU, or numpy array
    lat: object
    values: object         # [nchan][n] channel-major: torch bf16/f16 on the GPU, or numpy float16/32/64
    target: synth.Target   # square zenithal map (proj, side, cdelt, crval1, crval2, crpix)
    sigma: float           # GAUSS1D sigma in degrees; radius 3 sigma, HEALPix resolution sigma / 2
    dtype: str = "bf16"    # channel type of the contraction (host values are converted to it)
    geo_id: object = None  # the caller's name for the coordinate set; None = hash of the coordinates
    ready: object = None   # torch.cuda.Event after which device inputs are complete (None: the
                           # submitting thread's current stream at submit())
    repeat: bool = False   # bookkeeping of load generators
    # filled by the server
    arrival: float = 0.0
    admitted: float = 0.0
    done: float = 0.0
    geometry_hit: bool = False
    footprint: int = 0
    geo_key: object = None
    cube: object = None    # float32 [nchan][side*side] on the device (kept with keep_results)
    wsum: object = None
    error: object = None
    trace: str = ""        # the failure's formatted traceback (the exception keeps none)
    _pinned: object = None  # the cache entry this request holds a reference on
    _held: object = None    # device inputs kept alive until the round completes
    _shape: object = None   # (n, nchan), recorded at submit: the inputs are released on completion

    @property
    def n(self) -> int:
        if self._shape is not None:
            return self._shape[0]
        return int(len(self.lon)) if isinstance(self.lon, np.ndarray) else int(self.lon.numel())

    @property
    def nchan(self) -> int:
        if self._shape is not None:
            return self._shape[1]
        return int(self.values.shape[0])

    @property
    def updates(self) -> int:
        return self.n * self.nchan

    @property
    def on_device(self) -> bool:
        return not isinstance(self.lon, np.ndarray)


def _is_np(a) -> bool:
    return isinstance(a, np.ndarray)


def validate(r: Request) -> None:
    """Raises HcgError if the request cannot be gridded as given."""
    def bad(msg):
        raise capi.HcgError(capi.HCG_ERR_INVALID, f"request {r.rid}", msg)

    if r.dtype not in DTYPES:
        bad(f"dtype {r.dtype!r} (expected 'bf16' or 'f16')")
    if not (isinstance(r.sigma, (int, float)) and r.sigma > 0 and math.isfinite(r.sigma)):
        bad(f"sigma {r.sigma!r}")
    t = r.target
    if int(t.side) < 1 or not (t.cdelt > 0) or t.proj not in capi.PROJECTIONS:
        bad(f"target {t}")
    host = [_is_np(a) for a in (r.lon, r.lat, r.values)]
    if any(host) != all(host):
        bad("lon, lat and values must all be host arrays or all device tensors")
    if all(host):
        if r.lon.ndim != 1 or r.lat.shape != r.lon.shape or r.lon.size < 1:
            bad(f"lon/lat shapes {r.lon.shape}, {r.lat.shape}")
        if r.lon.dtype != np.float64 or r.lat.dtype != np.float64:
            bad("lon and lat must be float64")
        if r.values.ndim != 2 or r.values.shape[1] != r.lon.size or r.values.shape[0] < 1:
            bad(f"values shape {r.values.shape}, expected (nchan, {r.lon.size})")
        if r.values.dtype not in (np.float16, np.float32, np.float64):
            bad(f"host values of dtype {r.values.dtype}")
        return
    import torch
    from .gridder import check_tensor
    n = int(r.lon.numel())
    if n < 1 or r.lon.dim() != 1:
        bad("lon must be a non-empty 1-D tensor")
    check_tensor(r.lon, f"request {r.rid}", "lon", "f64", n)
    check_tensor(r.lat, f"request {r.rid}", "lat", "f64", n, r.lon.device)
    if r.values.dim() != 2 or int(r.values.shape[1]) != n or int(r.values.shape[0]) < 1:
        bad(f"values shape {tuple(r.values.shape)}, expected (nchan, {n})")
    check_tensor(r.values, f"request {r.rid}", "values", r.dtype, r.values.numel(), r.lon.device)
    if r.ready is not None and not isinstance(r.ready, torch.cuda.Event):
        bad("ready must be a torch.cuda.Event")


class TorchBackend:
    """Device side of the server: streams and events through torch,
    geometries and batched launches through the C ABI (hcgrid_amd.batch)."""

    def __init__(self):
        import torch
        self.torch = torch
        capi.load()
        self._kern = {}
        self._hash_stream = None
        self._lock = threading.Lock()

    def new_stream(self):
        return self.torch.cuda.Stream()

    def new_event(self):
        return self.torch.cuda.Event()

    def kernel(self, sigma: float):
        from . import gridder
        if sigma not in self._kern:
            self._kern[sigma] = gridder.prepare_grid_kernel(capi.GAUSS1D, [sigma], 3 * sigma, sigma / 2)
        return self._kern[sigma]

    def job_bytes(self, r: Request, waves: int) -> int:
        """HBM a request holds while it runs, its geometry aside."""
        m = r.target.side * r.target.side
        n_pad = int(capi.load().hcg_channels_padded_len(r.n))
        values = r.nchan * r.n * 2              # the device copy (or the caller's tensor)
        sorted_values = r.nchan * n_pad * 2 + 256
        cube = r.nchan * m * 4 + m * 4
        tiles = -(-r.target.side // 8) * -(-r.target.side // 8)
        # tile plans (3600 bytes per 64-cell tile) and work items (16 bytes per tile and channel group)
        plan = 4096 + tiles * 3600 + 16 * tiles * max(1, -(-r.nchan // (64 * max(1, waves))))
        coords = 16 * r.n
        # host values land on the device in their own type before the conversion
        upload = 0 if r.on_device else r.values.size * r.values.itemsize
        return values + sorted_values + cube + plan + coords + upload

    def geo_bytes(self, r: Request) -> int:
        from .batch import Geometry
        _, hp = self.kernel(r.sigma)
        return Geometry.bytes_for(r.n, r.target.side, hp) + 256

    def content_key(self, r: Request) -> tuple:
        """128-bit content hash of lon and lat: hcg_content_hash on a side
        stream for device inputs (after their ready event), blake2b for host
        arrays.  The two kinds never match each other (a repeat submitted the
        other way misses the cache, never hits a wrong entry)."""
        if not r.on_device:
            h = hashlib.blake2b(digest_size=16)
            h.update(np.ascontiguousarray(r.lon).view(np.uint8))
            h.update(np.ascontiguousarray(r.lat).view(np.uint8))
            return ("host", h.hexdigest())
        torch = self.torch
        if r.ready is None:
            r.ready = torch.cuda.Event()
            r.ready.record()  # the submitting thread's current stream
        with self._lock:
            if self._hash_stream is None:
                self._hash_stream = torch.cuda.Stream()
            s = self._hash_stream
            with torch.cuda.stream(s):
                out = torch.empty(4, dtype=torch.int64, device=r.lon.device)
                s.wait_event(r.ready)
                h = capi.C.c_void_p(s.cuda_stream)
                capi.call("hcg_content_hash", capi.C.c_void_p(r.lon.data_ptr()), r.lon.numel() * 8,
                          capi.C.c_void_p(out.data_ptr()), h)
                capi.call("hcg_content_hash", capi.C.c_void_p(r.lat.data_ptr()), r.lat.numel() * 8,
                          capi.C.c_void_p(out.data_ptr() + 16), h)
                words = out.cpu()  # synchronises the hash stream only
        return ("device",) + tuple(int(x) & 0xFFFFFFFFFFFFFFFF for x in words.tolist())

    def prepare_inputs(self, r: Request, stream):
        """Device lon, lat and channel values of the request, on `stream`."""
        torch = self.torch
        with torch.cuda.stream(stream):
            if r.on_device:
                if r.ready is not None:
                    stream.wait_event(r.ready)
                return r.lon, r.lat, r.values
            dev = torch.device("cuda", torch.cuda.current_device())

            def up(a):
                return torch.from_numpy(np.ascontiguousarray(a)).pin_memory().to(dev, non_blocking=True)

            tdt = torch.bfloat16 if r.dtype == "bf16" else torch.float16
            return up(r.lon), up(r.lat), up(r.values).to(tdt).contiguous()

    def build_geometry(self, r: Request, lon, lat, stream):
        from .batch import Geometry
        kern, hp = self.kernel(r.sigma)
        with self.torch.cuda.stream(stream):
            g = Geometry(kern, hp, r.n, r.target.side, in_torch=True)
            try:
                g.set_target(r.target, stream=stream)
                g.build(lon, lat, stream=stream)
            except Exception:
                g.close()
                raise
        return g

    def outputs(self, r: Request, stream):
        torch = self.torch
        m = r.target.side * r.target.side
        with torch.cuda.stream(stream):
            return (torch.empty((r.nchan, m), dtype=torch.float32, device="cuda"),
                    torch.empty(m, dtype=torch.float32, device="cuda"))

    def launch(self, jobs: list, waves: int, chunk: int, stream):
        from .batch import Batch
        b = Batch(jobs, waves=waves, chunk=chunk, stream=stream)
        try:
            b.enqueue(stream=stream)
        except Exception:
            b.close()
            raise
        return b

    def synchronize(self):
        self.torch.cuda.synchronize()


class _GeoEntry:
    __slots__ = ("geometry", "bytes", "refs", "last", "ready")

    def __init__(self, geometry, nbytes):
        self.geometry, self.bytes = geometry, nbytes
        self.refs = 0
        self.last = time.perf_counter()
        self.ready = None


class GriddingServer:
    def __init__(self, budget_bytes: float = 200e9, slots: int = 2, waves: int = 4, chunk: int = 256,
                 tuner: autotune.Autotuner | None = None, keep_results: bool = False, poll_s: float = 2e-4,
                 max_round: int = 64, backend=None):
        self.be = backend if backend is not None else TorchBackend()
        self.budget = int(budget_bytes)
        self.used = 0
        self.waves, self.chunk = int(waves), int(chunk)
        self.tuner = tuner
        self.keep = keep_results
        self.poll_s = poll_s
        self.max_round = int(max_round)
        self.inbox: queue.SimpleQueue = queue.SimpleQueue()
        self.pending: collections.deque = collections.deque()
        self.geoms: collections.OrderedDict = collections.OrderedDict()
        self.streams = [self.be.new_stream() for _ in range(int(slots))]
        self.rounds = [None] * int(slots)
        self.finished: list[Request] = []
        self.failed: list[Request] = []
        self.stats = collections.Counter()
        self._abort = None

    # -- client side ---------------------------------------------------------
    def submit(self, req: Request):
        """Queues a request (any thread).  Validates it and computes its
        geometry key here, on the caller's thread; an invalid request is
        queued already failed."""
        if not req.arrival:
            req.arrival = time.perf_counter()
        try:
            validate(req)
            req._shape = (req.n, req.nchan)
            t = req.target
            coords = ("id", req.geo_id) if req.geo_id is not None else self.be.content_key(req)
            req.geo_key = (coords, req.n, t.proj, t.side, t.cdelt, t.crval1, t.crval2, t.crpix, float(req.sigma))
        except Exception as exc:  # reported on the request, never raised into the loop
            self._set_error(req, exc)
        self.inbox.put(req)

    def abort(self, exc: BaseException):
        """Makes serve() raise (from any thread): e.g. the producer of the
        requests failed and the remaining ones will never arrive."""
        self._abort = exc

    # -- footprint and the geometry cache --------------------------------------
    def _evict(self, need: int) -> bool:
        """Frees unpinned cached geometries (LRU) until `need` more bytes fit."""
        for key in list(self.geoms.keys()):
            if self.used + need <= self.budget:
                break
            e = self.geoms[key]
            if e.refs == 0:
                self.used -= e.bytes
                e.geometry.close()
                del self.geoms[key]
                self.stats["evictions"] += 1
        return self.used + need <= self.budget

    @staticmethod
    def _release_inputs(r: Request):
        if r._shape is None:
            try:
                r._shape = (r.n, r.nchan)
            except Exception:  # malformed inputs (failed validation)
                r._shape = (0, 0)
        r.lon = r.lat = r.values = r.ready = None
        r._held = None

    @staticmethod
    def _set_error(r: Request, exc: BaseException):
        # the traceback as text: its frames would keep the round's other
        # requests' data (and this request's) alive as long as the error
        r.trace = "".join(traceback.format_exception(type(exc), exc, exc.__traceback__))
        e = exc
        for _ in range(8):  # the exception and its causes
            if e is None:
                break
            e.__traceback__ = None
            e = e.__cause__ or e.__context__
        r.error = exc

    def _fail(self, r: Request, exc: Exception):
        self._set_error(r, exc)
        r.done = time.perf_counter()
        self.used -= r.footprint
        r.footprint = 0
        if r._pinned is not None:
            r._pinned.refs -= 1
            r._pinned = None
        self._release_inputs(r)
        self.failed.append(r)
        self.stats["failed"] += 1

    # -- rounds ----------------------------------------------------------------
    def _admit(self) -> list:
        """Pending requests (FIFO) that fit the budget now, up to max_round.
        A request whose geometry is cached pins that entry here, so no later
        eviction in this pass (its own included) can close it."""
        take = []
        reserved = set()  # geometries that this round's earlier misses will build
        while self.pending and len(take) < self.max_round:
            r = self.pending[0]
            if r.error is not None:  # failed at submit
                self.pending.popleft()
                r.done = time.perf_counter()
                self._release_inputs(r)
                self.failed.append(r)
                self.stats["failed"] += 1
                continue
            need = self.be.job_bytes(r, self.waves)
            entry = self.geoms.get(r.geo_key)
            hit = entry is not None or r.geo_key in reserved
            if entry is not None:
                entry.refs += 1  # pin before any eviction below
            elif not hit:
                need += self.be.geo_bytes(r)
            if self.used + need > self.budget and not self._evict(need):
                if entry is not None:
                    entry.refs -= 1
                in_flight = any(x is not None for x in self.rounds)
                if not take and not in_flight:
                    # cannot fit even on an idle device with every geometry evicted
                    # (with no round in flight no request holds a geometry)
                    pinned = sum(e.refs for e in self.geoms.values())
                    if pinned:
                        raise RuntimeError(f"served loop: {pinned} geometry references held with no round in flight")
                    self.pending.popleft()
                    self._fail(r, capi.HcgError(capi.HCG_ERR_NOMEM, f"request {r.rid}",
                                                f"needs {need} bytes, budget {self.budget}"))
                    self.stats["rejected"] += 1
                    continue
                break
            self.pending.popleft()
            self.used += need
            r.footprint = need
            r.geometry_hit = hit
            r._pinned = entry
            if not hit:
                reserved.add(r.geo_key)
            take.append(r)
        return take

    def _geometry_for(self, r: Request, lon, lat, s, now: float) -> _GeoEntry:
        """The request's cache entry, pinned by it, built on `s` if new."""
        e = r._pinned
        if e is None:
            e = self.geoms.get(r.geo_key)
            if e is not None:  # built earlier in this round by a request with the same key
                e.refs += 1
                r._pinned = e
        if e is not None:
            self.geoms.move_to_end(r.geo_key)
            if e.ready is not None:
                s.wait_event(e.ready)  # possibly built on another slot: wait on the device
            self.stats["geometry_hits"] += 1
        else:
            gb = self.be.geo_bytes(r)
            if r.geometry_hit:  # its builder failed: build it here, charged now
                self.used += gb
                r.footprint += gb
            g = self.be.build_geometry(r, lon, lat, s)
            e = _GeoEntry(g, gb)
            e.ready = self.be.new_event()
            e.ready.record(s)
            e.refs = 1
            r._pinned = e
            self.geoms[r.geo_key] = e
            self.stats["geometry_builds"] += 1
            r.footprint -= gb  # the cached geometry stays charged until evicted
        e.last = now
        return e

    def _tile_cols(self, r: Request) -> int:
        if self.tuner is None:
            return 0
        return self.tuner.tiles.get(autotune.job_class(r.n, r.target.side, r.nchan, r.sigma, r.dtype), 0)

    def _launch(self, slot: int, reqs: list):
        s = self.streams[slot]
        now = time.perf_counter()
        jobs, ok = [], []
        for r in reqs:
            r.admitted = now
            try:
                lon, lat, vals = self.be.prepare_inputs(r, s)
                e = self._geometry_for(r, lon, lat, s, now)
                cube, wsum = self.be.outputs(r, s)
                jobs.append(dict(geometry=e.geometry, values=vals, cube=cube, wsum=wsum, dtype=r.dtype,
                                 tile_cols=self._tile_cols(r)))
                r._held = (lon, lat, vals, cube, wsum)
                ok.append(r)
            except Exception as exc:
                self._fail(r, exc)
        if not ok:
            return
        try:
            batches = [self.be.launch(jobs, self.waves, self.chunk, s)]
        except Exception:
            # find the request(s) at fault: each job in a launch of its own
            batches, keep = [], []
            for r, j in zip(ok, jobs):
                try:
                    batches.append(self.be.launch([j], self.waves, self.chunk, s))
                    keep.append(r)
                except Exception as exc:
                    self._fail(r, exc)
            ok = keep
            if not ok:
                return
        ev = self.be.new_event()
        ev.record(s)
        self.rounds[slot] = (ok, batches, ev)
        self.stats["rounds"] += 1
        self.stats["round_jobs_max"] = max(self.stats["round_jobs_max"], len(ok))

    def _reap(self) -> int:
        done = 0
        for slot, rd in enumerate(self.rounds):
            if rd is None:
                continue
            reqs, batches, ev = rd
            if not ev.query():
                continue
            t = time.perf_counter()
            for r in reqs:
                r.done = t
                self.used -= r.footprint
                r._pinned.refs -= 1
                r._pinned = None
                if self.keep:
                    r.cube, r.wsum = r._held[3], r._held[4]
                self._release_inputs(r)
                self.finished.append(r)
            for b in batches:
                b.close()
            self.rounds[slot] = None
            done += len(reqs)
        return done

    # -- main loop -----------------------------------------------------------------
    def serve(self, n_total: int, timeout_s: float = 3600.0, progress=None, every: int = 500):
        """Runs until n_total requests have completed (served or failed);
        progress(done, total) is called each time `every` more finish.
        Raises if abort() was called or on the timeout."""
        t_end = time.perf_counter() + timeout_s
        next_report = every
        while len(self.finished) + len(self.failed) < n_total:
            if progress is not None and len(self.finished) >= next_report:
                progress(len(self.finished), n_total)
                next_report += every
            if self._abort is not None:
                raise RuntimeError(f"served loop aborted after {len(self.finished)} of {n_total} requests: "
                                   f"{self._abort!r}") from self._abort
            if time.perf_counter() > t_end:
                raise TimeoutError(f"served {len(self.finished)} of {n_total} requests before the timeout")
            busy = False
            while True:
                try:
                    self.pending.append(self.inbox.get_nowait())
                except queue.Empty:
                    break
            if self._reap():
                busy = True
            for slot in range(len(self.rounds)):
                if self.rounds[slot] is None and self.pending:
                    reqs = self._admit()
                    if reqs:
                        self._launch(slot, reqs)
                        busy = True
            if not busy:
                time.sleep(self.poll_s)
        self.be.synchronize()
        return self.finished

    def close(self):
        self.be.synchronize()
        for e in self.geoms.values():
            e.geometry.close()
        self.geoms.clear()
        self.used = 0


# -- load generation (the configured served mix) ----------------------------------------

@dataclasses.dataclass
class RequestSpec:
    """A synthetic request of the batched-job distribution; make() creates
    its coordinates and values on the device."""
    rid: int
    n: int
    side: int
    nchan: int
    dtype: str
    sigma: float
    dec: float
    geo_seed: int          # same geo_seed and shape = same samples and target
    val_seed: int          # channel data
    repeat: bool = False

    def __post_init__(self):
        self.n = synth.field_size(self.n)  # the raster's actual sample count

    def make(self, device: str = "cuda") -> Request:
        import torch
        lon, lat, tgt = synth.device_field(self.n, self.side, self.dec, self.geo_seed, device=device)
        vals = synth.device_values(self.nchan, self.dtype, lon, lat, tgt, self.val_seed, device=device)
        ev = torch.cuda.Event()
        ev.record()
        return Request(self.rid, lon, lat, vals, tgt, self.sigma, self.dtype, ready=ev, repeat=self.repeat)


def request_stream(n_requests: int, rate: float, seed: int = 4, repeat_frac: float = 0.3, recent: int = 32,
                   **dist) -> list:
    """Poisson arrivals (rate per second) of jobs drawn like batch_specs;
    a repeat_frac share reuses the coordinates and target of one of the last
    `recent` distinct maps with new channel data.  Returns (offset_s, RequestSpec)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sides = dist.get("sides", (128, 256, 512, 1024, 2048))
    chans = dist.get("chans", (1, 16, 64, 256, 512))
    sigmas = dist.get("sigmas", (0.015, 0.02, 0.03))
    n_lo, n_hi = dist.get("n_range", (5e4, 5e6))
    dec_lo, dec_hi = dist.get("dec_range", (-60.0, 80.0))
    t = 0.0
    out, originals = [], []
    for i in range(n_requests):
        t += rng.exponential(1.0 / rate)
        nchan = int(rng.choice(chans))
        dtype = "bf16" if rng.integers(2) == 0 else "f16"
        if originals and rng.uniform() < repeat_frac:
            o = originals[-recent:][int(rng.integers(min(recent, len(originals))))]
            r = RequestSpec(i, o.n, o.side, nchan, dtype, o.sigma, o.dec, o.geo_seed, seed * 100003 + i,
                            repeat=True)
        else:
            n = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))))
            side = int(rng.choice(sides))
            sigma = float(rng.choice(sigmas))
            dec = float(rng.uniform(dec_lo, dec_hi))
            r = RequestSpec(i, n, side, nchan, dtype, sigma, dec, seed * 1000003 + i, seed * 100003 + i)
            originals.append(r)
        out.append((t, r))
    return out


def run_load(server: GriddingServer, stream: list, time_scale: float = 1.0, progress=None):
    """A client thread makes each request's inputs on the device at its
    arrival offset and submits it once they are complete (its arrival time);
    the server loop runs in this thread.  Returns the finished list."""
    import torch
    t0 = time.perf_counter()
    errors = []
    stop = threading.Event()

    def client():
        try:
            for off, spec in stream:
                wait = t0 + off * time_scale - time.perf_counter()
                if stop.wait(wait) if wait > 0 else stop.is_set():
                    return
                r = spec.make()
                r.ready.synchronize()
                r.arrival = time.perf_counter()
                server.submit(r)
                del r  # the server holds the request's inputs until it completes
        except Exception as exc:  # stops the loop, raised in its thread
            errors.append(exc)
            server.abort(exc)

    th = threading.Thread(target=client, daemon=True)
    th.start()
    try:
        done = server.serve(len(stream), progress=progress)
    finally:
        stop.set()  # the loop ended (or raised): the client stops producing
        th.join()
        torch.cuda.synchronize()
    if errors:
        raise errors[0]
    return done


def latency_summary(done: list) -> dict:
    lat = np.array([r.done - r.arrival for r in done])
    qd = np.array([r.admitted - r.arrival for r in done])
    t0 = min(r.arrival for r in done)
    t1 = max(r.done for r in done)
    upd = sum(r.updates for r in done)
    return {
        "jobs": len(done),
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "mean_ms": float(lat.mean() * 1e3),
        "queue_p50_ms": float(np.percentile(qd, 50) * 1e3),
        "wall_s": t1 - t0,
        "updates": upd,
        "value": upd / (t1 - t0),
        "repeats": int(sum(r.repeat for r in done)),
        "geometry_hits": int(sum(r.geometry_hit for r in done)),
    }
