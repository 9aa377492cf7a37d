"""The served loop's admission, geometry cache and failure handling on the
CPU, with a stand-in for the device backend (no GPU, no library calls).

The stand-in keeps the server's interface: streams, events that complete
after a set number of polls (so rounds stay in flight for a while),
geometries with a size and a closed flag, and launches that check every
geometry they are given is still open and that the budget holds.  Failures
are injected per request id at each stage.
"""
import itertools

import numpy as np
import pytest

from hcgrid_amd import capi, serve, synth


class FakeEvent:
    def __init__(self, polls):
        self.left = polls

    def record(self, stream):
        pass

    def query(self):
        self.left -= 1
        return self.left < 0

    def synchronize(self):
        self.left = -1


class FakeStream:
    def wait_event(self, ev):
        pass


class FakeGeometry:
    ids = itertools.count()

    def __init__(self, nbytes):
        self.nbytes = nbytes
        self.closed = False
        self.id = next(FakeGeometry.ids)

    def close(self):
        self.closed = True


class FakeBatch:
    def __init__(self, jobs):
        self.jobs = jobs
        self.closed = False

    def close(self):
        self.closed = True


class FakeBackend:
    """Sizes: a job holds n * nchan bytes, a geometry n bytes."""

    def __init__(self, polls=2, fail_prepare=(), fail_build=(), fail_launch=()):
        self.polls = polls
        self.fail_prepare, self.fail_build, self.fail_launch = set(fail_prepare), set(fail_build), set(fail_launch)
        self.server = None
        self.rid_of = {}
        self.launches = []
        self.built = []
        self.max_used = 0

    def new_stream(self):
        return FakeStream()

    def new_event(self):
        return FakeEvent(self.polls)

    def job_bytes(self, r, waves):
        return r.n * r.nchan

    def geo_bytes(self, r):
        return r.n

    def content_key(self, r):
        return ("fake", r.lon.tobytes(), r.lat.tobytes())

    def prepare_inputs(self, r, stream):
        if r.rid in self.fail_prepare:
            raise capi.HcgError(capi.HCG_ERR_HIP, "prepare", f"injected failure of request {r.rid}")
        self.rid_of[id(r.values)] = r.rid  # what launch() sees of a job is its arrays
        return r.lon, r.lat, r.values

    def build_geometry(self, r, lon, lat, stream):
        if r.rid in self.fail_build:
            raise capi.HcgError(capi.HCG_ERR_HIP, "build", f"injected failure of request {r.rid}")
        g = FakeGeometry(r.n)
        self.built.append(g)
        return g

    def outputs(self, r, stream):
        return np.zeros((r.nchan, r.target.side ** 2), np.float32), np.zeros(r.target.side ** 2, np.float32)

    def launch(self, jobs, waves, chunk, stream):
        srv = self.server
        assert srv.used <= srv.budget, "budget exceeded while a round is launched"
        self.max_used = max(self.max_used, srv.used)
        for j in jobs:
            assert not j["geometry"].closed, "a launch was given an evicted geometry"
        bad = [self.rid_of[id(j["values"])] for j in jobs if self.rid_of[id(j["values"])] in self.fail_launch]
        if bad:
            raise capi.HcgError(capi.HCG_ERR_INVALID, "launch", f"injected failure of {bad}")
        for j in jobs:  # the "result": each channel's sum, in every cell
            j["cube"][:] = j["values"].sum(axis=1, dtype=np.float64)[:, None]
            j["wsum"][:] = 1.0
        self.launches.append(len(jobs))
        return FakeBatch(jobs)

    def synchronize(self):
        pass


def _server(be, **kw):
    srv = serve.GriddingServer(backend=be, keep_results=True, poll_s=0, **kw)
    be.server = srv
    return srv


def _req(rid, n=100, nchan=4, seed=None, side=8, sigma=0.02, coords=None):
    rng = np.random.default_rng(rid if seed is None else seed)
    if coords is None:
        lon = 180 + rng.uniform(-1, 1, n)
        lat = 30 + rng.uniform(-1, 1, n)
    else:
        lon, lat = coords
    vals = np.random.default_rng(1000 + rid).normal(0, 1, (nchan, lon.size)).astype(np.float32)
    return serve.Request(rid, lon, lat, vals, synth.Target(side, 0.01, 180.0, 30.0, side / 2), sigma, "bf16")


def _run(srv, reqs):
    for r in reqs:
        srv.submit(r)
    srv.serve(len(reqs), timeout_s=10)
    return srv


def test_every_request_served_and_budget_held():
    be = FakeBackend()
    srv = _server(be, budget_bytes=3000, slots=2)
    reqs = [_req(i, n=100 + 10 * i, nchan=3) for i in range(12)]
    vals = {r.rid: r.values for r in reqs}
    _run(srv, reqs)
    assert len(srv.finished) == 12 and not srv.failed
    assert be.max_used <= 3000
    for r in reqs:
        assert r.error is None
        np.testing.assert_allclose(r.cube[:, 0], vals[r.rid].sum(axis=1, dtype=np.float64))
    # every round's memory returned: only the cached geometries stay charged
    assert srv.used == sum(e.bytes for e in srv.geoms.values())
    assert all(e.refs == 0 for e in srv.geoms.values())


def test_repeat_coordinates_hit_the_cache():
    be = FakeBackend()
    srv = _server(be, budget_bytes=10 ** 6, slots=1)
    base = _req(0)
    lon, lat = base.lon.copy(), base.lat.copy()  # the server releases base's inputs when it completes
    reqs = [base] + [_req(i, coords=(lon.copy(), lat.copy())) for i in range(1, 5)]
    _run(srv, reqs)
    assert srv.stats["geometry_builds"] == 1 and srv.stats["geometry_hits"] == 4
    assert len(be.built) == 1
    # same coordinates, another kernel width: another geometry
    srv2 = _server(FakeBackend(), budget_bytes=10 ** 6, slots=1)
    _run(srv2, [_req(10, coords=(lon, lat)), _req(11, coords=(lon, lat), sigma=0.03)])
    assert srv2.stats["geometry_builds"] == 2


def test_hit_geometry_is_pinned_against_eviction_in_the_same_pass():
    # A cached geometry G (unreferenced) and a tight budget.  Request A hits
    # G; request B, admitted in the same pass, needs an eviction.  G must not
    # be the one evicted: A's launch would get a closed geometry.
    be = FakeBackend(polls=0)
    srv = _server(be, budget_bytes=1000, slots=1, max_round=8)
    a0 = _req(0, n=300, nchan=1)
    lon0, lat0 = a0.lon.copy(), a0.lat.copy()
    _run(srv, [a0])  # builds G (300 bytes), now cached and unreferenced
    assert srv.used == 300
    srv.submit(_req(5, n=200, nchan=1))
    srv.serve(2, timeout_s=10)  # G and H (200) cached: 500 used
    assert srv.used == 500
    a = _req(1, n=300, nchan=1, coords=(lon0, lat0))  # hit on G: needs 300
    b = _req(2, n=150, nchan=1)  # miss: needs 150 + 150
    srv.submit(a)
    srv.submit(b)
    srv.serve(4, timeout_s=10)
    assert a.error is None and b.error is None
    assert a.geometry_hit
    assert be.max_used <= 1000
    # H was evicted to make room, G was not (the stand-in's launch also
    # asserts that no launch received a closed geometry)
    g, h = be.built[0], be.built[1]
    assert srv.stats["evictions"] == 1 and h.closed and not g.closed


def test_request_larger_than_budget_is_rejected_others_served():
    be = FakeBackend()
    srv = _server(be, budget_bytes=2000, slots=2)
    reqs = [_req(0, n=100), _req(1, n=1000, nchan=4), _req(2, n=100)]
    _run(srv, reqs)
    assert reqs[1].error is not None and reqs[1].error.code == capi.HCG_ERR_NOMEM
    assert reqs[1].cube is None
    assert reqs[0].error is None and reqs[2].error is None
    assert srv.stats["rejected"] == 1


def test_invalid_request_fails_at_submit_without_launching():
    be = FakeBackend()
    srv = _server(be, budget_bytes=10 ** 6)
    bad_shape = _req(0)
    bad_shape.values = bad_shape.values[:, :-1]
    bad_dtype = _req(1)
    bad_dtype.dtype = "f32"
    bad_coords = _req(2)
    bad_coords.lat = bad_coords.lat.astype(np.float32)
    good = _req(3)
    _run(srv, [bad_shape, bad_dtype, bad_coords, good])
    for r in (bad_shape, bad_dtype, bad_coords):
        assert isinstance(r.error, capi.HcgError) and r.error.code == capi.HCG_ERR_INVALID
        assert r.footprint == 0
    assert good.error is None and len(srv.finished) == 1
    assert be.launches == [1]


@pytest.mark.parametrize("stage", ["prepare", "build"])
def test_failure_while_setting_up_isolated(stage):
    kw = {"fail_prepare": {2}} if stage == "prepare" else {"fail_build": {2}}
    be = FakeBackend(**kw)
    srv = _server(be, budget_bytes=10 ** 6, slots=1, max_round=8)
    reqs = [_req(i) for i in range(5)]
    _run(srv, reqs)
    assert reqs[2].error is not None and "injected" in str(reqs[2].error)
    assert all(r.error is None for i, r in enumerate(reqs) if i != 2)
    assert len(srv.finished) == 4 and len(srv.failed) == 1
    assert srv.used == sum(e.bytes for e in srv.geoms.values())
    assert all(e.refs == 0 for e in srv.geoms.values())


def test_failure_of_the_round_launch_falls_back_to_one_launch_per_job():
    # the stand-in refuses any launch that holds request 3's job
    be = FakeBackend()
    be.fail_launch = {3}
    srv = _server(be, budget_bytes=10 ** 6, slots=1, max_round=8)
    reqs = [_req(i) for i in range(5)]
    _run(srv, reqs)
    assert reqs[3].error is not None
    assert all(r.error is None for i, r in enumerate(reqs) if i != 3)
    assert be.launches == [1, 1, 1, 1]  # the others, one launch each
    assert all(e.refs == 0 for e in srv.geoms.values())


def test_builder_failure_does_not_strand_same_key_requests():
    # two requests on the same new coordinates in one round; the first one's
    # set-up fails, so the second builds the geometry itself
    be = FakeBackend(fail_prepare={0})
    srv = _server(be, budget_bytes=10 ** 6, slots=1, max_round=8)
    a = _req(0)
    b = _req(1, coords=(a.lon.copy(), a.lat.copy()))
    _run(srv, [a, b])
    assert a.error is not None and b.error is None
    assert srv.stats["geometry_builds"] == 1
    assert srv.used == sum(e.bytes for e in srv.geoms.values())


def test_continuous_batching_overlaps_rounds():
    # rounds stay in flight for several polls; requests arriving meanwhile
    # form further rounds on the other slot
    be = FakeBackend(polls=5)
    srv = _server(be, budget_bytes=10 ** 6, slots=2, max_round=2)
    reqs = [_req(i) for i in range(8)]
    _run(srv, reqs)
    assert len(srv.finished) == 8
    assert srv.stats["rounds"] >= 4 and srv.stats["round_jobs_max"] <= 2


def test_inputs_released_on_completion():
    # served, failed at submit and failed at set-up: the server keeps no
    # reference to any request's inputs once it has completed
    import gc
    import weakref
    be = FakeBackend(fail_prepare={1})
    srv = _server(be, budget_bytes=10 ** 6, slots=1)
    reqs = [_req(i, n=50 + i, nchan=2) for i in range(4)]
    reqs[2].dtype = "f32"  # invalid: fails at submit
    refs = [weakref.ref(r.values) for r in reqs]
    shapes = [(r.n, r.nchan) for r in reqs]
    _run(srv, reqs)
    be.rid_of.clear()  # the stand-in's own bookkeeping
    gc.collect()
    assert [r.rid for r in srv.finished] == [0, 3] and len(srv.failed) == 2
    for r, ref, shp in zip(reqs, refs, shapes):
        assert r.lon is None and r.lat is None and r.values is None
        assert ref() is None, f"request {r.rid}: its values are still referenced"
        assert (r.n, r.nchan) == shp and r.updates == shp[0] * shp[1]


def test_abort_ends_the_loop():
    # a producer that fails calls abort(): the loop raises instead of
    # waiting for requests that will never arrive
    import threading
    be = FakeBackend()
    srv = _server(be, budget_bytes=10 ** 6, slots=1)
    srv.submit(_req(0))
    threading.Timer(0.2, srv.abort, args=(RuntimeError("producer failed"),)).start()
    with pytest.raises(RuntimeError, match="aborted after 1 of 5"):
        srv.serve(5, timeout_s=30)
