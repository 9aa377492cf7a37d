"""Resident geometries, batched multi-channel gridding and graph replay
(Python over the C ABI's hcg_geometry_*, hcg_batch_* and hcg_graph_*).

A Geometry is one map's sorted samples, ring LUT and target coordinates in
HBM; building it is the pre-sort of the reference's init_input_with_cpu
(HCGrid/gridding.cu:70-147) done on the device.  A Batch grids many
(geometry, channel values, cube) jobs of uneven sizes in one persistent
launch.  A Graph captures any sequence of these on a stream and replays it.
torch only allocates device memory and provides streams.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import BatchJob, BatchStats, GeometryView, call
from .gridder import _ptr, _stream_handle, _torch, check_device_of, check_tensor

DTYPES = {"bf16": capi.DTYPE_BF16, "f16": capi.DTYPE_F16}


def validate_job(i: int, j: dict) -> int:
    """Checks one Batch job dict before any pointer of it reaches the C ABI;
    returns its channel count.  values: contiguous [nchan][n] tensor of the
    declared dtype; cube: contiguous float32 [nchan][side*side]; wsum: None
    or contiguous float32 [side*side]; all on the geometry's device (any GPU
    when the geometry does not say)."""
    where = f"Batch job {i}"
    g = j["geometry"]
    dtype = j.get("dtype", "bf16")
    if dtype not in DTYPES:
        raise capi.HcgError(capi.HCG_ERR_INVALID, where, f"dtype {dtype!r} (expected 'bf16' or 'f16')")
    vals = j.get("values")
    device = getattr(g, "device", None)
    m = int(g.side) * int(g.side)
    if vals is None:
        return 0
    if vals.dim() != 2 or int(vals.shape[1]) != int(g.n):
        raise capi.HcgError(capi.HCG_ERR_INVALID, where,
                            f"values have shape {tuple(vals.shape)}, expected (nchan, {g.n})")
    nchan = int(vals.shape[0])
    check_tensor(vals, where, "values", dtype, nchan * int(g.n), check_device=False)
    if nchan == 0:
        return 0
    cube = j.get("cube")
    if cube is None or tuple(cube.shape) not in ((nchan, m), (nchan * m,)):
        raise capi.HcgError(capi.HCG_ERR_INVALID, where,
                            f"cube shape {None if cube is None else tuple(cube.shape)}, expected ({nchan}, {m})")
    check_tensor(cube, where, "cube", "f32", nchan * m, check_device=False)
    w = j.get("wsum")
    if w is not None:
        check_tensor(w, where, "wsum", "f32", m, check_device=False)
    for name, t in (("values", vals), ("cube", cube), ("wsum", w)):
        if t is not None:
            check_device_of(t, where, name, device)
    return nchan


class Geometry:
    """hcg_geometry: allocated at construction, filled by build()."""

    def __init__(self, kern, hp, n: int, side: int, in_torch: bool = False):
        """in_torch: place the geometry in a torch allocation (the caching
        allocator), so creating and dropping it makes no hipMalloc/hipFree."""
        capi.load()
        torch = _torch()
        self.kern, self.hp = kern, hp
        self.n, self.side = int(n), int(side)
        self.device = torch.device("cuda", torch.cuda.current_device())  # where hcg_geometry_* allocates
        self._h = C.c_void_p()
        self._mem = None
        if in_torch:
            nb = Geometry.bytes_for(self.n, self.side, hp)
            self._mem = torch.empty(nb + 256, dtype=torch.uint8, device="cuda")
            base = (self._mem.data_ptr() + 255) // 256 * 256
            call("hcg_geometry_create_in", C.byref(hp), C.byref(kern), self.n, self.side, C.c_void_p(base), nb,
                 C.byref(self._h))
        else:
            call("hcg_geometry_create", C.byref(hp), C.byref(kern), self.n, self.side, C.byref(self._h))
        v = GeometryView()
        call("hcg_geometry_arrays", self._h, C.byref(v))
        self.view = v

    @staticmethod
    def bytes_for(n: int, side: int, hp) -> int:
        return int(capi.load().hcg_geometry_bytes(int(n), int(side), C.byref(hp)))

    @property
    def handle(self) -> C.c_void_p:
        if not self._h:
            raise capi.HcgError(capi.HCG_ERR_INVALID, "Geometry", "destroyed")
        return self._h

    @property
    def nbytes(self) -> int:
        return int(self.view.bytes)

    def build(self, lon, lat, xwcs=None, ywcs=None, stream=None):
        """lon/lat: device float64 [n] degrees; xwcs/ywcs: device float64
        [side*side], or None when the geometry's own target arrays
        (view.xwcs/ywcs) have been filled already."""
        xp = C.c_void_p(self.view.xwcs) if xwcs is None else _ptr(xwcs)
        yp = C.c_void_p(self.view.ywcs) if ywcs is None else _ptr(ywcs)
        call("hcg_geometry_build", self.handle, _ptr(lon), _ptr(lat), xp, yp, _stream_handle(stream))

    def set_target(self, target, stream=None):
        """Fills the geometry's target coordinates on the device: a zenithal
        map (target.proj, default SIN) of target.side^2 cells, CDELT1 =
        -cdelt, CDELT2 = cdelt, LONPOLE 0 at CRVAL2 >= 90 and 180 otherwise
        (hcg_wcs_p2s_device)."""
        t = target
        if t.side != self.side:
            raise capi.HcgError(capi.HCG_ERR_INVALID, "Geometry.set_target", f"side {t.side} != {self.side}")
        proj = getattr(t, "proj", "SIN")
        lonpole = 0.0 if t.crval2 >= 90.0 else 180.0
        call("hcg_wcs_p2s_device", capi.PROJECTIONS[proj], self.side, self.side, t.crpix, t.crpix, -t.cdelt,
             t.cdelt, t.crval1, t.crval2, lonpole, C.c_void_p(self.view.xwcs), C.c_void_p(self.view.ywcs),
             _stream_handle(stream))

    def set_target_sin(self, target, stream=None):
        """set_target for a SIN map (the configured workloads' projection)."""
        self.set_target(target, stream)

    def _read(self, ptr, count, dtype):
        torch = _torch()
        out = np.empty(count, dtype=dtype)
        if count:
            torch.cuda.synchronize()
            call("hcg_memcpy_d2h", C.c_void_p(out.ctypes.data), C.c_void_p(ptr), out.nbytes, None)
            torch.cuda.synchronize()
        return out

    def perm(self) -> np.ndarray:
        return self._read(self.view.perm, self.n, np.uint32)

    def keys_sorted(self) -> np.ndarray:
        return self._read(self.view.keys_sorted, self.n + 1, np.uint32)

    def meta(self) -> dict:
        m = self._read(self.view.meta, 4, np.uint32)
        return dict(firstring=int(m[0]), usedrings=int(m[1]), n_valid=int(m[2]), n=int(m[3]))

    def close(self):
        if self._h:
            call("hcg_geometry_destroy", self._h)
            self._h = C.c_void_p()
        self._mem = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """hcg_batch over jobs = [dict(geometry, values, cube, wsum=None, dtype='bf16',
    tile_cols=0, normalize=True)].  values: device tensor [nchan][n] (bf16 or
    f16), cube: device float32 [nchan][side*side], wsum: float32 [side*side]
    or None.  The batch keeps references to everything it reads or writes."""

    def __init__(self, jobs, waves: int = 0, chunk: int = 0, stream=None):
        """stream: plan the batch in a torch allocation made on `stream` and
        upload the plan asynchronously there (no hipMalloc/hipFree, no
        synchronisation); the batch must then be enqueued on that stream."""
        capi.load()
        self.jobs = list(jobs)
        nchans = [validate_job(i, j) for i, j in enumerate(self.jobs)]
        arr = (BatchJob * max(1, len(self.jobs)))()
        for i, j in enumerate(self.jobs):
            g = j["geometry"]
            vals = j["values"]
            nchan = nchans[i]
            cube = j["cube"]
            arr[i].geometry = g.handle.value
            arr[i].nchan = nchan
            arr[i].dtype = DTYPES[j.get("dtype", "bf16")]
            arr[i].values = vals.data_ptr() if nchan else None
            arr[i].cube = cube.data_ptr() if nchan else None
            w = j.get("wsum")
            arr[i].wsum = w.data_ptr() if w is not None else None
            arr[i].tile_cols = int(j.get("tile_cols", 0))
            arr[i].normalize = 1 if j.get("normalize", True) else 0
        self._h = C.c_void_p()
        self._mem = None
        if stream is None:
            call("hcg_batch_create", len(self.jobs), arr, int(waves), int(chunk), C.byref(self._h))
        else:
            torch = _torch()
            nb = int(capi.load().hcg_batch_bytes(len(self.jobs), arr, int(waves), int(chunk)))
            if nb == 0:
                raise capi.HcgError(capi.HCG_ERR_INVALID, "Batch", "invalid jobs")
            with torch.cuda.stream(stream):
                self._mem = torch.empty(nb + 256, dtype=torch.uint8, device="cuda")
            base = (self._mem.data_ptr() + 255) // 256 * 256
            call("hcg_batch_create_in", len(self.jobs), arr, int(waves), int(chunk), C.c_void_p(base), nb,
                 _stream_handle(stream), C.byref(self._h))

    @property
    def handle(self) -> C.c_void_p:
        if not self._h:
            raise capi.HcgError(capi.HCG_ERR_INVALID, "Batch", "destroyed")
        return self._h

    def enqueue(self, stream=None):
        call("hcg_batch_enqueue", self.handle, _stream_handle(stream))

    def timing(self, enable: bool = True):
        """Records HIP events around the prep and gridding launches of every
        later eager enqueue (hcg_batch_timing)."""
        call("hcg_batch_timing", self.handle, 1 if enable else 0)

    def last_times(self) -> tuple[float, float]:
        """(prep_ms, grid_ms) of the last timed enqueue; waits for it."""
        p, g = C.c_float(), C.c_float()
        call("hcg_batch_last_times", self.handle, C.byref(p), C.byref(g))
        return float(p.value), float(g.value)

    def stats(self) -> dict:
        st = BatchStats()
        call("hcg_batch_info", self.handle, C.byref(st))
        return dict(njobs=st.njobs, nitems=st.nitems, blocks=st.blocks, waves=st.waves, chunk=st.chunk,
                    device_bytes=int(st.device_bytes))

    def nonfinite(self, stream=None) -> list[int]:
        n = len(self.jobs)
        out = (C.c_uint32 * max(1, n))()
        call("hcg_batch_nonfinite", self.handle, out, _stream_handle(stream))
        return [int(out[i]) for i in range(n)]

    def close(self):
        if self._h:
            call("hcg_batch_destroy", self._h)
            self._h = C.c_void_p()
        self._mem = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """A captured sequence of library calls, replayed with launch().  The
    stream must be a created stream (torch.cuda.Stream()), not the null one."""

    def __init__(self, fn, stream):
        capi.load()
        self.stream = stream
        self._exec = C.c_void_p()
        sh = _stream_handle(stream)
        call("hcg_graph_capture_begin", sh)
        try:
            fn(stream)
        finally:
            call("hcg_graph_capture_end", sh, C.byref(self._exec))

    def launch(self, stream=None):
        call("hcg_graph_launch", self._exec, _stream_handle(stream if stream is not None else self.stream))

    def close(self):
        if self._exec:
            call("hcg_graph_destroy", self._exec)
            self._exec = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
