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
from .gridder import _ptr, _stream_handle, _torch

DTYPES = {"bf16": capi.DTYPE_BF16, "f16": capi.DTYPE_F16}


class Geometry:
    """hcg_geometry: allocated at construction, filled by build()."""

    def __init__(self, kern, hp, n: int, side: int, in_torch: bool = False):
        """in_torch: place the geometry in a torch allocation (the caching
        allocator), so creating and dropping it makes no hipMalloc/hipFree."""
        capi.load()
        self.kern, self.hp = kern, hp
        self.n, self.side = int(n), int(side)
        self._h = C.c_void_p()
        self._mem = None
        if in_torch:
            torch = _torch()
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

    def set_target_sin(self, target, stream=None):
        """Fills the geometry's target coordinates on the device: a SIN map
        of target.side^2 cells (hcg_wcs_sin_p2s_device)."""
        t = target
        call("hcg_wcs_sin_p2s_device", self.side, self.side, t.crpix, t.crpix, -t.cdelt, t.cdelt, t.crval1,
             t.crval2, 180.0, C.c_void_p(self.view.xwcs), C.c_void_p(self.view.ywcs), _stream_handle(stream))

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
        arr = (BatchJob * max(1, len(self.jobs)))()
        for i, j in enumerate(self.jobs):
            g = j["geometry"]
            vals = j["values"]
            nchan = int(vals.shape[0]) if vals is not None else 0
            if vals is not None and int(vals.shape[1]) != g.n:
                raise capi.HcgError(capi.HCG_ERR_INVALID, "Batch", f"job {i}: values have {vals.shape[1]} samples, "
                                    f"geometry {g.n}")
            cube = j["cube"]
            if nchan and tuple(cube.shape) != (nchan, g.side * g.side):
                raise capi.HcgError(capi.HCG_ERR_INVALID, "Batch", f"job {i}: cube shape {tuple(cube.shape)}")
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
