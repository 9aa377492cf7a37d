"""Host-side driver of the device gridding path (Python over the C ABI).

Mirrors the reference's host sequence of solve_gridding
(HCGrid/gridding.cu:484-556): kernel set-up (_prepare_grid_kernel), the
pre-sort and ring LUT (init_input_with_cpu), zeroed output cubes
(init_output), the gridding kernel and the output normalisation
(write_output_map) -- with every array resident in HBM.  torch is only the
device allocator and stream provider; all compute is in libhcgrid_amd.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import capi
from .capi import GAUSS1D, Healpix, Kernel, RingMeta, call


def _torch():
    import torch
    return torch


def prepare_grid_kernel(kernel_type: int, kernel_params, sphere_radius: float, hpx_max_res: float,
                        factor: int = 1, selection: str = "compat"):
    """_prepare_grid_kernel (HCGrid/gmap.cpp:64-99) -> (Kernel, Healpix).
    selection: "compat" (the reference's windows) or "disc" (every sample
    within sphere_radius; HCG_SELECT_DISC)."""
    if selection not in capi.SELECTIONS:
        raise capi.HcgError(capi.HCG_ERR_INVALID, "prepare_grid_kernel", f"selection {selection!r}")
    capi.load()
    params = (C.c_double * 3)(*([float(x) for x in kernel_params] + [0.0] * (3 - len(kernel_params))))
    kern = Kernel()
    hp = Healpix()
    call("hcg_prepare_grid_kernel", kernel_type, params, float(sphere_radius), float(hpx_max_res),
         C.byref(kern), C.byref(hp))
    kern.factor = int(factor)
    kern.selection = capi.SELECTIONS[selection]
    return kern, hp


def wcs(proj: str, side_x: int, side_y: int, crpix1: float, crpix2: float, cdelt1: float, cdelt2: float,
        crval1: float, crval2: float, lonpole: float | None = None):
    """Target cell coordinates of a zenithal projection ("SIN", "TAN", "ARC",
    "ZEA", "STG"; set_WCS, HCGrid/gmap.cpp:183-250), host arrays.  lonpole
    defaults to 0 when crval2 >= 90 and to 180 otherwise (Calabretta &
    Greisen 2002, Paper II sec. 2.4)."""
    if proj not in capi.PROJECTIONS:
        raise ValueError(f"projection must be one of {sorted(capi.PROJECTIONS)}")
    if lonpole is None:
        lonpole = 0.0 if crval2 >= 90.0 else 180.0
    capi.load()
    xw = np.empty(side_x * side_y)
    yw = np.empty(side_x * side_y)
    call("hcg_wcs_p2s", capi.PROJECTIONS[proj], side_x, side_y, crpix1, crpix2, cdelt1, cdelt2, crval1, crval2,
         lonpole, xw.ctypes.data_as(C.c_void_p), yw.ctypes.data_as(C.c_void_p))
    return xw, yw


def wcs_sin(side_x: int, side_y: int, crpix1: float, crpix2: float, cdelt1: float, cdelt2: float,
            crval1: float, crval2: float, lonpole: float = 180.0):
    """SIN target cell coordinates (set_WCS, HCGrid/gmap.cpp:183-250), host arrays."""
    return wcs("SIN", side_x, side_y, crpix1, crpix2, cdelt1, cdelt2, crval1, crval2, lonpole)


def _ptr(t) -> C.c_void_p | None:
    return None if t is None else C.c_void_p(t.data_ptr())


_DTYPE_NAMES = {"f64": "float64", "f32": "float32", "bf16": "bfloat16", "f16": "float16", "i32": "int32"}


def check_tensor(t, where: str, name: str, dtypes, numel: int | None = None, device=None, min_numel=None,
                 check_device: bool = True):
    """Raises HcgError unless `t` is a contiguous torch tensor of one of
    `dtypes` (names: "f64", "f32", "bf16", "f16", "i32") with `numel`
    elements (or at least `min_numel`) on `device` (a torch.device; None =
    any CUDA device).  Called before any pointer goes to the C ABI, which
    sees only addresses and would read or write out of bounds otherwise.
    Dtype, size and layout are checked before the device, so the checks
    can be exercised with CPU tensors."""
    torch = _torch()
    if not isinstance(t, torch.Tensor):
        raise capi.HcgError(capi.HCG_ERR_INVALID, where, f"{name} must be a torch tensor, got {type(t).__name__}")
    if isinstance(dtypes, str):
        dtypes = (dtypes,)
    allowed = [getattr(torch, _DTYPE_NAMES[d]) for d in dtypes]
    if t.dtype not in allowed:
        raise capi.HcgError(capi.HCG_ERR_INVALID, where,
                            f"{name} has dtype {t.dtype}, expected {' or '.join(str(a) for a in allowed)}")
    if numel is not None and t.numel() != numel:
        raise capi.HcgError(capi.HCG_ERR_INVALID, where, f"{name} has {t.numel()} elements, expected {numel}")
    if min_numel is not None and t.numel() < min_numel:
        raise capi.HcgError(capi.HCG_ERR_INVALID, where,
                            f"{name} has {t.numel()} elements, expected at least {min_numel}")
    if not t.is_contiguous():
        raise capi.HcgError(capi.HCG_ERR_INVALID, where, f"{name} is not contiguous")
    if check_device:
        check_device_of(t, where, name, device)


def check_device_of(t, where: str, name: str, device=None):
    """Raises HcgError unless tensor `t` is on `device` (None = any GPU)."""
    if t.device.type != "cuda" or (device is not None and t.device != device):
        want = device if device is not None else "a GPU"
        raise capi.HcgError(capi.HCG_ERR_INVALID, where, f"{name} is on {t.device}, expected {want}")


def _stream_handle(stream) -> C.c_void_p:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


@dataclass
class PlainBuffers:
    keys: object        # int32 view of uint32 [n+1]
    perm: object        # int32 view of uint32 [n]
    workspace: object   # uint8
    records: object     # float64 [n*4]
    lon_s: object
    lat_s: object
    w_s: object
    lut: object         # int32 view of uint32 [nrings+2]
    ring_z: object
    ring_s: object
    meta: object        # int32 [4]
    datacube: object    # float64 [side*side]
    weightscube: object
    map: object


class PlainGridder:
    """Single-channel drop-in path with resident buffers for (n, side, nside)."""

    def __init__(self, n: int, side: int, kern: Kernel, hp: Healpix, device: int = 0,
                 with_weights: bool = False):
        torch = _torch()
        capi.load()
        if not torch.cuda.is_available():
            raise capi.HcgError(capi.HCG_ERR_HIP, "PlainGridder", "no GPU visible to torch")
        self.n, self.side, self.kern, self.hp = int(n), int(side), kern, hp
        dev = torch.device("cuda", device)
        self.device = dev
        i32, f64 = torch.int32, torch.float64
        ws_bytes = capi.load().hcg_presort_workspace_size(self.n)
        nr = int(hp.nrings) + 2
        m = self.side * self.side
        self.buf = PlainBuffers(
            keys=torch.empty(self.n + 1, dtype=i32, device=dev),
            perm=torch.empty(max(self.n, 1), dtype=i32, device=dev),
            workspace=torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev),
            records=torch.empty(max(self.n, 1) * 4, dtype=f64, device=dev),
            lon_s=torch.empty(max(self.n, 1), dtype=f64, device=dev),
            lat_s=torch.empty(max(self.n, 1), dtype=f64, device=dev),
            w_s=torch.empty(max(self.n, 1), dtype=f64, device=dev) if with_weights else None,
            lut=torch.empty(nr, dtype=i32, device=dev),
            ring_z=torch.empty(nr, dtype=f64, device=dev),
            ring_s=torch.empty(nr, dtype=f64, device=dev),
            meta=torch.zeros(4, dtype=i32, device=dev),
            datacube=torch.empty(max(m, 1), dtype=f64, device=dev),
            weightscube=torch.empty(max(m, 1), dtype=f64, device=dev),
            map=torch.empty(max(m, 1), dtype=f64, device=dev),
        )
        self.ws_bytes = ws_bytes

    # -- stages ------------------------------------------------------------
    def _check_coords(self, where, lon, lat):
        check_tensor(lon, where, "lon", "f64", self.n, self.device)
        check_tensor(lat, where, "lat", "f64", self.n, self.device)

    def presort(self, lon, lat, stream=None):
        b = self.buf
        self._check_coords("presort", lon, lat)
        call("hcg_presort", C.byref(self.hp), self.n, _ptr(lon), _ptr(lat), _ptr(b.keys), _ptr(b.perm),
             _ptr(b.workspace), self.ws_bytes, _stream_handle(stream))

    def gather(self, lon, lat, data, weights=None, stream=None):
        torch = _torch()
        b = self.buf
        dtype = capi.DTYPE_F64 if data.dtype == torch.float64 else capi.DTYPE_F32
        self._check_coords("gather", lon, lat)
        check_tensor(data, "gather", "data", ("f32", "f64"), self.n, self.device)
        if weights is not None and b.w_s is None:
            raise capi.HcgError(capi.HCG_ERR_INVALID, "gather", "gridder built without weights")
        if weights is not None:
            check_tensor(weights, "gather", "weights", "f64", self.n, self.device)
        call("hcg_gather_samples", self.n, _ptr(b.perm), _ptr(lon), _ptr(lat), _ptr(data), dtype,
             _ptr(weights), _ptr(b.records), _ptr(b.lon_s), _ptr(b.lat_s), None,
             _ptr(b.w_s) if weights is not None else None, _stream_handle(stream))

    def ring_lut(self, stream=None):
        b = self.buf
        call("hcg_ring_lut", C.byref(self.hp), self.n, _ptr(b.keys), _ptr(b.lut), _ptr(b.ring_z),
             _ptr(b.ring_s), _ptr(b.meta), _stream_handle(stream))

    def init_output(self, stream=None):
        b = self.buf
        m = self.side * self.side * 8
        sh = _stream_handle(stream)
        call("hcg_memset", _ptr(b.datacube), 0, m, sh)
        call("hcg_memset", _ptr(b.weightscube), 0, m, sh)

    def grid(self, xwcs, ywcs, weighted=False, stream=None):
        b = self.buf
        m = self.side * self.side
        check_tensor(xwcs, "grid", "xwcs", "f64", m, self.device)
        check_tensor(ywcs, "grid", "ywcs", "f64", m, self.device)
        call("hcg_grid_plain", C.byref(self.hp), C.byref(self.kern), self.side, self.n, _ptr(b.records),
             _ptr(b.keys), _ptr(b.lon_s), _ptr(b.lat_s), _ptr(b.w_s) if weighted else None, _ptr(b.lut),
             _ptr(b.ring_z), _ptr(b.ring_s), _ptr(b.meta), _ptr(xwcs), _ptr(ywcs), _ptr(b.datacube),
             _ptr(b.weightscube), _stream_handle(stream))

    def normalize(self, stream=None):
        b = self.buf
        call("hcg_normalize", self.side * self.side, _ptr(b.datacube), _ptr(b.weightscube), _ptr(b.map),
             _stream_handle(stream))

    def run(self, lon, lat, data, xwcs, ywcs, weights=None, stream=None, normalize=True):
        """One pass of the hot path; everything stays on the device."""
        self.presort(lon, lat, stream)
        self.gather(lon, lat, data, weights, stream)
        self.ring_lut(stream)
        self.init_output(stream)
        self.grid(xwcs, ywcs, weighted=weights is not None, stream=stream)
        if normalize:
            self.normalize(stream)

    # -- results -------------------------------------------------------------
    def meta(self) -> dict:
        m = self.buf.meta.cpu().numpy().view(np.uint32)
        return dict(firstring=int(m[0]), usedrings=int(m[1]), n_valid=int(m[2]), n=int(m[3]))

    def keys_sorted(self) -> np.ndarray:
        return self.buf.keys.cpu().numpy().view(np.uint32).copy()

    def perm(self) -> np.ndarray:
        return self.buf.perm[: self.n].cpu().numpy().view(np.uint32).copy()

    def lut(self) -> np.ndarray:
        used = self.meta()["usedrings"]
        return self.buf.lut[: used + 1].cpu().numpy().view(np.uint32).copy()


def grid_plain(lon, lat, data, xwcs, ywcs, side: int, sigma: float, radius: float, hpx_res: float,
               weights=None, factor: int = 1, kernel_type: int = GAUSS1D, kernel_params=None,
               device: int = 0, selection: str = "compat") -> dict:
    """Host arrays in, host arrays out: one full device pass (for tests and tools)."""
    torch = _torch()
    if kernel_params is None:
        kernel_params = [sigma]
    kern, hp = prepare_grid_kernel(kernel_type, kernel_params, radius, hpx_res, factor, selection)
    g = PlainGridder(len(lon), side, kern, hp, device=device, with_weights=weights is not None)
    dev = g.device
    t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)  # noqa: E731
    lon_d, lat_d = t(lon, torch.float64), t(lat, torch.float64)
    data_np = np.asarray(data)
    data_d = t(data_np, torch.float64 if data_np.dtype == np.float64 else torch.float32)
    xw_d, yw_d = t(xwcs, torch.float64), t(ywcs, torch.float64)
    w_d = None if weights is None else t(weights, torch.float64)
    g.run(lon_d, lat_d, data_d, xw_d, yw_d, weights=w_d)
    torch.cuda.synchronize()
    out = dict(meta=g.meta(), keys_sorted=g.keys_sorted(), perm=g.perm(), lut=g.lut(),
               datacube=g.buf.datacube.cpu().numpy().copy(), weightscube=g.buf.weightscube.cpu().numpy().copy(),
               map=g.buf.map.cpu().numpy().copy(), kern=kern, hp=hp)
    return out


class CubeGridder:
    """Multi-channel path: the HCGrid+ channel-by-channel job as one pass.

    Keys, sort, LUT and records as the single-channel path; channel values
    (bf16/f16, channel-major [nchan][n] or sample-major [n][nchan]) are
    gathered into sorted order (blocked by 8 samples) and contracted on MFMA.
    """

    def __init__(self, n: int, side: int, nchan: int, kern: Kernel, hp: Healpix, dtype: str = "bf16",
                 device: int = 0):
        torch = _torch()
        capi.load()
        if not torch.cuda.is_available():
            raise capi.HcgError(capi.HCG_ERR_HIP, "CubeGridder", "no GPU visible to torch")
        if dtype not in ("bf16", "f16"):
            raise capi.HcgError(capi.HCG_ERR_INVALID, "CubeGridder", "dtype must be bf16 or f16")
        self.plain = PlainGridder(n, side, kern, hp, device=device)
        self.n, self.side, self.nchan, self.kern, self.hp = int(n), int(side), int(nchan), kern, hp
        self.dtype = dtype
        self.dcode = capi.DTYPE_BF16 if dtype == "bf16" else capi.DTYPE_F16
        tdt = torch.bfloat16 if dtype == "bf16" else torch.float16
        dev = self.plain.device
        self.n_pad = int(capi.load().hcg_channels_padded_len(self.n))
        self.values_sorted = torch.empty((self.nchan, self.n_pad), dtype=tdt, device=dev)
        self.cube = torch.empty((self.nchan, self.side * self.side), dtype=torch.float32, device=dev)
        self.wsum = torch.empty(self.side * self.side, dtype=torch.float32, device=dev)
        self.nonfinite = torch.zeros(1, dtype=torch.int32, device=dev)
        self._ws = None

    def gather_channels(self, values, layout: int = 0, stream=None):
        if layout not in (0, 1):
            raise capi.HcgError(capi.HCG_ERR_INVALID, "gather_channels", "layout must be 0 or 1")
        check_tensor(values, "gather_channels", "values", self.dtype, self.nchan * self.n, self.plain.device)
        shape = (self.nchan, self.n) if layout == 0 else (self.n, self.nchan)
        if values.dim() == 2 and tuple(values.shape) != shape:
            raise capi.HcgError(capi.HCG_ERR_INVALID, "gather_channels",
                                f"values have shape {tuple(values.shape)}, expected {shape} for layout {layout}")
        call("hcg_gather_channels", self.n, self.nchan, _ptr(self.plain.buf.perm), _ptr(values), self.dcode, layout,
             _ptr(self.values_sorted), _ptr(self.nonfinite), _stream_handle(stream))

    def grid(self, xwcs, ywcs, normalize: bool = True, stream=None, shape=None, plan: bool = True):
        """shape: None (library defaults) or (tile_cols, waves, chunk), 0 = default.
        plan: precompute the tile plan in a workspace (hcg_grid_mfma_ex)."""
        torch = _torch()
        b = self.plain.buf
        m = self.side * self.side
        check_tensor(xwcs, "grid", "xwcs", "f64", m, self.plain.device)
        check_tensor(ywcs, "grid", "ywcs", "f64", m, self.plain.device)
        shp = capi.MfmaShape(*(shape if shape is not None else (0, 0, 0)), 0)
        ws, nb = None, 0
        if plan:
            nb = int(capi.load().hcg_grid_mfma_workspace_size(self.side, C.byref(shp)))
            if self._ws is None or self._ws.numel() < nb:
                self._ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=self.plain.device)
            ws = self._ws
        call("hcg_grid_mfma_ex", C.byref(self.hp), C.byref(self.kern), self.side, self.n, self.nchan,
             self.dcode, _ptr(b.records), _ptr(b.keys), _ptr(b.lon_s), _ptr(b.lat_s), None,
             _ptr(self.values_sorted), _ptr(b.lut), _ptr(b.ring_z), _ptr(b.ring_s), _ptr(b.meta), _ptr(xwcs),
             _ptr(ywcs), _ptr(self.cube), _ptr(self.wsum), 1 if normalize else 0, C.byref(shp), _ptr(ws), nb,
             _stream_handle(stream))

    def gather_records(self, lon, lat, stream=None):
        """Sorted sample records and coordinates (the value slot stays 0)."""
        b = self.plain.buf
        self.plain._check_coords("gather_records", lon, lat)
        call("hcg_gather_samples", self.n, _ptr(b.perm), _ptr(lon), _ptr(lat), None, capi.DTYPE_F64, None,
             _ptr(b.records), _ptr(b.lon_s), _ptr(b.lat_s), None, None, _stream_handle(stream))

    def run(self, lon, lat, values, xwcs, ywcs, layout: int = 0, stream=None, normalize: bool = True,
            shape=None):
        """One pass: keys + sort, records, LUT, channel gather, contraction."""
        self.plain.presort(lon, lat, stream)
        self.gather_records(lon, lat, stream)
        self.plain.ring_lut(stream)
        self.gather_channels(values, layout, stream)
        self.grid(xwcs, ywcs, normalize, stream, shape)


def grid_cube(lon, lat, values_cm, xwcs, ywcs, side: int, sigma: float, radius: float, hpx_res: float,
              dtype: str = "bf16", layout: int = 0, device: int = 0, shape=None,
              selection: str = "compat") -> dict:
    """Host arrays in/out: multi-channel pass (values_cm: torch tensor on the device).
    shape: None or (tile_cols, waves, chunk) for hcg_grid_mfma_shaped."""
    torch = _torch()
    kern, hp = prepare_grid_kernel(GAUSS1D, [sigma], radius, hpx_res, 1, selection)
    nchan = values_cm.shape[0] if layout == 0 else values_cm.shape[1]
    g = CubeGridder(len(lon), side, nchan, kern, hp, dtype=dtype, device=device)
    dev = g.plain.device
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)  # noqa: E731
    g.run(t(lon), t(lat), values_cm.contiguous(), t(xwcs), t(ywcs), layout=layout, shape=shape)
    torch.cuda.synchronize()
    return dict(cube=g.cube.float().cpu().numpy().copy(), wsum=g.wsum.cpu().numpy().copy(),
                perm=g.plain.perm(), nonfinite=int(g.nonfinite.item()), meta=g.plain.meta())
