"""ctypes binding of the C ABI in include/hcgrid_amd.h.

This module only loads ``hcgrid_amd/libhcgrid_amd.so`` and declares the
prototypes.  It never falls back to anything: if the library is missing the
import of a function raises ``HcgError``.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhcgrid_amd.so")

HCG_OK = 0
HCG_ERR_INVALID = -1
HCG_ERR_HIP = -2
HCG_ERR_NOMEM = -3
HCG_ERR_UNSUPPORTED = -4
HCG_ERR_WORKSPACE = -5
HCG_ERR_IO = -6

MAX_Y = 1073741824
GAUSS1D, GAUSS2D, TAPERED_SINC = 1, 2, 3
SELECT_COMPAT, SELECT_DISC = 0, 1
SELECTIONS = {"compat": SELECT_COMPAT, "disc": SELECT_DISC}
# zenithal target projections (HCG_PROJ_*)
PROJECTIONS = {"SIN": 0, "TAN": 1, "ARC": 2, "ZEA": 3, "STG": 4}
DTYPE_F64, DTYPE_F32, DTYPE_BF16, DTYPE_F16 = 0, 1, 2, 3


class HcgError(RuntimeError):
    def __init__(self, code: int, where: str, detail: str = ""):
        self.code = code
        super().__init__(f"{where}: error {code}{(' - ' + detail) if detail else ''}")


class Healpix(C.Structure):
    _fields_ = [
        ("nside", C.c_uint64), ("order", C.c_uint64), ("nrings", C.c_uint64),
        ("npface", C.c_uint64), ("ncap", C.c_uint64), ("npix", C.c_uint64),
        ("fact1", C.c_double), ("fact2", C.c_double), ("omega", C.c_double),
        ("resolution", C.c_double),
    ]


class Kernel(C.Structure):
    _fields_ = [
        ("kernel_type", C.c_uint32), ("bearing_needed", C.c_uint32),
        ("factor", C.c_uint32), ("selection", C.c_uint32),
        ("params", C.c_double * 3), ("sphere_radius", C.c_double),
        ("disc_size", C.c_double),
    ]


class MfmaShape(C.Structure):
    _fields_ = [("tile_cols", C.c_uint32), ("waves", C.c_uint32), ("chunk", C.c_uint32),
                ("reserved", C.c_uint32)]


class RingMeta(C.Structure):
    _fields_ = [("firstring", C.c_uint32), ("usedrings", C.c_uint32),
                ("n_valid", C.c_uint32), ("n", C.c_uint32)]


class GeometryView(C.Structure):
    _fields_ = [("n", C.c_uint32), ("side", C.c_uint32), ("bytes", C.c_size_t),
                ("keys_sorted", C.c_void_p), ("perm", C.c_void_p), ("records", C.c_void_p),
                ("lon_sorted", C.c_void_p), ("lat_sorted", C.c_void_p), ("lut", C.c_void_p),
                ("ring_z", C.c_void_p), ("ring_s", C.c_void_p), ("meta", C.c_void_p),
                ("xwcs", C.c_void_p), ("ywcs", C.c_void_p)]


class BatchJob(C.Structure):
    _fields_ = [("geometry", C.c_void_p), ("nchan", C.c_uint32), ("dtype", C.c_int),
                ("values", C.c_void_p), ("cube", C.c_void_p), ("wsum", C.c_void_p),
                ("tile_cols", C.c_uint32), ("normalize", C.c_int)]


class BatchStats(C.Structure):
    _fields_ = [("njobs", C.c_uint32), ("nitems", C.c_uint32), ("blocks", C.c_uint32),
                ("waves", C.c_uint32), ("chunk", C.c_uint32), ("reserved", C.c_uint32),
                ("device_bytes", C.c_size_t)]


SAMPLE_RECORD_BYTES = 32  # hcg_sample: x, y, z, v doubles

_vp = C.c_void_p
_u32 = C.c_uint32
_u64 = C.c_uint64
_int = C.c_int
_dbl = C.c_double
_sz = C.c_size_t

# name -> (restype, argtypes); every int-returning function is error-checked.
PROTOTYPES = {
    "hcg_strerror": (C.c_char_p, [_int]),
    "hcg_last_error_message": (C.c_char_p, []),
    "hcg_device_count": (_int, [C.POINTER(_int)]),
    "hcg_set_device": (_int, [_int]),
    "hcg_malloc": (_int, [C.POINTER(_vp), _sz]),
    "hcg_free": (_int, [_vp]),
    "hcg_memcpy_h2d": (_int, [_vp, _vp, _sz, _vp]),
    "hcg_memcpy_d2h": (_int, [_vp, _vp, _sz, _vp]),
    "hcg_memset": (_int, [_vp, _int, _sz, _vp]),
    "hcg_stream_create": (_int, [C.POINTER(_vp)]),
    "hcg_stream_destroy": (_int, [_vp]),
    "hcg_stream_synchronize": (_int, [_vp]),
    "hcg_healpix_init": (_int, [_u64, C.POINTER(Healpix)]),
    "hcg_set_optimal_nside": (_u64, [_dbl]),
    "hcg_prepare_grid_kernel": (_int, [_u32, C.POINTER(_dbl), _dbl, _dbl, C.POINTER(Kernel),
                                       C.POINTER(Healpix)]),
    "hcg_host_ang2pix": (_u64, [C.POINTER(Healpix), _dbl, _dbl]),
    "hcg_host_pix2ring": (_u64, [C.POINTER(Healpix), _u64]),
    "hcg_host_ring_info": (None, [C.POINTER(Healpix), _u64, C.POINTER(_u64), C.POINTER(_u64)]),
    "hcg_host_pix2theta": (_dbl, [C.POINTER(Healpix), _u64]),
    "hcg_presort_workspace_size": (_sz, [_u32]),
    "hcg_presort": (_int, [C.POINTER(Healpix), _u32, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "hcg_presort_noalloc": (_int, [C.POINTER(Healpix), _u32, _vp, _vp, _vp, _vp, _vp]),
    "hcg_gather_samples": (_int, [_u32, _vp, _vp, _vp, _vp, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "hcg_ring_lut": (_int, [C.POINTER(Healpix), _u32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "hcg_grid_plain": (_int, [C.POINTER(Healpix), C.POINTER(Kernel), _u32, _u32, _vp, _vp, _vp, _vp, _vp,
                              _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "hcg_channels_padded_len": (_u32, [_u32]),
    "hcg_gather_channels": (_int, [_u32, _u32, _vp, _vp, _int, _int, _vp, _vp, _vp]),
    "hcg_grid_mfma": (_int, [C.POINTER(Healpix), C.POINTER(Kernel), _u32, _u32, _u32, _int, _vp, _vp, _vp, _vp,
                             _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _int, _vp]),
    "hcg_grid_mfma_workspace_size": (_sz, [_u32, C.POINTER(MfmaShape)]),
    "hcg_grid_mfma_ex": (_int, [C.POINTER(Healpix), C.POINTER(Kernel), _u32, _u32, _u32, _int, _vp, _vp, _vp,
                                _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _int,
                                C.POINTER(MfmaShape), _vp, _sz, _vp]),
    "hcg_grid_mfma_shaped": (_int, [C.POINTER(Healpix), C.POINTER(Kernel), _u32, _u32, _u32, _int, _vp, _vp, _vp,
                                    _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _int,
                                    C.POINTER(MfmaShape), _vp]),
    "hcg_normalize": (_int, [_u64, _vp, _vp, _vp, _vp]),
    "hcg_eval_exp_neg": (_int, [_u32, _vp, _vp, _vp]),
    "hcg_wcs_p2s": (_int, [_u32, _u32, _u32, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, _vp, _vp]),
    "hcg_wcs_p2s_device": (_int, [_u32, _u32, _u32, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, _vp, _vp, _vp]),
    "hcg_wcs_sin_p2s": (_int, [_u32, _u32, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, _vp, _vp]),
    "hcg_wcs_sin_p2s_device": (_int, [_u32, _u32, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, _vp, _vp, _vp]),
    "hcg_geometry_bytes": (_sz, [_u32, _u32, C.POINTER(Healpix)]),
    "hcg_geometry_create": (_int, [C.POINTER(Healpix), C.POINTER(Kernel), _u32, _u32, C.POINTER(_vp)]),
    "hcg_geometry_create_in": (_int, [C.POINTER(Healpix), C.POINTER(Kernel), _u32, _u32, _vp, _sz,
                                      C.POINTER(_vp)]),
    "hcg_geometry_build": (_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "hcg_geometry_arrays": (_int, [_vp, C.POINTER(GeometryView)]),
    "hcg_geometry_destroy": (_int, [_vp]),
    "hcg_batch_job_bytes": (_sz, [_vp, _u32]),
    "hcg_batch_create": (_int, [_u32, C.POINTER(BatchJob), _u32, _u32, C.POINTER(_vp)]),
    "hcg_batch_bytes": (_sz, [_u32, C.POINTER(BatchJob), _u32, _u32]),
    "hcg_batch_create_in": (_int, [_u32, C.POINTER(BatchJob), _u32, _u32, _vp, _sz, _vp, C.POINTER(_vp)]),
    "hcg_batch_enqueue": (_int, [_vp, _vp]),
    "hcg_batch_info": (_int, [_vp, C.POINTER(BatchStats)]),
    "hcg_batch_nonfinite": (_int, [_vp, C.POINTER(_u32), _vp]),
    "hcg_batch_timing": (_int, [_vp, _int]),
    "hcg_batch_last_times": (_int, [_vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "hcg_batch_destroy": (_int, [_vp]),
    "hcg_content_hash": (_int, [_vp, _sz, _vp, _vp]),
    "hcg_graph_capture_begin": (_int, [_vp]),
    "hcg_graph_capture_end": (_int, [_vp, C.POINTER(_vp)]),
    "hcg_graph_launch": (_int, [_vp, _vp]),
    "hcg_graph_destroy": (_int, [_vp]),
}

_lib = None


def header_symbols(header_path: str | None = None) -> list[str]:
    """Function names declared in include/hcgrid_amd.h."""
    import re
    if header_path is None:
        header_path = os.path.join(os.path.dirname(_HERE), "include", "hcgrid_amd.h")
    text = open(header_path).read()
    return sorted(set(re.findall(r"\b(hcg_[a-z0-9_]+)\s*\(", text)))


def load(path: str = LIB_PATH):
    """Load the gfx950 library and declare prototypes (no GPU call is made)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise HcgError(HCG_ERR_INVALID, "load", f"{path} is not built (run __graft_entry__.build())")
    # torch bundles its own libamdhip64.so.7.  The library must bind to the
    # same HIP runtime instance that owns torch's allocations and streams, so
    # torch is loaded first and the dynamic linker reuses its runtime (same
    # SONAME); loading ours first would start a second runtime in the process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, where: str) -> None:
    if rc != HCG_OK:
        lib = load()
        raise HcgError(rc, where, lib.hcg_last_error_message().decode(errors="replace"))


def call(name: str, *args) -> int:
    """Call an int-returning ABI function and raise HcgError on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    check(rc, name)
    return rc
