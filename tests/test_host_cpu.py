"""The C++ drop-in host layer on the CPU: FITS input/target/output and the
target geometry, checked against cfitsio (an independent reader/writer that
ships in this image) and the oracle's SIN deprojection.  Nothing here runs
on the GPU."""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
from hcgrid_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_LIB = os.path.join(ROOT, "hcgrid_amd", "libhcgrid_host.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")
CFITSIO = "/opt/conda/lib/libcfitsio.so"


def _host():
    capi.load()  # one HIP runtime per process (see capi.load)
    lib = C.CDLL(HOST_LIB)
    out = subprocess.run(["nm", "-D", "--defined-only", HOST_LIB], capture_output=True, text=True, check=True).stdout
    dem = subprocess.run(["c++filt"], input=out, capture_output=True, text=True, check=True).stdout
    syms = {}
    for raw, pretty in zip(out.splitlines(), dem.splitlines()):
        m = re.match(r"\S+ T (\S+)", raw)
        p = re.match(r"\S+ T (.+)$", pretty)
        if m and p:
            syms[p.group(1)] = m.group(1)
    lib.hcg_host_set_exit_on_error(0)
    lib.hcg_host_error.restype = C.c_char_p
    return lib, syms


def _fn(lib, syms, pretty, argtypes, restype=None):
    f = getattr(lib, syms[pretty])
    f.argtypes = argtypes
    f.restype = restype
    return f


class _Fits:
    """Minimal cfitsio access (the checker)."""

    def __init__(self):
        if not os.path.exists(CFITSIO):
            pytest.skip("cfitsio not present")
        self.L = C.CDLL(CFITSIO)

    def read(self, path, hdu, n):
        f, st = C.c_void_p(), C.c_int(0)
        self.L.ffopen(C.byref(f), path.encode(), 0, C.byref(st))
        assert st.value == 0
        self.L.ffmahd(f, hdu, None, C.byref(st))
        out = np.empty(n)
        anynul = C.c_int(0)
        nul = C.c_double(0)
        self.L.ffgpvd(f, 1, C.c_longlong(1), C.c_longlong(n), nul, out.ctypes.data_as(C.c_void_p),
                      C.byref(anynul), C.byref(st))
        assert st.value == 0, st.value
        self.L.ffclos(f, C.byref(st))
        return out

    def header(self, path, hdu=1):
        f, st = C.c_void_p(), C.c_int(0)
        self.L.ffopen(C.byref(f), path.encode(), 0, C.byref(st))
        self.L.ffmahd(f, hdu, None, C.byref(st))
        hdr, nk = C.c_char_p(), C.c_int()
        self.L.ffhdr2str.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_char_p),
                                     C.POINTER(C.c_int), C.POINTER(C.c_int)]
        self.L.ffhdr2str(f, 0, None, 0, C.byref(hdr), C.byref(nk), C.byref(st))
        assert st.value == 0
        s = hdr.value.decode()
        self.L.ffclos(f, C.byref(st))
        return s


def test_read_input_map_matches_cfitsio():
    lib, syms = _host()
    rd = _fn(lib, syms, "read_input_map(char const*)", [C.c_char_p])
    rd(os.path.join(GOLDEN, "input1.fits").encode())
    assert lib.hcg_host_error() == b""
    n = 10000
    data = np.ctypeslib.as_array(C.cast(C.c_void_p.in_dll(lib, "h_data").value, C.POINTER(C.c_double)), (n,))
    lons = np.ctypeslib.as_array(C.cast(C.c_void_p.in_dll(lib, "h_lons").value, C.POINTER(C.c_double)), (n,))
    lats = np.ctypeslib.as_array(C.cast(C.c_void_p.in_dll(lib, "h_lats").value, C.POINTER(C.c_double)), (n,))
    fx = _Fits()
    p = os.path.join(GOLDEN, "input1.fits")
    assert np.array_equal(data, fx.read(p, 1, n))
    coords = fx.read(p, 2, 2 * n)
    assert np.array_equal(lons, coords[:n]) and np.array_equal(lats, coords[n:])
    gm = (C.c_uint32 * 4).in_dll(lib, "h_GMaps")  # block_warp_num, factor, data_shape, spec_dim
    assert gm[2] == n and gm[3] == 1


def test_target_header_and_wcs():
    lib, syms = _host()
    ro = _fn(lib, syms, "read_output_map(char const*)", [C.c_char_p])
    sw = _fn(lib, syms, "set_WCS()", [], C.c_int)
    path = os.path.join(GOLDEN, "target1.fits")
    ro(path.encode())
    assert lib.hcg_host_error() == b""
    hdr = C.c_char_p.in_dll(lib, "h_header").value.decode()
    assert hdr == _Fits().header(path)  # what fits_hdr2str returns, END included
    zyx = np.ctypeslib.as_array(C.cast(C.c_void_p.in_dll(lib, "h_zyx").value, C.POINTER(C.c_uint32)), (3,))
    assert list(zyx) == [1, 90, 90]
    assert sw() == 1
    m = 90 * 90
    xw = np.ctypeslib.as_array(C.cast(C.c_void_p.in_dll(lib, "h_xwcs").value, C.POINTER(C.c_double)), (m,))
    yw = np.ctypeslib.as_array(C.cast(C.c_void_p.in_dll(lib, "h_ywcs").value, C.POINTER(C.c_double)), (m,))
    xo, yo = O.wcs_sin(90, 90, 45.0, 45.0, -0.055555555555556, 0.055555555555556, 60.0, 30.0)
    assert np.abs(xw - xo).max() < 1e-10 and np.abs(yw - yo).max() < 1e-10


def _card(key, value):
    return f"{key:<8}= {value:>20}".ljust(80).encode()


def _target_variant(tmp_path, name, *subs):
    src = open(os.path.join(GOLDEN, "target1.fits"), "rb").read()
    for a, b in subs:
        assert len(a) == len(b) and src.count(a) == 1, a
        src = src.replace(a, b)
    p = tmp_path / name
    p.write_bytes(src)
    return p


def _set_wcs(lib, syms, path):
    _fn(lib, syms, "read_output_map(char const*)", [C.c_char_p])(str(path).encode())
    return _fn(lib, syms, "set_WCS()", [], C.c_int)()


@pytest.mark.parametrize("proj", ["TAN", "ARC", "ZEA", "STG"])
def test_set_wcs_zenithal_targets(tmp_path, proj):
    lib, syms = _host()
    p = _target_variant(tmp_path, proj + ".fits", (b"RA---SIN", b"RA---" + proj.encode()),
                        (b"DEC--SIN", b"DEC--" + proj.encode()))
    assert _set_wcs(lib, syms, p) == 1, lib.hcg_host_error()
    m = 90 * 90
    xw = np.ctypeslib.as_array(C.cast(C.c_void_p.in_dll(lib, "h_xwcs").value, C.POINTER(C.c_double)), (m,))
    yw = np.ctypeslib.as_array(C.cast(C.c_void_p.in_dll(lib, "h_ywcs").value, C.POINTER(C.c_double)), (m,))
    xo, yo = O.wcs_zenithal(proj, 90, 90, 45.0, 45.0, -0.055555555555556, 0.055555555555556, 60.0, 30.0)
    assert np.abs(xw - xo).max() < 1e-10 and np.abs(yw - yo).max() < 1e-10


def test_set_wcs_default_lonpole_at_the_pole(tmp_path):
    # no LONPOLE card: 0 for a target centred on the pole, else 180
    lib, syms = _host()
    src = open(os.path.join(GOLDEN, "target1.fits"), "rb").read()
    lp = src[src.index(b"LONPOLE "):src.index(b"LONPOLE ") + 80]
    crv = src[src.index(b"CRVAL2  "):src.index(b"CRVAL2  ") + 80]
    p = _target_variant(tmp_path, "pole.fits", (lp, b"COMMENT".ljust(80)), (crv, _card("CRVAL2", "90.0")))
    assert _set_wcs(lib, syms, p) == 1, lib.hcg_host_error()
    m = 90 * 90
    xw = np.ctypeslib.as_array(C.cast(C.c_void_p.in_dll(lib, "h_xwcs").value, C.POINTER(C.c_double)), (m,))
    yw = np.ctypeslib.as_array(C.cast(C.c_void_p.in_dll(lib, "h_ywcs").value, C.POINTER(C.c_double)), (m,))
    xo, yo = O.wcs_zenithal("SIN", 90, 90, 45.0, 45.0, -0.055555555555556, 0.055555555555556, 60.0, 90.0, 0.0)
    off = yo < 90.0 - 1e-9  # the longitude of the pole itself is arbitrary
    assert off.sum() == m - 1
    dx = np.abs((xw[off] - xo[off] + 180.0) % 360.0 - 180.0)
    assert dx.max() < 1e-9 and np.abs(yw - yo).max() < 1e-10
    p = _target_variant(tmp_path, "nolp.fits", (lp, b"COMMENT".ljust(80)))
    assert _set_wcs(lib, syms, p) == 1
    xw = np.ctypeslib.as_array(C.cast(C.c_void_p.in_dll(lib, "h_xwcs").value, C.POINTER(C.c_double)), (m,))
    yw = np.ctypeslib.as_array(C.cast(C.c_void_p.in_dll(lib, "h_ywcs").value, C.POINTER(C.c_double)), (m,))
    xo, yo = O.wcs_sin(90, 90, 45.0, 45.0, -0.055555555555556, 0.055555555555556, 60.0, 30.0)
    assert np.abs(xw - xo).max() < 1e-10 and np.abs(yw - yo).max() < 1e-10


def test_set_wcs_rejects_other_projections(tmp_path):
    lib, syms = _host()
    p = _target_variant(tmp_path, "car.fits", (b"RA---SIN", b"RA---CAR"), (b"DEC--SIN", b"DEC--CAR"))
    assert _set_wcs(lib, syms, p) == 0
    assert b"only zenithal" in lib.hcg_host_error()
    p = _target_variant(tmp_path, "mixed.fits", (b"DEC--SIN", b"DEC--TAN"))
    assert _set_wcs(lib, syms, p) == 0
    # SIN with a slant (PV2_1 != 0) is not the plain orthographic projection
    src = open(os.path.join(GOLDEN, "target1.fits"), "rb").read()
    rd = src[src.index(b"RADESYS "):src.index(b"RADESYS ") + 80]
    p = _target_variant(tmp_path, "slant.fits", (rd, _card("PV2_1", "0.1")))
    assert _set_wcs(lib, syms, p) == 0
    assert b"no PV2" in lib.hcg_host_error()


def test_write_output_map_roundtrip(tmp_path):
    lib, syms = _host()
    _fn(lib, syms, "read_output_map(char const*)", [C.c_char_p])(os.path.join(GOLDEN, "target1.fits").encode())
    _fn(lib, syms, "init_output()", [])()
    m = 90 * 90
    dc = np.ctypeslib.as_array(C.cast(C.c_void_p.in_dll(lib, "h_datacube").value, C.POINTER(C.c_double)), (m,))
    wc = np.ctypeslib.as_array(C.cast(C.c_void_p.in_dll(lib, "h_weightscube").value, C.POINTER(C.c_double)), (m,))
    rng = np.random.default_rng(3)
    d0 = rng.normal(size=m)
    w0 = rng.uniform(0.5, 2, size=m)
    w0[::17] = 0.0  # empty cells -> NaN
    dc[:] = d0
    wc[:] = w0
    out = tmp_path / "out1.fits"
    _fn(lib, syms, "write_output_map(char const*)", [C.c_char_p])(("!" + str(out)).encode())
    assert lib.hcg_host_error() == b""
    got = _Fits().read(str(out), 1, m)
    with np.errstate(invalid="ignore", divide="ignore"):
        expect = d0 / w0
    assert np.array_equal(np.isnan(got), np.isnan(expect))
    assert np.array_equal(got[~np.isnan(got)], expect[~np.isnan(expect)])
    hdr = _Fits().header(str(out))
    assert "BITPIX  =                  -64" in hdr and "CRVAL1  =                 60.0" in hdr
    assert "NAXIS1  =                   90" in hdr and "NAXIS3  =                    1" in hdr


def test_write_ordered_map_roundtrip(tmp_path):
    lib, syms = _host()
    src = os.path.join(GOLDEN, "input1.fits")
    _fn(lib, syms, "read_input_map(char const*)", [C.c_char_p])(src.encode())
    n = 10000
    out = tmp_path / "sorted1.fits"
    _fn(lib, syms, "write_ordered_map(char const*, char const*)", [C.c_char_p, C.c_char_p])(
        src.encode(), ("!" + str(out)).encode())
    assert lib.hcg_host_error() == b""
    fx = _Fits()
    assert np.array_equal(fx.read(str(out), 1, n), fx.read(src, 1, n))
    assert np.array_equal(fx.read(str(out), 2, 2 * n), fx.read(src, 2, 2 * n))


def test_missing_file_reports_error():
    lib, syms = _host()
    _fn(lib, syms, "read_input_map(char const*)", [C.c_char_p])(b"/nonexistent/input.fits")
    assert b"cannot open" in lib.hcg_host_error()


def test_cli_help_and_missing_args():
    exe = os.path.join(ROOT, "hcgrid_amd", "HCGrid")
    r = subprocess.run([exe, "-h"], capture_output=True, text=True)
    assert r.returncode == 1 and "useage" in r.stderr
    r = subprocess.run([exe, "--fits_path", "/tmp/"], capture_output=True, text=True)
    assert r.returncode == 1 and "required" in r.stderr
    args = ["--fits_path", "/nonexistent/", "--input_file", "in", "--target_file", "t", "--output_file", "o",
            "--fits_id", "1", "--order_arg", "1"]
    r = subprocess.run([exe, *args, "--selection", "box"], capture_output=True, text=True)
    assert r.returncode == 1 and "--selection must be compat or disc" in r.stderr
