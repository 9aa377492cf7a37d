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


def _target_cards(tmp_path, name, set_cards=(), drop=()):
    """target1.fits with cards replaced, added (set_cards: (key, value text))
    or dropped; string values carry their quotes."""
    src = open(os.path.join(GOLDEN, "target1.fits"), "rb").read()
    cards = [src[i:i + 80] for i in range(0, 2880, 80)]
    end = next(k for k, c in enumerate(cards) if c.startswith(b"END"))
    keys = {k for k, _ in set_cards} | set(drop) | {"CHECKSUM", "DATASUM"}
    out = [c for c in cards[:end] if c[:8].decode().strip() not in keys]
    out += [f"{k:<8}= {v:>20}".ljust(80).encode() for k, v in set_cards]
    assert len(out) < 36
    p = tmp_path / name
    p.write_bytes(b"".join(out + [b"END".ljust(80)]).ljust(2880, b" ") + src[2880:])
    return p


def _wcs_arrays(lib, m=90 * 90):
    xw = np.ctypeslib.as_array(C.cast(C.c_void_p.in_dll(lib, "h_xwcs").value, C.POINTER(C.c_double)), (m,))
    yw = np.ctypeslib.as_array(C.cast(C.c_void_p.in_dll(lib, "h_ywcs").value, C.POINTER(C.c_double)), (m,))
    return xw.copy(), yw.copy()


def _close(got, ref, tol=1e-10):
    fin = np.isfinite(ref[1])
    assert np.array_equal(fin, np.isfinite(got[1])) and np.array_equal(fin, np.isfinite(got[0]))
    dx = np.abs((got[0][fin] - ref[0][fin] + 180.0) % 360.0 - 180.0)
    keep = np.abs(ref[1][fin]) < 89.99  # the longitude at a pole is arbitrary
    assert dx[keep].max() < tol and np.abs(got[1][fin] - ref[1][fin]).max() < tol


C1 = 0.055555555555556


def _ctype(code):
    return [("CTYPE1", f"'RA---{code}'"), ("CTYPE2", f"'DEC--{code}'")]


# (projection, crval2, lonpole card or None, latpole, pv cards, oracle pv)
GENERAL = [("CAR", 0.0, None, 90.0, [], (0.0, 0.0)), ("CAR", 30.0, None, 30.0, [], (0.0, 0.0)),
           ("CAR", -35.0, None, -90.0, [], (0.0, 0.0)), ("MER", 20.0, None, 90.0, [], (0.0, 0.0)),
           ("CEA", 10.0, None, 90.0, [("PV2_1", "0.6")], (0.6, 0.0)), ("CEA", -5.0, None, 90.0, [], (1.0, 0.0)),
           ("SFL", 45.0, None, 90.0, [], (0.0, 0.0)), ("AIT", 60.0, None, 90.0, [], (0.0, 0.0)),
           ("GLS", 30.0, None, 90.0, [], (0.0, 0.0)), ("AIT", 10.0, 0.0, 90.0, [], (0.0, 0.0)),
           ("SIN", 30.0, 180.0, 30.0, [("PV2_1", "0.12"), ("PV2_2", "-0.2")], (0.12, -0.2)),
           ("TAN", -60.0, 170.0, 30.0, [], (0.0, 0.0))]


@pytest.mark.parametrize("proj,d0,lonpole,latpole,pvc,pv", GENERAL)
def test_set_wcs_general_targets(tmp_path, proj, d0, lonpole, latpole, pvc, pv):
    # each target against the oracle's general WCS at the same matrix, with
    # a CDELT-only header and with a 20 deg PC rotation
    lib, syms = _host()
    lp = [("LONPOLE", repr(lonpole))] if lonpole is not None else []
    base = _ctype(proj) + [("CRVAL2", repr(d0)), ("LATPOLE", repr(latpole))] + lp + pvc
    drop = () if lonpole is not None else ("LONPOLE",)
    p = _target_cards(tmp_path, "g.fits", base, drop)
    assert _set_wcs(lib, syms, p) == 1, lib.hcg_host_error()
    ref = O.wcs_celestial(proj, 90, 90, 45.0, 45.0, [-C1, 0, 0, C1], 60.0, d0, lonpole, latpole, pv)
    _close(_wcs_arrays(lib), ref)
    r = np.radians(20.0)
    pc = [np.cos(r), -np.sin(r), np.sin(r), np.cos(r)]
    p = _target_cards(tmp_path, "pc.fits", base + [(k, repr(float(v))) for k, v in zip(["PC1_1", "PC1_2", "PC2_1", "PC2_2"],
                                                                                pc)], drop)
    assert _set_wcs(lib, syms, p) == 1, lib.hcg_host_error()
    ref = O.wcs_celestial(proj, 90, 90, 45.0, 45.0, [-C1 * pc[0], -C1 * pc[1], C1 * pc[2], C1 * pc[3]], 60.0, d0,
                          lonpole, latpole, pv)
    _close(_wcs_arrays(lib), ref)


def test_set_wcs_linear_forms_agree(tmp_path):
    # the same rotated grid written as PC + CDELT, as CD and as a legacy
    # CROTA2 (Paper II eq. 188): one map
    lib, syms = _host()
    rho = 17.0
    r = np.radians(rho)
    c1, c2 = -C1, C1
    pc = [np.cos(r), -np.sin(r) * c2 / c1, np.sin(r) * c1 / c2, np.cos(r)]
    base = _ctype("CAR") + [("CRVAL2", "20.0"), ("LATPOLE", "90.0")]
    forms = {"pc": [(k, repr(float(v))) for k, v in zip(["PC1_1", "PC1_2", "PC2_1", "PC2_2"], pc)],
             "cd": [(k, repr(float(v))) for k, v in zip(["CD1_1", "CD1_2", "CD2_1", "CD2_2"],
                                                 [c1 * pc[0], c1 * pc[1], c2 * pc[2], c2 * pc[3]])],
             "crota": [("CROTA2", repr(rho))]}
    maps = {}
    for name, cards in forms.items():
        drop = ("LONPOLE", "CDELT1", "CDELT2") if name == "cd" else ("LONPOLE",)
        p = _target_cards(tmp_path, name + ".fits", base + cards, drop)
        assert _set_wcs(lib, syms, p) == 1, (name, lib.hcg_host_error())
        maps[name] = _wcs_arrays(lib)
    ref = O.wcs_celestial("CAR", 90, 90, 45.0, 45.0, [c1 * pc[0], c1 * pc[1], c2 * pc[2], c2 * pc[3]], 60.0, 20.0)
    for name in forms:
        _close(maps[name], ref, 1e-11)


@pytest.mark.parametrize("d0", [30.0, -65.0])
def test_set_wcs_ncp_closed_form(tmp_path, d0):
    # NCP: l = cos(dec) sin(ra - ra0), m = (cos(dec0) - cos(dec) cos(ra - ra0)) / sin(dec0)
    lib, syms = _host()
    p = _target_cards(tmp_path, "ncp.fits", _ctype("NCP") + [("CRVAL2", repr(d0))])
    assert _set_wcs(lib, syms, p) == 1, lib.hcg_host_error()
    xw, yw = _wcs_arrays(lib)
    fin = np.isfinite(yw)
    assert fin.sum() == fin.size
    a, d, dd = np.radians(xw - 60.0), np.radians(yw), np.radians(d0)
    lm_l = np.degrees(np.cos(d) * np.sin(a))
    lm_m = np.degrees((np.cos(dd) - np.cos(d) * np.cos(a)) / np.sin(dd))
    px = (np.arange(90 * 90) % 90) + 1 - 45.0
    py = (np.arange(90 * 90) // 90) + 1 - 45.0
    assert np.abs(lm_l - (-C1) * px).max() < 1e-10 and np.abs(lm_m - C1 * py).max() < 1e-10


def test_set_wcs_rejects_what_it_cannot_map(tmp_path):
    lib, syms = _host()
    cases = [("mixed", [("CTYPE2", "'DEC--TAN'")], (), b"supported targets"),
             ("latfirst", [("CTYPE1", "'DEC--SIN'"), ("CTYPE2", "'RA---SIN'")], (), b"supported targets"),
             ("zpn", _ctype("ZPN"), (), b"supported targets"),
             ("lonpole", _ctype("CAR") + [("LONPOLE", "180.0")], (), b"LONPOLE admits no native pole"),
             ("cea", _ctype("CEA") + [("PV2_1", "1.5")], (), b"CEA needs"),
             ("phi0", [("PV1_1", "10.0")], (), b"native reference point"),
             ("singular", [("PC1_1", "1.0"), ("PC1_2", "1.0"), ("PC2_1", "1.0"), ("PC2_2", "1.0")], (), b"singular"),
             ("nocdelt", [], ("CDELT1",), b"lacks CTYPE/CRPIX/CDELT/CRVAL"),
             ("ncp0", _ctype("NCP") + [("CRVAL2", "0.0")], (), b"NCP needs")]
    for name, cards, drop, msg in cases:
        p = _target_cards(tmp_path, name + ".fits", cards, drop)
        assert _set_wcs(lib, syms, p) == 0, name
        assert msg in lib.hcg_host_error(), (name, lib.hcg_host_error())


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
