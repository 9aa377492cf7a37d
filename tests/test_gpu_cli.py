"""End to end through the drop-in: the HCGrid command line (C++ host layer over
the C ABI) on the reference's own input and target files, against the oracle."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O

from helpers import map_errors, normalise, read_fits

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "hcgrid_amd", "HCGrid")


def _oracle_map(d):
    kv, vals = read_fits(os.path.join(d, "input1.fits"), 0)
    _, coords = read_fits(os.path.join(d, "input1.fits"), 1)
    n = vals.size
    lon, lat = coords[:n], coords[n:]
    tk, _ = read_fits(os.path.join(d, "target1.fits"), 0)
    nx, ny = int(tk["NAXIS1"]), int(tk["NAXIS2"])
    xw, yw = O.wcs_sin(nx, ny, float(tk["CRPIX1"]), float(tk["CRPIX2"]), float(tk["CDELT1"]),
                       float(tk["CDELT2"]), float(tk["CRVAL1"]), float(tk["CRVAL2"]))
    sigma = (300.0 / 3600.0) / math.sqrt(8 * math.log(2))  # HCGrid/HCGrid.cpp:141-151
    r = O.pipeline(lon, lat, vals, xw, yw, ny, sigma, 5 * sigma, sigma / 2)
    return r, lon, lat, vals


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_cli_on_reference_fixtures(gpu, tmp_path, order):
    for f in ("input1.fits", "target1.fits"):
        shutil.copy(os.path.join(ROOT, "tests", "golden", f), tmp_path / f)
    d = str(tmp_path) + "/"
    print(f"case: HCGrid order_arg={order}", flush=True)
    p = subprocess.run([EXE, "--fits_path", d, "--input_file", "input", "--target_file", "target",
                        "--output_file", "output", "--sorted_file", "sorted", "--fits_id", "1",
                        "--beam_size", "300", "--order_arg", str(order)], capture_output=True, text=True,
                       timeout=120)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0
    assert "kernel elapsed time=" in p.stdout and "solving_gridding time=" in p.stdout
    kv, m = read_fits(str(tmp_path / "output1.fits"), 0)
    assert kv["BITPIX"] == "-64" and kv["NAXIS1"] == "90" and kv["NAXIS2"] == "90"
    r, lon, lat, vals = _oracle_map(d)
    o = normalise(r["datacube"], r["weightscube"])
    e = map_errors(m, o)
    print(e, flush=True)
    assert e["nan_match"] and e["max_rel"] <= 1e-6 and e["cell_rel"] <= 1e-6
    assert e["cells"] > 1000
    # the sorted input: values and coordinates in the stable HEALPix order
    _, sv = read_fits(str(tmp_path / "sorted1.fits"), 0)
    _, sc = read_fits(str(tmp_path / "sorted1.fits"), 1)
    n = vals.size
    perm = r["perm"]
    assert np.array_equal(sv, vals[perm].astype(np.float32).astype(np.float64))
    assert np.array_equal(sc[:n], lon[perm].astype(np.float32).astype(np.float64))
    assert np.array_equal(sc[n:], lat[perm].astype(np.float32).astype(np.float64))


def _rewrite_target(src, dst, set_cards, drop=()):
    """A copy of a one-block FITS target with header cards replaced or added."""
    raw = open(src, "rb").read()
    cards = [raw[i:i + 80] for i in range(0, 2880, 80)]
    end = next(k for k, c in enumerate(cards) if c.startswith(b"END"))
    keys = {k for k, _ in set_cards} | set(drop) | {"CHECKSUM", "DATASUM"}
    out = [c for c in cards[:end] if c[:8].decode().strip() not in keys]
    out += [f"{k:<8}= {v:>20}".ljust(80).encode() for k, v in set_cards]
    with open(dst, "wb") as f:
        f.write(b"".join(out + [b"END".ljust(80)]).ljust(2880, b" ") + raw[2880:])


# (CTYPE code, extra cards, oracle matrix rotation in degrees, oracle pv)
GENERAL_TARGETS = [("CAR", [("LATPOLE", "90.0")], 0.0, (0.0, 0.0)),
                   ("AIT", [("LATPOLE", "90.0"), ("PC1_1", repr(math.cos(math.radians(30)))),
                            ("PC1_2", repr(-math.sin(math.radians(30)))), ("PC2_1", repr(math.sin(math.radians(30)))),
                            ("PC2_2", repr(math.cos(math.radians(30))))], 30.0, (0.0, 0.0)),
                   ("SIN", [("PV2_1", "0.05"), ("PV2_2", "-0.1")], 0.0, (0.05, -0.1))]


@pytest.mark.parametrize("code,cards,rot,pv", GENERAL_TARGETS)
def test_cli_general_targets(gpu, tmp_path, code, cards, rot, pv):
    # the reference's input on CAR, rotated AIT and slanted SIN targets: the
    # drop-in's map against the oracle's pipeline on the oracle's WCS
    shutil.copy(os.path.join(ROOT, "tests", "golden", "input1.fits"), tmp_path / "input1.fits")
    lp = ("LONPOLE",) if code != "SIN" else ()
    _rewrite_target(os.path.join(ROOT, "tests", "golden", "target1.fits"), str(tmp_path / "target1.fits"),
                    [("CTYPE1", f"'RA---{code}'"), ("CTYPE2", f"'DEC--{code}'")] + cards, lp)
    d = str(tmp_path) + "/"
    print(f"case: HCGrid target {code}", flush=True)
    p = subprocess.run([EXE, "--fits_path", d, "--input_file", "input", "--target_file", "target",
                        "--output_file", "output", "--fits_id", "1", "--beam_size", "300", "--order_arg", "1"],
                       capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0
    _, m = read_fits(str(tmp_path / "output1.fits"), 0)
    kv, vals = read_fits(str(tmp_path / "input1.fits"), 0)
    _, coords = read_fits(str(tmp_path / "input1.fits"), 1)
    n = vals.size
    lon, lat = coords[:n], coords[n:]
    c1, c2 = -0.055555555555556, 0.055555555555556
    r = math.radians(rot)
    mat = [c1 * math.cos(r), -c1 * math.sin(r), c2 * math.sin(r), c2 * math.cos(r)]
    xw, yw = O.wcs_celestial(code, 90, 90, 45.0, 45.0, mat, 60.0, 30.0, 180.0 if code == "SIN" else None, 90.0,
                             pv)
    sigma = (300.0 / 3600.0) / math.sqrt(8 * math.log(2))
    ro = O.pipeline(lon, lat, vals, xw, yw, 90, sigma, 5 * sigma, sigma / 2)
    e = map_errors(m, normalise(ro["datacube"], ro["weightscube"]))
    print(e, flush=True)
    assert e["nan_match"] and e["max_rel"] <= 1e-6 and e["cell_rel"] <= 1e-6
    assert e["cells"] > 1000


EXE_PLUS = os.path.join(ROOT, "hcgrid_amd", "HCGridPlus")


def _write_h5(path, lon, lat, vals):
    import ctypes as C
    lib = os.path.join(ROOT, "hcgrid_amd", "libhcgrid_h5.so")
    if not os.path.exists(lib):
        pytest.skip("libhcgrid_h5.so not built")
    L = C.CDLL(lib)
    L.hcg_h5_write_input.argtypes = [C.c_char_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                     C.c_int, C.c_char_p, C.c_size_t]
    err = C.create_string_buffer(512)
    v = np.ascontiguousarray(vals, dtype=np.float64)
    assert L.hcg_h5_write_input(path.encode(), lon.size, lon.ctypes.data, lat.ctypes.data, v.shape[0],
                                v.ctypes.data, 0, err, 512) == 0, err.value


@pytest.mark.parametrize("cube_dtype", ["f64", "bf16"])
def test_plus_cli_hdf5_cube(gpu, tmp_path, cube_dtype):
    # HCGrid+'s flow on the reference's own field: the coordinates and values
    # of input1.fits as channel 0 of a 5-channel HDF5 file, target1.fits as
    # the target; every channel of the output cube against the oracle
    import torch
    shutil.copy(os.path.join(ROOT, "tests", "golden", "target1.fits"), tmp_path / "target1.fits")
    _, vals = read_fits(os.path.join(ROOT, "tests", "golden", "input1.fits"), 0)
    _, coords = read_fits(os.path.join(ROOT, "tests", "golden", "input1.fits"), 1)
    n = vals.size
    lon, lat = coords[:n].copy(), coords[n:].copy()
    rng = np.random.default_rng(3)
    V = np.stack([vals] + [vals * (1 + 0.5 * c) + rng.standard_normal(n) for c in range(1, 5)])
    _write_h5(str(tmp_path / "input1.hdf5"), lon, lat, V)
    d = str(tmp_path) + "/"
    print(f"case: HCGridPlus cube_dtype={cube_dtype} n={n}", flush=True)
    p = subprocess.run([EXE_PLUS, "--fits_path", d, "--input_file", "input", "--target_file", "target",
                        "--output_file", "output", "--sorted_file", "sorted", "--fits_id", "1", "--beam_size", "300",
                        "--order_arg", "2", "--cube_dtype", cube_dtype], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0
    assert "channels=5" in p.stdout and "solving_gridding time=" in p.stdout
    kv, m = read_fits(str(tmp_path / "output1.fits"), 0)
    assert kv["BITPIX"] == "-64" and kv["NAXIS"] == "3" and kv["NAXIS3"] == "5"
    tk, _ = read_fits(str(tmp_path / "target1.fits"), 0)
    nx, ny = int(tk["NAXIS1"]), int(tk["NAXIS2"])
    xw, yw = O.wcs_sin(nx, ny, float(tk["CRPIX1"]), float(tk["CRPIX2"]), float(tk["CDELT1"]),
                       float(tk["CDELT2"]), float(tk["CRVAL1"]), float(tk["CRVAL2"]))
    sigma = (300.0 / 3600.0) / math.sqrt(8 * math.log(2))
    if cube_dtype == "bf16":  # both sides grid the values the device sees
        Vg = torch.tensor(V.astype(np.float32)).to(torch.bfloat16).float().numpy().astype(np.float64)
        tol = 1e-4
    else:
        Vg, tol = V, 1e-6
    r = O.pipeline(lon, lat, Vg[0], xw, yw, ny, sigma, 5 * sigma, sigma / 2)
    if cube_dtype == "bf16":  # 16-bit values are exact in the FP32 multi-channel restatement
        out, ws = O.grid_compat_multi(r["h"], r["g"], r["lon_s"], r["lat_s"], Vg[:, r["perm"]], None,
                                      r["keys_sorted"], r["firstring"], r["usedrings"], r["lut"], xw, yw, 0, ny)
        maps = [normalise(out[c], ws) for c in range(5)]
    else:  # the FP64 restatement per channel
        maps = [normalise(r["datacube"], r["weightscube"])]
        for c in range(1, 5):
            rc = O.pipeline(lon, lat, Vg[c], xw, yw, ny, sigma, 5 * sigma, sigma / 2)
            maps.append(normalise(rc["datacube"], rc["weightscube"]))
    cube = m.reshape(5, nx * ny)
    for c in range(5):
        e = map_errors(cube[c], maps[c])
        print(f"  channel {c}: {e}", flush=True)
        assert e["nan_match"] and e["max_rel"] <= tol and e["cells"] > 1000
    # the sorted input: 1 x N x 5 values and N x 2 coordinates in the stable order
    skv, sv = read_fits(str(tmp_path / "sorted1.fits"), 0)
    _, sc = read_fits(str(tmp_path / "sorted1.fits"), 1)
    perm = r["perm"]
    assert skv["NAXIS3"] == "5"
    assert np.array_equal(sv.reshape(5, n), V[:, perm])
    assert np.array_equal(sc[:n], lon[perm]) and np.array_equal(sc[n:], lat[perm])


def test_plus_cli_reference_target_1895(gpu, tmp_path):
    # the reference's HCGrid+ target fixture (test_data/HDF5/hdf_target1895.fits:
    # a 500 x 500 x 1 SIN cube header at 0.01 deg around (60, 30)); its HDF5
    # input is not in the reference, so a raster over the field (about 200k
    # samples, 3 channels) is written with the library's own HDF5 writer.
    # Every channel of the output cube against the oracle's FP64 pipeline.
    shutil.copy(os.path.join(ROOT, "tests", "golden", "hdf_target1895.fits"), tmp_path / "hdf_target1895.fits")
    tk, _ = read_fits(str(tmp_path / "hdf_target1895.fits"), 0)
    assert tk["NAXIS"] == "3" and tk["NAXIS1"] == "500" and tk["NAXIS2"] == "500"
    nx, ny = int(tk["NAXIS1"]), int(tk["NAXIS2"])
    ra0, dec0 = float(tk["CRVAL1"]), float(tk["CRVAL2"])
    rng = np.random.default_rng(1895)
    dec = np.arange(dec0 - 2.7, dec0 + 2.7, 0.015)
    lon = np.concatenate([np.arange(ra0 - 2.7, ra0 + 2.7, 0.01) for _ in dec])
    per = lon.size // dec.size
    lat = np.repeat(dec, per)
    lon = ra0 + (lon - ra0) / np.cos(np.radians(lat)) + rng.uniform(-0.002, 0.002, lon.size)
    n = lon.size
    base = np.exp(-((lon - ra0) ** 2 + (lat - dec0) ** 2) / 2.0) * 10.0
    V = np.stack([base + rng.standard_normal(n) * s for s in (0.1, 0.5, 1.0)])
    _write_h5(str(tmp_path / "input1895.hdf5"), lon, lat, V)
    d = str(tmp_path) + "/"
    print(f"case: HCGridPlus on hdf_target1895.fits, n={n}, 3 channels", flush=True)
    p = subprocess.run([EXE_PLUS, "--fits_path", d, "--input_file", "input", "--target_file", "hdf_target",
                        "--output_file", "output", "--sorted_file", "sorted", "--fits_id", "1895",
                        "--beam_size", "300", "--order_arg", "1", "--cube_dtype", "f64"],
                       capture_output=True, text=True, timeout=300)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0 and "channels=3" in p.stdout
    kv, m = read_fits(str(tmp_path / "output1895.fits"), 0)
    assert kv["NAXIS1"] == "500" and kv["NAXIS2"] == "500" and kv["NAXIS3"] == "3"
    xw, yw = O.wcs_sin(nx, ny, float(tk["CRPIX1"]), float(tk["CRPIX2"]), float(tk["CDELT1"]),
                       float(tk["CDELT2"]), ra0, dec0)
    sigma = (300.0 / 3600.0) / math.sqrt(8 * math.log(2))
    cube = m.reshape(3, nx * ny)
    for c in range(3):
        r = O.pipeline(lon, lat, V[c], xw, yw, ny, sigma, 5 * sigma, sigma / 2)
        e = map_errors(cube[c], normalise(r["datacube"], r["weightscube"]))
        print(f"  channel {c}: {e}", flush=True)
        assert e["nan_match"] and e["max_rel"] <= 1e-6 and e["cells"] > 200000
