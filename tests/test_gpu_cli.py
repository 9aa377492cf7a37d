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
