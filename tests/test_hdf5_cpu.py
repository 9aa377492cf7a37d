"""The HCGrid+ HDF5 input (HCGrid+/gmap.cpp:109-191) on the CPU.

The layout, from the reference's reader: group /coords with the attribute
num_samples and the datasets xcoords, ycoords; group /signal with the
datasets signal_0, signal_1, ...  The independent side is h5py (in the
image's conda Python 3.9, run as a subprocess): files it writes in that
layout are read by libhcgrid_h5.so and by the C++ host API, and files
libhcgrid_h5.so writes are read back by h5py.  Nothing here uses the GPU.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from hcgrid_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
H5_LIB = os.path.join(ROOT, "hcgrid_amd", "libhcgrid_h5.so")
HOST_LIB = os.path.join(ROOT, "hcgrid_amd", "libhcgrid_host.so")
CONDA_PY = "/opt/conda/bin/python3.9"


def _h5():
    if not os.path.exists(H5_LIB):
        pytest.skip("libhcgrid_h5.so not built (no HDF5 headers)")
    L = C.CDLL(H5_LIB)
    vp, u32, dp = C.c_void_p, C.c_uint32, C.POINTER(C.c_double)
    L.hcg_h5_read_coords.argtypes = [C.c_char_p, C.POINTER(u32), C.POINTER(u32), C.POINTER(dp), C.POINTER(dp),
                                     C.c_char_p, C.c_size_t]
    L.hcg_h5_read_signal.argtypes = [C.c_char_p, u32, u32, vp, C.c_char_p, C.c_size_t]
    L.hcg_h5_write_input.argtypes = [C.c_char_p, u32, vp, vp, u32, vp, C.c_int, C.c_char_p, C.c_size_t]
    return L


def _h5py(code):
    if not os.path.exists(CONDA_PY):
        pytest.skip("no h5py interpreter in this image")
    r = subprocess.run([CONDA_PY, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _read(L, path):
    n, nchan = C.c_uint32(), C.c_uint32()
    lon, lat = C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
    err = C.create_string_buffer(512)
    rc = L.hcg_h5_read_coords(str(path).encode(), C.byref(n), C.byref(nchan), C.byref(lon), C.byref(lat), err, 512)
    if rc != 0:
        return rc, err.value.decode(), None
    N, K = n.value, nchan.value
    x = np.ctypeslib.as_array(lon, (max(N, 1),))[:N].copy()
    y = np.ctypeslib.as_array(lat, (max(N, 1),))[:N].copy()
    libc = C.CDLL(None)
    libc.free(C.cast(lon, C.c_void_p))
    libc.free(C.cast(lat, C.c_void_p))
    v = np.empty((K, N))
    rc = L.hcg_h5_read_signal(str(path).encode(), N, K, v.ctypes.data, err, 512)
    return rc, err.value.decode(), (x, y, v)


def _sample(n=1234, nchan=3, seed=0):
    rng = np.random.default_rng(seed)
    lon = rng.uniform(170, 190, n)
    lat = rng.uniform(20, 40, n)
    v = rng.standard_normal((nchan, n))
    return lon, lat, v


@pytest.mark.parametrize("dtype", ["f8", "f4"])
def test_reader_on_h5py_files(tmp_path, dtype):
    lon, lat, v = _sample()
    np.save(tmp_path / "a.npy", np.stack([lon, lat, *v]))
    path = tmp_path / "in.hdf5"
    _h5py(f"""
import h5py, numpy as np
a = np.load({str(tmp_path / 'a.npy')!r})
with h5py.File({str(path)!r}, 'w') as f:
    g = f.create_group('coords'); g.attrs['num_samples'] = a.shape[1]
    g.create_dataset('xcoords', data=a[0].astype('{dtype}')); g.create_dataset('ycoords', data=a[1].astype('{dtype}'))
    s = f.create_group('signal')
    for c in range(a.shape[0] - 2): s.create_dataset('signal_%d' % c, data=a[2 + c].astype('{dtype}'))
""")
    rc, err, got = _read(_h5(), path)
    assert rc == 0, err
    x, y, vv = got
    cast = (lambda a: a.astype(np.float32).astype(np.float64)) if dtype == "f4" else (lambda a: a)
    assert np.array_equal(x, cast(lon)) and np.array_equal(y, cast(lat))
    assert vv.shape == v.shape and np.array_equal(vv, cast(v))


def test_writer_read_by_h5py(tmp_path):
    L = _h5()
    lon, lat, v = _sample(n=777, nchan=4, seed=1)
    path = tmp_path / "out.hdf5"
    err = C.create_string_buffer(512)
    assert L.hcg_h5_write_input(str(path).encode(), lon.size, lon.ctypes.data, lat.ctypes.data, v.shape[0],
                                np.ascontiguousarray(v).ctypes.data, 0, err, 512) == 0, err.value
    out = _h5py(f"""
import h5py, numpy as np
with h5py.File({str(path)!r}, 'r') as f:
    c = f['coords']; s = f['signal']
    a = np.stack([c['xcoords'][:], c['ycoords'][:]] + [s['signal_%d' % i][:] for i in range(len(s))])
    np.save({str(tmp_path / 'b.npy')!r}, a)
    print(int(c.attrs['num_samples']), len(s), c['xcoords'].dtype)
""")
    n, k, dt = out.split()
    assert int(n) == lon.size and int(k) == v.shape[0] and dt == "float64"
    b = np.load(tmp_path / "b.npy")
    assert np.array_equal(b[0], lon) and np.array_equal(b[1], lat) and np.array_equal(b[2:], v)


def test_reader_errors(tmp_path):
    L = _h5()
    rc, err, _ = _read(L, tmp_path / "missing.hdf5")
    assert rc != 0 and "cannot open" in err
    lon, lat, v = _sample(n=100, nchan=2)
    path = tmp_path / "x.hdf5"
    e = C.create_string_buffer(512)
    assert L.hcg_h5_write_input(str(path).encode(), 100, lon.ctypes.data, lat.ctypes.data, 2,
                                np.ascontiguousarray(v).ctypes.data, 1, e, 512) == 0
    out = np.empty((3, 100))
    assert L.hcg_h5_read_signal(str(path).encode(), 100, 3, out.ctypes.data, e, 512) != 0
    assert "signal_2 not found" in e.value.decode()
    assert L.hcg_h5_read_signal(str(path).encode(), 99, 1, out.ctypes.data, e, 512) != 0
    assert "has 100 elements" in e.value.decode()
    # no coords group
    bad = tmp_path / "bad.hdf5"
    _h5py(f"import h5py\nwith h5py.File({str(bad)!r}, 'w') as f: f.create_group('signal')")
    rc, err, _ = _read(L, bad)
    assert rc != 0 and "coords not found" in err


class _GMaps(C.Structure):  # include/hcgrid_host.h, class GMaps
    _fields_ = [("block_warp_num", C.c_uint32), ("factor", C.c_uint32), ("data_shape", C.c_uint32),
                ("spec_dim", C.c_uint32), ("kernel_type", C.c_uint32), ("disc_size", C.c_double),
                ("sphere_radius", C.c_double), ("bearing_needed", C.c_bool), ("selection", C.c_uint32),
                ("spec_dtype", C.c_uint32)]


def test_host_api_reads_hcgrid_plus_input(tmp_path):
    import re
    _h5()
    lon, lat, v = _sample(n=500, nchan=5, seed=2)
    path = tmp_path / "in1.hdf5"
    e = C.create_string_buffer(512)
    assert _h5().hcg_h5_write_input(str(path).encode(), 500, lon.ctypes.data, lat.ctypes.data, 5,
                                    np.ascontiguousarray(v).ctypes.data, 0, e, 512) == 0
    capi.load()
    lib = C.CDLL(HOST_LIB)
    lib.hcg_host_set_exit_on_error(0)
    lib.hcg_host_error.restype = C.c_char_p
    syms = subprocess.run(["nm", "-D", "--defined-only", HOST_LIB], capture_output=True, text=True,
                          check=True).stdout
    def sym(pat):
        m = re.search(r"\S+ T (" + pat + r")\n", syms)
        assert m, pat
        return getattr(lib, m.group(1))
    rc = sym(r"_Z21read_input_coordinatePKc")
    rd = sym(r"_Z15read_input_dataPKc")
    rc.argtypes = rd.argtypes = [C.c_char_p]
    rc(str(path).encode())
    assert lib.hcg_host_error() == b""
    g = _GMaps.in_dll(lib, "h_GMaps")
    assert g.data_shape == 500 and g.spec_dim == 5
    hl = C.POINTER(C.c_double).in_dll(lib, "h_lons")
    assert np.array_equal(np.ctypeslib.as_array(hl, (500,)), lon)
    rd(str(path).encode())
    assert lib.hcg_host_error() == b""
    hs = C.POINTER(C.c_double).in_dll(lib, "h_spec_data")
    assert np.array_equal(np.ctypeslib.as_array(hs, (5 * 500,)).reshape(5, 500), v)
    rc(str(tmp_path / "none.hdf5").encode())
    assert b"cannot open HDF5 file" in lib.hcg_host_error()


def test_plus_cli_arguments(tmp_path):
    exe = os.path.join(ROOT, "hcgrid_amd", "HCGridPlus")
    r = subprocess.run([exe, "-h"], capture_output=True, text=True)
    assert r.returncode == 1 and "useage" in r.stderr
    args = ["--fits_path", str(tmp_path) + "/", "--input_file", "in", "--target_file", "t", "--output_file", "o",
            "--fits_id", "1", "--order_arg", "1"]
    r = subprocess.run([exe, *args, "--cube_dtype", "f32"], capture_output=True, text=True)
    assert r.returncode == 1 and "--cube_dtype must be" in r.stderr
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "cannot open HDF5 file" in r.stderr
