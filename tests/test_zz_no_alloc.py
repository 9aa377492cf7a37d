"""hcg_presort_noalloc (the pre-sort with its scratch in a static device
array, for callers that cannot allocate) against hcg_presort (workspace
argument), through the C ABI: keys and order bit-identical on two calls in a
row, the second with fewer samples (the array keeps the first call's
contents, none of which may leak into the second), plus the size limit."""
import ctypes as C

import numpy as np
import pytest

from hcgrid_amd import capi, gridder, synth

pytestmark = pytest.mark.gpu


def _presort_pair(torch, wl, hp):
    n = wl.n
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")  # noqa: E731
    lon, lat = d(wl.lon), d(wl.lat)
    ws_bytes = capi.load().hcg_presort_workspace_size(n)
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device="cuda")
    k0 = torch.empty(n + 1, dtype=torch.int32, device="cuda")
    p0 = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
    k1 = torch.full((n + 1,), -1, dtype=torch.int32, device="cuda")
    p1 = torch.full((max(n, 1),), -1, dtype=torch.int32, device="cuda")
    ptr = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    torch.cuda.synchronize()
    capi.call("hcg_presort", C.byref(hp), n, ptr(lon), ptr(lat), ptr(k0), ptr(p0), ptr(ws), ws_bytes, None)
    capi.call("hcg_presort_noalloc", C.byref(hp), n, ptr(lon), ptr(lat), ptr(k1), ptr(p1), None)
    torch.cuda.synchronize()
    return k0.cpu().numpy(), p0.cpu().numpy()[:n], k1.cpu().numpy(), p1.cpu().numpy()[:n]


def test_noalloc_presort_equals_workspace_presort_twice(gpu):
    import torch
    _, hp = gridder.prepare_grid_kernel(1, [0.02], 0.06, 0.01)
    # first a 2M-sample call, then a smaller one on the same scratch array
    for i, wl in enumerate((synth.raster_small(400, 5000, 64, seed=3), synth.raster_250k())):
        print(f"case: noalloc call {i + 1}, n={wl.n}", flush=True)
        k0, p0, k1, p1 = _presort_pair(torch, wl, hp)
        assert np.array_equal(k0, k1), "keys differ"
        assert np.array_equal(p0, p1), "order differs"
        assert k1[wl.n] == int(hp.npix)
        assert np.all(np.diff(k1[:wl.n].view(np.uint32).astype(np.int64)) >= 0)


def test_noalloc_presort_limits(gpu):
    import torch
    _, hp = gridder.prepare_grid_kernel(1, [0.02], 0.06, 0.01)
    dummy = torch.empty(16, dtype=torch.float64, device="cuda")
    p = C.c_void_p(dummy.data_ptr())
    n_over = 8388608 + 1
    rc = capi.load().hcg_presort_noalloc(C.byref(hp), n_over, p, p, p, p, None)
    assert rc == capi.HCG_ERR_UNSUPPORTED
    assert b"HCG_PRESORT_NOALLOC_MAX_N" in capi.load().hcg_last_error_message()
    # n = 0 writes only the sentinel
    keys = torch.zeros(1, dtype=torch.int32, device="cuda")
    capi.call("hcg_presort_noalloc", C.byref(hp), 0, None, None, C.c_void_p(keys.data_ptr()), p, None)
    torch.cuda.synchronize()
    assert int(keys.item()) == int(hp.npix)
