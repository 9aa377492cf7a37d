"""Argument checks of the Python layer, on the CPU: tensors of the wrong
dtype, size, layout or device raise HcgError before any pointer reaches the
C ABI (which sees only addresses and would read or write out of bounds)."""
import pytest
import torch

from hcgrid_amd import batch, capi
from hcgrid_amd.gridder import check_tensor


class _Geo:
    """Stands in for a built Geometry: only what validate_job reads."""

    def __init__(self, n, side):
        self.n, self.side = n, side
        self.device = None


def _job(n=100, side=8, nchan=3, dtype="bf16", **over):
    tdt = torch.bfloat16 if dtype == "bf16" else torch.float16
    j = dict(geometry=_Geo(n, side), values=torch.zeros((nchan, n), dtype=tdt),
             cube=torch.zeros((nchan, side * side), dtype=torch.float32),
             wsum=torch.zeros(side * side, dtype=torch.float32), dtype=dtype)
    j.update(over)
    return j


def _raises(fn, text):
    with pytest.raises(capi.HcgError) as ei:
        fn()
    assert text in str(ei.value), str(ei.value)
    assert ei.value.code == capi.HCG_ERR_INVALID


def test_check_tensor_order_and_messages():
    t = torch.zeros(10, dtype=torch.float64)
    _raises(lambda: check_tensor(t, "w", "x", "f32"), "dtype")
    _raises(lambda: check_tensor(t, "w", "x", "f64", numel=11), "11")
    _raises(lambda: check_tensor(torch.zeros((4, 4), dtype=torch.float64).t(), "w", "x", "f64"), "contiguous")
    # everything else right: a CPU tensor is refused for not being on a GPU
    _raises(lambda: check_tensor(t, "w", "x", "f64", numel=10), "expected a GPU")
    _raises(lambda: check_tensor(t.numpy(), "w", "x", "f64"), "torch tensor")


def test_batch_job_half_precision_cube_rejected():
    # a float16 cube of the right shape would receive FP32 stores: twice its size
    j = _job(cube=torch.zeros((3, 64), dtype=torch.float16))
    _raises(lambda: batch.validate_job(0, j), "cube has dtype")


def test_batch_job_values_dtype_must_match_declared():
    j = _job(dtype="bf16", values=torch.zeros((3, 100), dtype=torch.float16))
    _raises(lambda: batch.validate_job(0, j), "values has dtype")


def test_batch_job_shapes():
    _raises(lambda: batch.validate_job(0, _job(values=torch.zeros((3, 99), dtype=torch.bfloat16))), "values have")
    _raises(lambda: batch.validate_job(0, _job(cube=torch.zeros((2, 64), dtype=torch.float32))), "cube shape")
    _raises(lambda: batch.validate_job(0, _job(wsum=torch.zeros(63, dtype=torch.float32))), "wsum has 63")
    _raises(lambda: batch.validate_job(0, _job(dtype="f32")), "dtype 'f32'")


def test_batch_job_noncontiguous_values_rejected():
    v = torch.zeros((100, 3), dtype=torch.bfloat16).t()  # (3, 100) view, transposed
    _raises(lambda: batch.validate_job(0, _job(values=v)), "not contiguous")


def test_batch_job_host_tensors_rejected_last():
    # a job whose tensors are all right except for living in host memory
    _raises(lambda: batch.validate_job(0, _job()), "expected a GPU")


def test_batch_job_without_channels_passes():
    assert batch.validate_job(0, dict(geometry=_Geo(10, 4), values=None, cube=None, wsum=None)) == 0
