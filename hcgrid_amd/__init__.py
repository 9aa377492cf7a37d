"""hcgrid_amd: HCGrid's convolution-gridding path, native to MI355X (gfx950).

The compute lives in ``libhcgrid_amd.so`` (hand-written HIP, C ABI declared in
``include/hcgrid_amd.h``).  This package is the Python host side over that
ABI; it has no CPU compute path and raises if the library is not built.
"""
from . import capi  # noqa: F401
from .capi import GAUSS1D, GAUSS2D, TAPERED_SINC, HcgError  # noqa: F401

__all__ = ["capi", "GAUSS1D", "GAUSS2D", "TAPERED_SINC", "HcgError"]
