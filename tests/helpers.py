"""Shared comparison helpers for the parity tests."""
from __future__ import annotations

import numpy as np


def map_errors(m: np.ndarray, o: np.ndarray) -> dict:
    """Peak-normalised max and RMS relative error, NaN agreement, and the
    per-cell relative error over cells with |o| >= 1e-2 max|o|."""
    nan_m, nan_o = np.isnan(m), np.isnan(o)
    good = ~nan_o & ~nan_m
    out = {"nan_match": bool(np.array_equal(nan_m, nan_o)), "cells": int(good.sum())}
    if not good.any():
        out.update(max_rel=0.0, rms_rel=0.0, cell_rel=0.0)
        return out
    mg, og = m[good], o[good]
    peak = float(np.max(np.abs(og)))
    diff = np.abs(mg - og)
    out["max_rel"] = float(diff.max() / peak) if peak > 0 else float(diff.max())
    den = float(np.sqrt(np.sum(og * og)))
    out["rms_rel"] = float(np.sqrt(np.sum(diff * diff)) / den) if den > 0 else 0.0
    big = np.abs(og) >= 1e-2 * peak
    out["cell_rel"] = float(np.max(diff[big] / np.abs(og[big]))) if big.any() else 0.0
    return out


def normalise(dc: np.ndarray, wc: np.ndarray) -> np.ndarray:
    with np.errstate(invalid="ignore", divide="ignore"):
        return dc / wc
