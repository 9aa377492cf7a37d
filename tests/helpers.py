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


def read_fits(path: str, hdu: int = 0):
    """(cards, data) of an image HDU; a small reader independent of the product's."""
    raw = open(path, "rb").read()
    off = 0
    for i in range(hdu + 1):
        cards = []
        while True:
            block = raw[off:off + 2880]
            off += 2880
            done = False
            for k in range(36):
                c = block[80 * k:80 * k + 80].decode("ascii")
                if c.startswith("END"):
                    done = True
                    break
                cards.append(c)
            if done:
                break
        kv = {}
        for c in cards:
            if c[8:10] == "= ":
                v = c[10:].split("/")[0].strip()
                kv[c[:8].strip()] = v.strip("'").strip()
        bitpix = int(kv["BITPIX"])
        naxes = [int(kv[f"NAXIS{j}"]) for j in range(1, int(kv["NAXIS"]) + 1)]
        nel = int(np.prod(naxes)) if naxes else 0
        nb = abs(bitpix) // 8
        dt = {-32: ">f4", -64: ">f8", 16: ">i2", 32: ">i4", 8: "u1", 64: ">i8"}[bitpix]
        data = np.frombuffer(raw[off:off + nel * nb], dtype=dt).astype(np.float64)
        off += (nel * nb + 2879) // 2880 * 2880
        if i == hdu:
            return kv, data
