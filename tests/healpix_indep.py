"""An independent RING pixelisation used to pin the oracle's ang2pix.

It shares no steps with HCGrid's loc2pix.  A point is mapped with the
HEALPix projection (Calabretta & Roukema 2007, MNRAS 381, 865; H = 4, K = 3):
  equatorial |z| <= 2/3 : x = phi, y = (3 pi / 8) z
  polar      |z| >  2/3 : s = sqrt(3 (1 - |z|)),
                          x = phi_c + (phi - phi_c) s, y = sign(z) (pi / 4)(2 - s)
with phi_c the centre of the polar facet containing phi.  In units of
pi / (4 nside) the pixel centres form the lattice X + Y = nside + 1 (mod 2)
and every pixel is the diamond |dX| + |dY| <= 1 around its centre, so the
pixel is found by rounding the rotated coordinates A = (X + Y - nside - 1)/2,
B = (X - Y - nside - 1)/2.  Ring i = 2 nside - Y_c; the index inside the ring
follows from X_c (Gorski et al. 2005, ApJ 622, 759, section 4).
"""
from __future__ import annotations

import numpy as np


def ang2pix_ring(nside: int, theta: np.ndarray, phi: np.ndarray) -> np.ndarray:
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.mod(np.asarray(phi, dtype=np.float64), 2 * np.pi)
    N = nside
    z = np.cos(theta)
    unit = np.pi / (4 * N)
    polar = np.abs(z) > 2.0 / 3.0
    # sigma from half-angles for accuracy near the poles: 1 - |z| = 2 sin^2(t/2)
    half = np.where(z > 0, np.sin(theta / 2), np.cos(theta / 2))
    s = np.sqrt(6.0) * np.abs(half)
    facet = np.floor(phi / (np.pi / 2)).clip(0, 3)
    phi_c = (2 * facet + 1) * np.pi / 4
    x = np.where(polar, phi_c + (phi - phi_c) * s, phi)
    y = np.where(polar, np.sign(z) * (np.pi / 4) * (2 - s), (3 * np.pi / 8) * z)
    X = x / unit
    Y = y / unit
    A = np.floor((X + Y - (N + 1)) / 2 + 0.5)
    B = np.floor((X - Y - (N + 1)) / 2 + 0.5)
    Xc = (A + B + N + 1).astype(np.int64)
    Yc = (A - B).astype(np.int64)
    ring = 2 * N - Yc
    npix = 12 * N * N
    ncap = 2 * N * (N - 1)
    out = np.empty(theta.shape, dtype=np.int64)

    north = ring < N
    south = ring > 3 * N
    eq = ~north & ~south
    # equatorial: X_c = 2 iphi - 2 or 2 iphi - 1 -> iphi - 1 = floor(X_c / 2) (mod 4N)
    iphi0 = np.mod(np.floor_divide(Xc, 2), 4 * N)
    out[eq] = (ncap + (ring - N) * 4 * N + iphi0)[eq]
    # north cap: X_c = (2f + 1)(N - i) + 2 iphi - 1
    f = np.floor_divide(Xc, 2 * N).clip(0, 3)
    i_n = ring
    iphi_n = (Xc - (2 * f + 1) * (N - i_n) + 1) // 2
    out[north] = (2 * i_n * (i_n - 1) + iphi_n - 1)[north]
    i_s = 4 * N - ring
    iphi_s = (Xc - (2 * f + 1) * (N - i_s) + 1) // 2
    out[south] = (npix - 2 * i_s * (i_s + 1) + iphi_s - 1)[south]
    return out.astype(np.uint64)


def ring_start_closed_form(nside: int, rings: np.ndarray) -> np.ndarray:
    """startpix(ring) as the running sum of ring sizes 4i, 4N, 4(4N - i)."""
    N = nside
    sizes = np.empty(4 * N, dtype=np.int64)
    i = np.arange(4 * N)
    sizes[:] = np.where(i < N, 4 * i, np.where(i <= 3 * N, 4 * N, 4 * (4 * N - i)))
    sizes[0] = 0
    starts = np.concatenate([[0], np.cumsum(sizes)])  # starts[r] = sum of sizes of rings < r
    return starts[np.asarray(rings, dtype=np.int64)].astype(np.uint64)
