"""Seeded synthetic raster-scan workloads (numpy).

Raster scans: constant-Dec scan lines, samples evenly spaced in RA plus a
uniform jitter; values are Gaussian point sources (a 0.025 deg FWHM beam)
plus white noise.  Sky coordinates are FP64 degrees.  The target is a square
SIN-projection map.  Generated with numpy's PCG64 so that every run with a
given seed produces the same bytes.

Named workloads:
  raster_250k  : 200 lines x 1250 samples over 4 x 4 deg at (180, 30);
                 400 x 400 map, 0.01 deg pixels; sigma 0.02 deg, R = 3 sigma.
  raster_20m   : 1000 lines x 20000 samples over 20 x 20 deg at (180, 30);
                 2048 x 2048 map, 0.01 deg pixels; same kernel.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

BEAM_FWHM_DEG = 0.025
BEAM_SIGMA_DEG = BEAM_FWHM_DEG / math.sqrt(8.0 * math.log(2.0))


@dataclasses.dataclass
class Target:
    side: int            # NAXIS1 = NAXIS2
    cdelt: float         # deg, |CDELT1| = CDELT2
    crval1: float
    crval2: float
    crpix: float         # CRPIX1 = CRPIX2 (the reference generator's N/2)
    proj: str = "SIN"    # zenithal projection: SIN, TAN, ARC, ZEA or STG

    def header_cards(self) -> dict:
        return {
            "NAXIS": 3, "NAXIS1": self.side, "NAXIS2": self.side, "NAXIS3": 1,
            "CTYPE1": f"RA---{self.proj}", "CTYPE2": f"DEC--{self.proj}", "CUNIT1": "deg", "CUNIT2": "deg",
            "CDELT1": -self.cdelt, "CDELT2": self.cdelt, "CRPIX1": self.crpix,
            "CRPIX2": self.crpix, "CRVAL1": self.crval1, "CRVAL2": self.crval2,
        }


@dataclasses.dataclass
class Workload:
    name: str
    lon: np.ndarray      # f64 deg
    lat: np.ndarray      # f64 deg
    values: np.ndarray   # f32
    target: Target
    sigma: float = 0.02
    radius: float = 0.06
    hpx_res: float = 0.01

    @property
    def n(self) -> int:
        return int(self.lon.size)


def _haversine_deg(l1, b1, l2, b2):
    d2r = math.pi / 180.0
    l1, b1, l2, b2 = (np.asarray(v, dtype=np.float64) * d2r for v in (l1, b1, l2, b2))
    a = np.sin((b1 - b2) / 2.0) ** 2 + np.cos(b1) * np.cos(b2) * np.sin((l1 - l2) / 2.0) ** 2
    return 2.0 * np.arcsin(np.sqrt(a)) / d2r


def raster(n_lines: int, n_per_line: int, dec0: float, ddec: float, ra0: float, dra: float,
           jitter: float, n_src: int, seed: int, noise: float = 1.0):
    """Raster-scan coordinates and single-channel values."""
    rng = np.random.Generator(np.random.PCG64(seed))
    k = np.arange(n_lines, dtype=np.float64)
    j = np.arange(n_per_line, dtype=np.float64)
    lat = np.repeat(dec0 + ddec * k, n_per_line)
    lon = np.tile(ra0 + dra * j, n_lines) + rng.uniform(-jitter, jitter, size=n_lines * n_per_line)
    lo_ra, hi_ra = ra0, ra0 + dra * (n_per_line - 1)
    lo_de, hi_de = dec0, dec0 + ddec * (n_lines - 1)
    amp = rng.uniform(1.0, 10.0, size=n_src)
    sra = rng.uniform(lo_ra, hi_ra, size=n_src)
    sde = rng.uniform(lo_de, hi_de, size=n_src)
    values = (noise * rng.standard_normal(lon.size)).astype(np.float64)
    reach = 6.0 * BEAM_SIGMA_DEG
    for a, r, d in zip(amp, sra, sde):
        # only lines within reach of the source can be touched
        rows = np.nonzero(np.abs(dec0 + ddec * k - d) < reach)[0]
        if rows.size == 0:
            continue
        sl = slice(int(rows[0]) * n_per_line, (int(rows[-1]) + 1) * n_per_line)
        dist = _haversine_deg(lon[sl], lat[sl], r, d)
        values[sl] += a * np.exp(-dist * dist / (2.0 * BEAM_SIGMA_DEG ** 2))
    return lon, lat, values.astype(np.float32), (amp, sra, sde)


def raster_250k(seed: int = 0) -> Workload:
    lon, lat, val, _ = raster(200, 1250, 28.01, 0.02, 178.0016, 0.0032, 0.0005, 20, seed)
    return Workload("raster_250k", lon, lat, val, Target(400, 0.01, 180.0, 30.0, 200.0))


def raster_20m(seed: int = 1) -> Workload:
    lon, lat, val, _ = raster(1000, 20000, 20.01, 0.02, 170.0005, 0.001, 0.0005, 20, seed)
    return Workload("raster_20m", lon, lat, val, Target(2048, 0.01, 180.0, 30.0, 1024.0))


def raster_small(n_lines: int, n_per_line: int, side: int, dec_c: float = 30.0, ra_c: float = 180.0,
                 spacing_dec: float = 0.02, spacing_ra: float = 0.0032, cdelt: float = 0.01,
                 seed: int = 7, n_src: int = 5) -> Workload:
    """A small raster centred on (ra_c, dec_c), sized for CPU-speed tests."""
    dec0 = dec_c - spacing_dec * (n_lines - 1) / 2.0
    ra0 = ra_c - spacing_ra * (n_per_line - 1) / 2.0
    lon, lat, val, _ = raster(n_lines, n_per_line, dec0, spacing_dec, ra0, spacing_ra,
                              spacing_ra / 6.4, n_src, seed)
    return Workload(f"raster_{n_lines}x{n_per_line}", lon, lat, val,
                    Target(side, cdelt, ra_c, dec_c, side / 2.0))


@dataclasses.dataclass
class CubeWorkload:
    name: str
    lon: np.ndarray          # f64 deg
    lat: np.ndarray          # f64 deg
    values: object           # torch tensor (nchan, n) bf16 on the device, channel-major
    target: Target
    nchan: int
    sigma: float = 0.02
    radius: float = 0.06
    hpx_res: float = 0.01

    @property
    def n(self) -> int:
        return int(self.lon.size)


def cube_coords(n_lines: int = 500, n_per_line: int = 4000, dec0: float = 25.01, ddec: float = 0.02,
                ra0: float = 175.00125, dra: float = 0.0025, seed: int = 2):
    """The sample coordinates of raster_cube (same arguments, same bytes),
    without the channel values; also returns the generator, positioned where
    raster_cube continues with the sources."""
    rng = np.random.Generator(np.random.PCG64(seed))
    j = np.arange(n_per_line, dtype=np.float64)
    k = np.arange(n_lines, dtype=np.float64)
    lat = np.repeat(dec0 + ddec * k, n_per_line)
    lon = np.tile(ra0 + dra * j, n_lines) + rng.uniform(-dra / 6.4, dra / 6.4, size=n_lines * n_per_line)
    return lon, lat, rng


def raster_cube(n_lines: int = 500, n_per_line: int = 4000, dec0: float = 25.01, ddec: float = 0.02,
                ra0: float = 175.00125, dra: float = 0.0025, nchan: int = 1024, side: int = 1024,
                n_src: int = 20, seed: int = 2, device: str = "cuda", dtype: str = "bf16") -> CubeWorkload:
    """Spectral raster: per-source Gaussian line profiles (centre U[200, 800]
    channels scaled to nchan/1024, width U[5, 30], amplitude U[1, 10]) on a
    0.025 deg beam, plus N(0, 0.5) noise.  Default: 500 lines x 4000 samples
    over 10 x 10 deg at (180, 30), 1024 channels, 1024 x 1024 map at 0.01 deg.
    Values are generated on the device (torch, seeded) in channel-major order."""
    import torch

    lon, lat, rng = cube_coords(n_lines, n_per_line, dec0, ddec, ra0, dra, seed)
    k = np.arange(n_lines, dtype=np.float64)
    n = lon.size
    tdt = torch.bfloat16 if dtype == "bf16" else torch.float16
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    vals = torch.empty((nchan, n), dtype=tdt, device=device)
    chunk = max(1, (1 << 26) // max(n, 1))
    for c0 in range(0, nchan, chunk):
        c1 = min(nchan, c0 + chunk)
        vals[c0:c1] = (0.5 * torch.randn((c1 - c0, n), generator=gen, device=device)).to(tdt)
    amp = rng.uniform(1.0, 10.0, n_src)
    sra = rng.uniform(ra0, ra0 + dra * (n_per_line - 1), n_src)
    sde = rng.uniform(dec0, dec0 + ddec * (n_lines - 1), n_src)
    mu = rng.uniform(200.0, 800.0, n_src) * (nchan / 1024.0)
    wid = rng.uniform(5.0, 30.0, n_src) * max(nchan / 1024.0, 1.0 / 30.0)
    reach = 6.0 * BEAM_SIGMA_DEG
    ch = torch.arange(nchan, device=device, dtype=torch.float32)[:, None]
    for a, r, d, m_, w_ in zip(amp, sra, sde, mu, wid):
        rows = np.nonzero(np.abs(dec0 + ddec * k - d) < reach)[0]
        if rows.size == 0:
            continue
        sl = slice(int(rows[0]) * n_per_line, (int(rows[-1]) + 1) * n_per_line)
        dist = _haversine_deg(lon[sl], lat[sl], r, d)
        idx = np.nonzero(dist < reach)[0] + sl.start
        if idx.size == 0:
            continue
        spat = a * np.exp(-dist[idx - sl.start] ** 2 / (2.0 * BEAM_SIGMA_DEG ** 2))
        prof = torch.exp(-((ch - float(m_)) ** 2) / (2.0 * float(w_) ** 2))
        add = prof * torch.as_tensor(spat, dtype=torch.float32, device=device)[None, :]
        it = torch.as_tensor(idx, device=device)
        vals[:, it] = (vals[:, it].float() + add).to(tdt)
    dec_c = dec0 + ddec * (n_lines - 1) / 2.0
    ra_c = ra0 + dra * (n_per_line - 1) / 2.0
    return CubeWorkload(f"raster_cube_{n // 1000}k_x{nchan}", lon, lat, vals,
                        Target(side, 0.01, round(ra_c, 6), round(dec_c, 6), side / 2.0), nchan)


@dataclasses.dataclass
class BatchSpec:
    """One job of a batched workload (coordinates on the host, values made
    on the device by batch_values)."""
    name: str
    lon: np.ndarray
    lat: np.ndarray
    target: Target
    nchan: int
    dtype: str
    sigma: float
    seed: int

    @property
    def n(self) -> int:
        return int(self.lon.size)

    @property
    def radius(self) -> float:
        return 3.0 * self.sigma

    @property
    def hpx_res(self) -> float:
        return self.sigma / 2.0


def batch_field(n: int, side: int, dec_c: float, ra_c: float = 180.0, cdelt: float = 0.01, seed: int = 0):
    """Raster over a square field of side * cdelt degrees on the sky centred
    on (ra_c, dec_c): round(sqrt(n)) constant-Dec lines, samples evenly
    spaced in RA (spacing scaled by 1/cos(dec_c) so the field is square on
    the sky) with a uniform jitter of +-1/6.4 of the spacing.  The centre is
    moved towards the equator where the field would pass 89 deg."""
    rng = np.random.Generator(np.random.PCG64(seed))
    field = side * cdelt
    half = field / 2.0
    dec_c = float(np.clip(dec_c, -89.0 + half, 89.0 - half))
    lines = max(1, int(round(math.sqrt(n))))
    per_line = max(1, int(round(n / lines)))
    ddec = field / lines
    dra = field / per_line / math.cos(math.radians(dec_c))
    k = np.arange(lines, dtype=np.float64)
    j = np.arange(per_line, dtype=np.float64)
    lat = np.repeat(dec_c - half + ddec * (k + 0.5), per_line)
    lon = (np.tile(ra_c - dra * per_line / 2.0 + dra * (j + 0.5), lines)
           + rng.uniform(-dra / 6.4, dra / 6.4, size=lines * per_line))
    return lon, lat, Target(side, cdelt, ra_c, round(dec_c, 6), side / 2.0)


def batch_specs(njobs: int = 64, seed: int = 3, sides=(128, 256, 512, 1024, 2048), chans=(1, 16, 64, 256, 512),
                sigmas=(0.015, 0.02, 0.03), n_range=(5e4, 5e6), dec_range=(-60.0, 80.0)) -> list:
    """Uneven jobs: N log-uniform in n_range, map side, channel count, value
    type (bf16/f16) and sigma drawn uniformly from the lists, field centre
    Dec uniform in dec_range at RA 180 deg (never near RA 0)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(njobs):
        n = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        side = int(rng.choice(sides))
        nchan = int(rng.choice(chans))
        dtype = "bf16" if rng.integers(2) == 0 else "f16"
        sigma = float(rng.choice(sigmas))
        dec = float(rng.uniform(*dec_range))
        lon, lat, tgt = batch_field(n, side, dec, seed=seed * 1000 + i)
        out.append(BatchSpec(f"job{i}", lon, lat, tgt, nchan, dtype, sigma, seed * 1000 + i))
    return out


def batch_values(spec: BatchSpec, device: str = "cuda"):
    """Channel-major values [nchan][n] of a job on the device: N(0, 0.5)
    noise plus one Gaussian line (centre and width drawn per job) on a
    spatial Gaussian of 0.3 x the field width around the field centre."""
    import torch

    gen = torch.Generator(device=device)
    gen.manual_seed(spec.seed)
    tdt = torch.bfloat16 if spec.dtype == "bf16" else torch.float16
    n, c = spec.n, spec.nchan
    rng = np.random.Generator(np.random.PCG64(spec.seed + 1))
    mu = rng.uniform(0.2, 0.8) * c
    wid = max(1.0, rng.uniform(0.02, 0.1) * c)
    amp = rng.uniform(1.0, 10.0)
    t = spec.target
    scale = 0.3 * t.side * t.cdelt
    dy = spec.lat - t.crval2
    dx = (spec.lon - t.crval1) * math.cos(math.radians(t.crval2))
    spat = torch.as_tensor(amp * np.exp(-(dx * dx + dy * dy) / (2 * scale * scale)), dtype=torch.float32,
                           device=device)
    ch = torch.arange(c, device=device, dtype=torch.float32)
    prof = torch.exp(-((ch - mu) ** 2) / (2 * wid * wid))
    out = torch.empty((c, n), dtype=tdt, device=device)
    step = max(1, (1 << 26) // max(n, 1))
    for c0 in range(0, c, step):
        c1 = min(c, c0 + step)
        blk = 0.5 * torch.randn((c1 - c0, n), generator=gen, device=device) + prof[c0:c1, None] * spat[None, :]
        out[c0:c1] = blk.to(tdt)
    return out


def field_size(n: int) -> int:
    """Samples in a batch_field/device_field raster asked for n samples:
    round(sqrt(n)) lines of round(n / lines) samples."""
    lines = max(1, int(round(math.sqrt(n))))
    return lines * max(1, int(round(n / lines)))


def device_field(n: int, side: int, dec_c: float, seed: int, ra_c: float = 180.0, cdelt: float = 0.01,
                 device: str = "cuda"):
    """batch_field's raster made on the device (torch, on the current
    stream; the jitter comes from torch's generator, so the samples differ
    from batch_field's).  Returns (lon, lat, Target), float64 tensors."""
    import torch

    field = side * cdelt
    half = field / 2.0
    dec_c = float(np.clip(dec_c, -89.0 + half, 89.0 - half))
    lines = max(1, int(round(math.sqrt(n))))
    per_line = max(1, int(round(n / lines)))
    ddec = field / lines
    dra = field / per_line / math.cos(math.radians(dec_c))
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    k = torch.arange(lines, dtype=torch.float64, device=device)
    j = torch.arange(per_line, dtype=torch.float64, device=device)
    lat = (dec_c - half + ddec * (k + 0.5)).repeat_interleave(per_line)
    jit = (torch.rand(lines * per_line, generator=gen, device=device, dtype=torch.float64) * 2 - 1) * (dra / 6.4)
    lon = (ra_c - dra * per_line / 2.0 + dra * (j + 0.5)).repeat(lines) + jit
    return lon, lat, Target(side, cdelt, ra_c, round(dec_c, 6), side / 2.0)


def device_values(nchan: int, dtype: str, lon, lat, target: Target, seed: int, device: str = "cuda"):
    """batch_values made entirely on the device from device coordinates."""
    import torch

    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed) + 7)
    tdt = torch.bfloat16 if dtype == "bf16" else torch.float16
    n = int(lon.numel())
    rng = np.random.Generator(np.random.PCG64(int(seed) + 1))
    mu = rng.uniform(0.2, 0.8) * nchan
    wid = max(1.0, rng.uniform(0.02, 0.1) * nchan)
    amp = rng.uniform(1.0, 10.0)
    scale = 0.3 * target.side * target.cdelt
    dy = lat - target.crval2
    dx = (lon - target.crval1) * math.cos(math.radians(target.crval2))
    spat = (amp * torch.exp(-(dx * dx + dy * dy) / (2 * scale * scale))).float()
    ch = torch.arange(nchan, device=device, dtype=torch.float32)
    prof = torch.exp(-((ch - mu) ** 2) / (2 * wid * wid))
    out = torch.empty((nchan, n), dtype=tdt, device=device)
    step = max(1, (1 << 26) // max(n, 1))
    for c0 in range(0, nchan, step):
        c1 = min(nchan, c0 + step)
        blk = 0.5 * torch.randn((c1 - c0, n), generator=gen, device=device) + prof[c0:c1, None] * spat[None, :]
        out[c0:c1] = blk.to(tdt)
    return out
