"""Seeded synthetic inputs for the accelerated models.

Every generator draws from SplitMix64 (Steele, Lea, Flood 2014) turned into
uniform doubles as (x >> 11) * 2**-53, so the same seed gives the same arrays
on any host.  The arrays are generated once on the host and the same buffers
feed both the device path and the CPU baseline.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n consecutive SplitMix64 outputs for `seed` (uint64)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        x = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * _GOLDEN
        x = (x ^ (x >> np.uint64(30))) * _M1
        x = (x ^ (x >> np.uint64(27))) * _M2
        x = x ^ (x >> np.uint64(31))
    return x


def uniform(seed: int, n: int) -> np.ndarray:
    """n uniform doubles in [0, 1)."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


class Stream:
    """Sequential draws from one seed."""

    def __init__(self, seed: int):
        self.seed = seed
        self.pos = 0

    def take(self, n: int) -> np.ndarray:
        out = uniform(self.seed + self.pos * 0x9E3779B97F4A7C15, n) if n else np.zeros(0)
        self.pos += n + 1
        return out

    def between(self, lo: float, hi: float, n: int = 1) -> np.ndarray:
        return lo + (hi - lo) * self.take(n)


def dem(rows: int, cols: int, seed: int, base: float = 0.0, slope: float = 500.0,
        n_bumps: int = 20) -> np.ndarray:
    """Altitude: base + linear slope falling with the row index, plus Gaussian
    bumps of height U(5,50) m and width sigma U(10,80) cells at uniform centres."""
    s = Stream(seed)
    i = np.arange(rows, dtype=np.float64)[:, None]
    j = np.arange(cols, dtype=np.float64)[None, :]
    z = base + slope * (1.0 - i / max(rows - 1, 1)) + 0.0 * j
    amp = s.between(5.0, 50.0, n_bumps)
    sig = s.between(10.0, 80.0, n_bumps)
    ci = s.between(0.0, rows, n_bumps)
    cj = s.between(0.0, cols, n_bumps)
    for a, sg, a_i, a_j in zip(amp, sig, ci, cj):
        z = z + a * np.exp(-((i - a_i) ** 2 + (j - a_j) ** 2) / (2.0 * sg * sg))
    return np.ascontiguousarray(z)


def source_patch(rows: int, cols: int, seed: int, patch: int = 40,
                 lo: float = 1.0, hi: float = 10.0):
    """Debris thickness: zero except a patch x patch block of U(lo,hi) metres at
    a seeded position in the upper (upslope) quarter of the grid.
    Returns (h, (r0, c0, size))."""
    s = Stream(seed ^ 0x5DEECE66D)
    p = min(patch, max(rows // 4, 1), max(cols // 2, 1))
    r0 = int(s.between(0.0, max(rows // 4 - p, 0) + 1)[0])
    c0 = int(s.between(0.0, max(cols - p, 0) + 1)[0])
    h = np.zeros((rows, cols), dtype=np.float64)
    h[r0:r0 + p, c0:c0 + p] = s.between(lo, hi, p * p).reshape(p, p)
    return h, (r0, c0, p)


def sciddicat_inputs(rows: int, cols: int, seed: int, patch: int = 40):
    """(z, h, patch_box) for a SciddicaT run: 500 m slope + 20 bumps, 40x40 source."""
    z = dem(rows, cols, seed)
    h, box = source_patch(rows, cols, seed, patch)
    return z, h, box


def life_inputs(rows: int, cols: int, seed: int, p_alive: float = 0.3) -> np.ndarray:
    """uint8 0/1 grid, alive iff u < p_alive."""
    return (uniform(seed, rows * cols) < p_alive).astype(np.uint8).reshape(rows, cols)


def batch_instances(n: int, seed: int, lo: int = 128, hi: int = 1024):
    """Per-instance (rows, cols, p_r, p_eps, seed) for the batched ensemble:
    sides uniform in lo..hi independently, p_r in U(0.3,0.7), p_eps in U(1e-4,1e-2)."""
    s = Stream(seed)
    rows = (lo + np.floor(s.take(n) * (hi - lo + 1))).astype(int)
    cols = (lo + np.floor(s.take(n) * (hi - lo + 1))).astype(int)
    pr = s.between(0.3, 0.7, n)
    pe = s.between(1e-4, 1e-2, n)
    return [(int(rows[k]), int(cols[k]), float(pr[k]), float(pe[k]), seed * 1000 + k)
            for k in range(n)]


def read_matrix_gz(path: str) -> np.ndarray:
    """Whitespace text matrix (the reference's calLoadMatrix2Dr format), gzip'd."""
    import gzip
    with gzip.open(path, "rt") as f:
        rows = [np.array(line.split(), dtype=np.float64) for line in f if line.strip()]
    return np.vstack(rows)


# SCIARA-fv2 parameters in the order of the reference configuration file
# (OpenCAL-OMP/examples/calomp_sciara-fv2/data/2006/2006_000000000000.cfg):
# Pclock [s], PTsol [K], PTvent [K], Pr(Tsol), Pr(Tvent), Phc(Tsol) [m],
# Phc(Tvent) [m], Pcool, Prho [kg/m3], Pepsilon, Psigma [W/m2K4], Pcv [J/kgK].
# PTsol is 1100 K here (the sample file has 1143 K).
SCIARA_PARAMS = (60.0, 1100.0, 1360.0, 0.075, 0.9, 60.0, 0.4, 9.0, 2600.0, 0.9, 5.68e-8, 1150.0)


def sciara_inputs(rows: int, cols: int, seed: int, n_vents: int = 3, n_rates: int = 10,
                  steps_per_rate: int = 1000, vol_lo: float = 5.0, vol_hi: float = 20.0,
                  cell_size: float = 10.0, params=SCIARA_PARAMS):
    """Inputs of a SCIARA-fv2 run: DEM of 2000 m base + 500 m slope + 20 bumps,
    `n_vents` vents at seeded cells in the upslope quarter (kept 8 cells off the
    border), each emitting a seeded volume U(vol_lo, vol_hi) m3 per step during
    each of `n_rates` intervals of `steps_per_rate` steps.  The schedule is
    given, as the reference reads it, in m3/s (volume / Pclock), with
    emission_time = steps_per_rate * Pclock seconds; the stop condition then
    ends the run after n_rates * steps_per_rate steps.
    Returns a dict with z0, cell_size, params, vents (n_vents x 2 int32,
    row/col), rates (n_vents x n_rates), emission_time."""
    s = Stream(seed ^ 0x5C1A4A)
    z0 = dem(rows, cols, seed, base=2000.0)
    m = 8
    r_hi = max(m + 1, rows // 4)
    vr = (m + np.floor(s.take(n_vents) * (r_hi - m))).astype(np.int32)
    vc = (m + np.floor(s.take(n_vents) * max(cols - 2 * m, 1))).astype(np.int32)
    vents = np.ascontiguousarray(np.stack([vr, vc], axis=1).astype(np.int32))
    pclock = float(params[0])
    vol = s.between(vol_lo, vol_hi, n_vents * n_rates).reshape(n_vents, n_rates)
    rates = np.ascontiguousarray(vol / pclock)
    return dict(z0=z0, cell_size=float(cell_size), params=np.array(params, dtype=np.float64),
                vents=vents, rates=rates, emission_time=int(steps_per_rate * pclock))
