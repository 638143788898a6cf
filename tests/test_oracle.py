"""Pins of the oracle restatement (CPU only).

SciddicaT: bit-exact against the reference OpenCAL-OMP build (oracle/_ref,
when built in this tree) on seeded inputs, full grid and active cells,
including the per-step active-cell counts; the reference's shipped 610x496
result (text-identical); mass conservation; flat-DEM / dry-cell closed forms.
Life: reference build, numpy roll Life, blinker/block/glider.
Adjoint: central finite differences of the oracle forward on tiny grids, and
linearity of the flows in p_r."""
import gzip
import os

import numpy as np
import pytest

import native
from opencal_amd import synth

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
HAVE_REF = native.ref("sciddicat") is not None


@pytest.mark.skipif(not HAVE_REF, reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("active", [False, True])
@pytest.mark.parametrize("rows,cols,steps,seed,pr,pe", [
    (64, 96, 300, 7, 0.5, 0.001), (33, 47, 200, 8, 0.35, 0.004), (80, 40, 150, 9, 0.68, 2e-4),
])
def test_sciddicat_oracle_vs_reference(rows, cols, steps, seed, pr, pe, active):
    z, h, _ = synth.sciddicat_inputs(rows, cols, seed, patch=12)
    a = native.oracle_sciddicat(z, h, pr, pe, steps, active, counts=True)
    b = native.ref_sciddicat(z, h, pr, pe, steps, active, counts=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    if active:
        assert np.array_equal(a[2][:steps], b[2][:steps])


def test_sciddicat_oracle_shipped_result():
    """OpenCAL-OMP/examples/calomp_sciddicaT/data/width_final.txt (610x496,
    4000 steps), reproduced as text."""
    z = synth.read_matrix_gz(os.path.join(GOLDEN, "sciddicaT_610x496_dem.txt.gz"))
    h = synth.read_matrix_gz(os.path.join(GOLDEN, "sciddicaT_610x496_source.txt.gz"))
    _, oh = native.oracle_sciddicat(z, h, 0.5, 0.001, 4000, False)
    text = gzip.open(os.path.join(GOLDEN, "sciddicaT_610x496_width_final.txt.gz"), "rt").read()
    assert "".join("".join("%.9f  " % v for v in row) + "\n" for row in oh) == text
    assert abs(oh.sum() - 5557.0) < 1e-6


def test_sciddicat_oracle_closed_forms():
    rows, cols = 20, 24
    # flat DEM, uniform thickness: every neighbour level equals the centre's,
    # so all are eliminated and nothing flows
    z = np.full((rows, cols), 100.0)
    h = np.full((rows, cols), 2.0)
    _, oh = native.oracle_sciddicat(z, h, 0.5, 0.001, 10, False)
    assert np.array_equal(oh, h)
    # dry grid (h <= eps): nothing moves
    h = np.full((rows, cols), 0.0005)
    _, oh = native.oracle_sciddicat(z, h, 0.5, 0.001, 10, False)
    assert np.array_equal(oh, h)
    # single wet cell on a flat plane (the init lowers z by h, so the terrain
    # under the debris is z - h): spreads symmetrically, mass conserved, the
    # outflow never exceeds r*(h - eps)
    h = np.zeros((rows, cols))
    h[10, 12] = 4.0
    _, oh = native.oracle_sciddicat(z + h, h, 0.5, 0.001, 1, False)
    assert abs(oh.sum() - 4.0) < 1e-12
    assert oh[9, 12] == oh[11, 12] == oh[10, 11] == oh[10, 13] > 0
    assert 4.0 - oh[10, 12] <= 0.5 * (4.0 - 0.001) + 1e-12


@pytest.mark.parametrize("active", [False, True])
def test_sciddicat_mass_conservation(active):
    z, h, _ = synth.sciddicat_inputs(120, 90, 31, patch=20)
    _, oh = native.oracle_sciddicat(z, h, 0.45, 0.002, 400, active)
    assert abs(oh.sum() - h.sum()) <= 1e-10 * h.sum()
    assert oh.min() >= 0.0


@pytest.mark.skipif(native.ref("life") is None, reason="reference build (oracle/_ref) not present")
def test_life_oracle_vs_reference():
    a0 = synth.life_inputs(64, 80, 3)
    assert np.array_equal(native.oracle_life(a0, 100), native.ref_life(a0, 100))


def test_life_oracle_closed_forms():
    g = np.zeros((16, 16), np.uint8)
    g[5, 4:7] = 1
    assert np.array_equal(native.oracle_life(g, 2), g)
    gl = np.zeros((16, 16), np.uint8)
    gl[1, 2] = gl[2, 3] = gl[3, 1] = gl[3, 2] = gl[3, 3] = 1
    assert np.array_equal(native.oracle_life(gl, 4), np.roll(np.roll(gl, 1, 0), 1, 1))
    a0 = synth.life_inputs(30, 34, 9)
    a = a0.astype(int)
    for _ in range(20):
        s = sum(np.roll(np.roll(a, di, 0), dj, 1) for di in (-1, 0, 1) for dj in (-1, 0, 1) if (di, dj) != (0, 0))
        a = ((s == 3) | ((s == 2) & (a == 1))).astype(int)
    assert np.array_equal(native.oracle_life(a0, 20), a.astype(np.uint8))


def _loss_and_grads(z, h0, pr, pe, steps, target):
    hT = h0.copy()
    for _ in range(steps):
        hT = native.oracle_step(z, hT, pr, pe)
    lam = 2.0 * (hT - target)
    lh, lz, dr, de = native.oracle_adjoint(z, h0, pr, pe, steps, lam)
    return ((hT - target) ** 2).sum(), lh, lz, dr, de


def test_adjoint_oracle_finite_differences():
    """Central differences (h=1e-6) of L = sum (h_T - target)^2 w.r.t. p_r,
    eps, one h0 cell and one z cell, away from elimination switches."""
    rng = np.random.default_rng(4)
    rows, cols, steps = 12, 14, 6
    z = 10.0 + rng.random((rows, cols))
    h0 = 0.5 + rng.random((rows, cols))
    target = 0.5 + rng.random((rows, cols))
    pr, pe = 0.5, 0.01
    L, lh, lz, dr, de = _loss_and_grads(z, h0, pr, pe, steps, target)

    def loss(z_, h_, pr_, pe_):
        hT = h_.copy()
        for _ in range(steps):
            hT = native.oracle_step(z_, hT, pr_, pe_)
        return ((hT - target) ** 2).sum()

    eps = 1e-6
    fd_r = (loss(z, h0, pr + eps, pe) - loss(z, h0, pr - eps, pe)) / (2 * eps)
    fd_e = (loss(z, h0, pr, pe + eps) - loss(z, h0, pr, pe - eps)) / (2 * eps)
    assert abs(fd_r - dr) <= 1e-5 * max(1.0, abs(fd_r))
    assert abs(fd_e - de) <= 1e-5 * max(1.0, abs(fd_e))
    for (i, j) in [(3, 4), (7, 9), (0, 13)]:
        hp, hm = h0.copy(), h0.copy()
        hp[i, j] += eps
        hm[i, j] -= eps
        fd = (loss(z, hp, pr, pe) - loss(z, hm, pr, pe)) / (2 * eps)
        assert abs(fd - lh[i, j]) <= 1e-5 * max(1.0, abs(fd))
        zp, zm = z.copy(), z.copy()
        zp[i, j] += eps
        zm[i, j] -= eps
        fd = (loss(zp, h0, pr, pe) - loss(zm, h0, pr, pe)) / (2 * eps)
        assert abs(fd - lz[i, j]) <= 1e-5 * max(1.0, abs(fd))


def test_step_linear_in_p_r():
    """With frozen eliminations the outflows are linear in p_r:
    h'(2r) - h = 2 (h'(r) - h) for small r on an interior-only configuration."""
    rng = np.random.default_rng(5)
    z = 10.0 + rng.random((10, 10))
    h = 0.5 + rng.random((10, 10))
    d1 = native.oracle_step(z, h, 0.1, 0.01) - h
    d2 = native.oracle_step(z, h, 0.2, 0.01) - h
    assert np.allclose(d2, 2 * d1, rtol=1e-9, atol=1e-13)


def test_adjoint_oracle_all_levels_eliminated():
    """h just above eps at ~500 m: rounding makes the level average equal the
    levels, every level is eliminated (count 0) and the reference moves
    nothing; the adjoint of that cell is zero, not 0/0."""
    z = np.full((6, 6), 500.0)
    h = np.zeros((6, 6))
    h[2, 2] = 0.001 + 1e-14
    z[2, 2] -= h[2, 2]
    assert native.oracle_step(z, h, 0.5, 0.001)[2, 2] == h[2, 2]
    lh, lz, dr, de = native.oracle_adjoint(z, h, 0.5, 0.001, 3, np.ones((6, 6)))
    assert np.isfinite(lh).all() and np.isfinite(lz).all()
    assert np.array_equal(lh, np.ones((6, 6))) and dr == 0.0 and de == 0.0
