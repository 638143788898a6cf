"""Batched SciddicaT ensemble: every instance equals its own single run
(bit-exact), against the oracle and against the single-model GPU path;
uneven sizes including odd column counts and grids narrower than a strip;
the wet-tile queue (dry tiles skipped from step 2 on) under several tile
row counts and launch forms, including runs of 2 and 3 steps (the queue's
first steps)."""
import numpy as np
import pytest

import native
from opencal_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("steps,rpw,per_launch,bpc", [(150, "32", "0", "0"), (151, "8", "0", "3"), (151, "16", "1", "0"),
                                                     (150, "8", "7", "1"), (2, "16", "0", "0"), (3, "32", "1", "0")])
def test_batch_vs_oracle(steps, rpw, per_launch, bpc, monkeypatch):
    """Uneven sizes, including grids shorter than a chunk and narrower than
    a strip, over an even and an odd number of steps; all steps in one
    cooperative launch (two or three blocks per CU), launches of 7 steps
    (one block per CU), or one launch per step."""
    from opencal_amd.models import sciddicat_batch
    monkeypatch.setenv("CALHIP_BATCH_ROWS_PER_WAVE", rpw)
    monkeypatch.setenv("CALHIP_BATCH_STEPS_PER_LAUNCH", per_launch)
    monkeypatch.setenv("CALHIP_BATCH_BLOCKS_PER_CU", bpc)
    specs = synth.batch_instances(12, seed=77, lo=24, hi=180)
    specs += [(6, 40, 0.5, 0.002, 901), (20, 6, 0.45, 0.001, 902), (9, 8, 0.6, 0.004, 903), (8, 130, 0.5, 0.001, 904)]
    insts = []
    for rows, cols, pr, pe, seed in specs:
        z, h, _ = synth.sciddicat_inputs(rows, cols, seed, patch=min(12, rows // 4 or 1))
        insts.append((z, h, pr, pe))
    print("case batch steps", steps, "rows", rpw, "per launch", per_launch, "blocks/CU", bpc,
          "sizes", [(r, c) for r, c, *_ in specs], flush=True)
    outs, ms = sciddicat_batch(insts, steps)
    assert ms > 0
    for (z, h, pr, pe), got in zip(insts, outs):
        _, want = native.oracle_sciddicat(z, h, pr, pe, steps, False)
        assert np.array_equal(got, want)


def test_batch_vs_single_runs():
    from opencal_amd.models import sciddicat_batch, sciddicat_run
    specs = [(37, 41, 0.5, 0.001, 1), (130, 257, 0.33, 0.006, 2), (96, 128, 0.69, 0.0002, 3), (5, 3, 0.5, 0.001, 4)]
    insts = []
    for rows, cols, pr, pe, seed in specs:
        z, h, _ = synth.sciddicat_inputs(rows, cols, seed, patch=min(10, rows // 4 or 1))
        insts.append((z, h, pr, pe))
    outs, _ = sciddicat_batch(insts, 200)
    for (z, h, pr, pe), got in zip(insts, outs):
        _, single, _ = sciddicat_run(z, h, pr, pe, 200, False)
        assert np.array_equal(got, single)


def test_batch_zero_steps_returns_input():
    from opencal_amd.models import sciddicat_batch
    z, h, _ = synth.sciddicat_inputs(40, 50, 3, patch=8)
    outs, _ = sciddicat_batch([(z, h, 0.5, 0.001)], 0)
    assert np.array_equal(outs[0], h)


def test_batch_special_values():
    """Instances with infinite altitudes (a +inf wall below the source, a
    -inf sink beside it) and rows of negative-zero thickness beside plain
    instances: the batch's minimisation takes the select-and-add form in the
    waves that meet an infinite level, and each instance still equals the
    oracle bit for bit (zero signs and infinities included)."""
    from opencal_amd.models import sciddicat_batch
    insts = []
    for k, (rows, cols) in enumerate([(64, 96), (40, 130), (72, 72), (33, 47)]):
        z, h, (r0, c0, p) = synth.sciddicat_inputs(rows, cols, 500 + k, patch=10)
        if k != 3:
            z[min(r0 + p + 2, rows - 1), max(c0 - 8, 0):min(c0 + p + 8, cols)] = np.inf
            z[r0 + p // 2, (c0 + p + 2) % cols] = -np.inf
            h[min(r0 + p + 5, rows - 1), :] = -0.0
            h[:, (c0 + p + 4) % cols] = -0.0
        insts.append((z, h, 0.5, 0.001))
    print("case batch special values", [z.shape for z, *_ in insts], flush=True)
    for steps in (1, 40):
        outs, _ = sciddicat_batch(insts, steps)
        for (z, h, pr, pe), got in zip(insts, outs):
            _, want = native.oracle_sciddicat(z, h, pr, pe, steps, False)
            assert np.array_equal(got.view(np.int64), want.view(np.int64)), (z.shape, steps)
            assert not np.isnan(got).any()
