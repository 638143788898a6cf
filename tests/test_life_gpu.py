"""Life on the GPU through the C ABI: bit-exact against the oracle, numpy
and closed forms; per-step launches and the LDS-resident multi-step path."""
import numpy as np
import pytest

import native
from opencal_amd import synth

pytestmark = pytest.mark.gpu


def np_life(a, steps):
    a = a.astype(np.int32)
    for _ in range(steps):
        s = sum(np.roll(np.roll(a, di, 0), dj, 1) for di in (-1, 0, 1) for dj in (-1, 0, 1) if (di, dj) != (0, 0))
        a = ((s == 3) | ((s == 2) & (a == 1))).astype(np.int32)
    return a.astype(np.uint8)


@pytest.mark.parametrize("rows,cols,steps", [(512, 512, 1000), (64, 128, 333), (37, 53, 101), (256, 1024, 77)])
def test_life_vs_oracle(rows, cols, steps):
    from opencal_amd.models import life_run
    print(f"case life {rows}x{cols} steps={steps}", flush=True)
    a0 = synth.life_inputs(rows, cols, seed=rows * 7 + cols)
    got = life_run(a0, steps)
    want = native.oracle_life(a0, steps)
    assert np.array_equal(got, want)
    assert int(got.sum()) == int(want.sum())


@pytest.mark.parametrize("rows,cols,steps", [(64, 64, 200), (8, 192, 100), (100, 320, 66), (33, 128, 2),
                                             (70, 256, 33), (3, 64, 40)])
def test_life_tiles_edges(rows, cols, steps):
    """calRun2D's temporally blocked path: boards narrower than a tile or
    shorter than a region (the region wraps onto itself), step counts that
    leave a partial last launch, and a single multi-step step."""
    from opencal_amd.models import life_run
    print(f"case life tiles {rows}x{cols} steps={steps}", flush=True)
    a0 = synth.life_inputs(rows, cols, seed=rows * 11 + cols)
    assert np.array_equal(life_run(a0, steps), np_life(a0, steps))


def test_life_stepwise_equals_batched():
    from opencal_amd.models import Life
    a0 = synth.life_inputs(128, 128, seed=5)
    m = Life(a0, 50)
    m.steps(50)
    step_wise = m.read()
    m.close()
    from opencal_amd.models import life_run
    assert np.array_equal(step_wise, life_run(a0, 50))
    assert np.array_equal(step_wise, np_life(a0, 50))


def test_life_closed_forms():
    from opencal_amd.models import life_run
    g = np.zeros((64, 64), np.uint8)
    g[10, 10:13] = 1                      # blinker, period 2
    g[30:32, 30:32] = 1                   # block, still life
    assert np.array_equal(life_run(g, 2), g)
    assert np.array_equal(life_run(g, 1)[9:12, 11], [1, 1, 1])
    gl = np.zeros((64, 64), np.uint8)     # glider: shifts (+1, +1) every 4 steps on the torus
    gl[1, 2] = gl[2, 3] = gl[3, 1] = gl[3, 2] = gl[3, 3] = 1
    assert np.array_equal(life_run(gl, 4 * 64), gl)
    assert np.array_equal(life_run(gl, 4), np.roll(np.roll(gl, 1, 0), 1, 1))
