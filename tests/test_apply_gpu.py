"""calApplyElementaryProcess2D on the hand-written device programs
(OpenCAL-OMP/source/cal2D.c:717-734): a model whose own global transition
runs the bound processes one at a time, each followed by calUpdate2D or by
calUpdateSubstate2D* of what it wrote (the pattern of
OpenCAL/examples/cal_sciddicaS3hex-unsafe-explicit-glut/source/
sciddicaS3Hex.c:218-229), driven by calRun2D's loop.

Pins: bit-exact against the oracle (which runs the reference sequence:
every process, then the update) for SciddicaT full grid and active cells,
including the final active-cell count, and for Life against the oracle's
Life, each under both update forms; the result equals the fused program's
own run."""
import numpy as np
import pytest

import native
from opencal_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("active,mode", [(False, 1), (False, 2), (True, 1)])
def test_sciddicat_processes_applied_one_at_a_time(active, mode):
    from opencal_amd.models import SciddicaT, sciddicat_run
    print(f"case apply sciddicaT active={active} mode={mode}", flush=True)
    z, h, _ = synth.sciddicat_inputs(96, 130, 31, patch=20)
    steps = 60
    m = SciddicaT(z, h, 0.5, 0.001, steps, active)
    try:
        m.set_explicit(mode)
        m.run()
        gz, gh, cnt = m.read()
        st = m.stats()
    finally:
        m.close()
    oz, oh, counts = native.oracle_sciddicat(z, h, 0.5, 0.001, steps, active, counts=True)
    assert np.array_equal(gz, oz)
    assert np.array_equal(gh, oh), f"max |dh| = {np.abs(gh - oh).max()}"
    if active:
        assert cnt == counts[steps - 1]
    assert st["steps"] == 0   # no step of the fused program ran
    fz, fh, fc = sciddicat_run(z, h, 0.5, 0.001, steps, active)
    assert np.array_equal(fh, gh) and fc == cnt


@pytest.mark.parametrize("mode", [1, 2])
def test_life_transition_applied_alone(mode):
    from opencal_amd.models import Life
    print(f"case apply life mode={mode}", flush=True)
    a0 = synth.life_inputs(70, 128, 9)
    m = Life(a0, 50)
    try:
        m.set_explicit(mode)
        m.run()
        got = m.read()
    finally:
        m.close()
    assert np.array_equal(got, native.oracle_life(a0, 50))
