"""SciddicaT with active cells in one launch per K steps (K = 1 for each
calRunCAStep2D, up to 4 for the steps calhipRunSteps2D hands over together),
checked bit for bit against the oracle (thickness, zero signs included, and
the set's size) and, for the flows, against the engine's two-launch path.

Cases: tile-ragged grid shapes (partial tiles whose rows or columns limit the
steps per launch), step counts that are not multiples of the launch's steps,
r >= 1 (cells that leave the set while wet keep non-zero flows: ghosts), sets
that empty part-way through a launch (every later step then covers the grid
and turns -0.0 thickness into +0.0), runs interleaved with host reads, and
an entry check that refuses the mode (eps < 0: dry cells outside the set with
h > eps) and falls back to the two-launch path."""
import os

import numpy as np
import pytest

import native
from opencal_amd import synth

pytestmark = pytest.mark.gpu

NO_FUSED = "CALHIP_SCIDDICAT_NO_FUSED_ACTIVE"


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.int64), np.asarray(b).view(np.int64))


def run_plan(z, h, p_r, p_eps, plan, fused=True, flows=False):
    """plan: ("run", n) = calhipRunSteps2D(n), ("step", n) = n calRunCAStep2D.
    Returns (h, count, launches, flows or None)."""
    from opencal_amd.models import SciddicaT
    total = sum(n for _, n in plan)
    if fused:
        os.environ.pop(NO_FUSED, None)
    else:
        os.environ[NO_FUSED] = "1"
    m = SciddicaT(z, h, p_r, p_eps, total + 8, active=True)
    try:
        m.init()
        m.clear_stop()
        for kind, n in plan:
            if kind == "run":
                m.run_steps(n)
            else:
                m.steps(n)
        launches = m.stats()["launches"]
        _, gh, cnt = m.read()
        fl = [m.read_flow(k) for k in range(4)] if flows else None
        return gh, cnt, launches, fl
    finally:
        m.close()
        os.environ.pop(NO_FUSED, None)


@pytest.mark.parametrize("rows,cols,steps,seed", [
    (64, 64, 61, 1),      # whole tiles, K = 4
    (37, 53, 45, 2),      # partial tiles of 5 rows and 5 columns
    (100, 70, 33, 3),     # 4 and 6
    (17, 17, 40, 4),      # two tile rows and columns, the second of one cell
    (48, 35, 50, 5),      # last tile column of 3 cells: K <= 2... 3 -> 2
    (130, 260, 90, 6),    # last tile row of 2 rows: K <= 2
    (5, 7, 30, 7),        # one tile, region larger than the grid
    (3, 3, 12, 8),
])
def test_fused_active_vs_oracle(rows, cols, steps, seed):
    print(f"case fused active {rows}x{cols} steps={steps} seed={seed}", flush=True)
    z, h, _ = synth.sciddicat_inputs(rows, cols, seed, patch=max(1, min(12, rows // 2)))
    _, oh, counts = native.oracle_sciddicat(z, h, 0.5, 0.001, steps, True, counts=True)
    for plan in ([("run", steps)], [("step", steps)], [("run", 7), ("step", 3), ("run", steps - 10)]):
        gh, cnt, launches, _ = run_plan(z, h, 0.5, 0.001, plan)
        assert same(gh, oh), f"plan={plan} max |dh| = {np.abs(gh - oh).max()}"
        assert cnt == counts[-1], f"plan={plan}"


def test_fused_active_launch_counts():
    """calhipRunSteps2D at 64 x 64: the first step uploads the model (one
    launch), the entry check and setup add two, then 4 steps per launch."""
    z, h, _ = synth.sciddicat_inputs(64, 64, 9, patch=10)
    _, _, fused_launches, _ = run_plan(z, h, 0.5, 0.001, [("run", 41)])
    _, _, list_launches, _ = run_plan(z, h, 0.5, 0.001, [("run", 41)], fused=False)
    print(f"launches fused {fused_launches}, two-launch path {list_launches}", flush=True)
    assert fused_launches <= 2 + 1 + 10 + 3   # entry check and setup, first step, 40 steps / 4, leave, flow reset
    assert list_launches >= 2 * 41


@pytest.mark.parametrize("p_r", [1.0, 1.25, 2.5])
def test_fused_active_ghost_flows(p_r):
    """r >= 1: a wet cell can leave the set in one step and keep the flows it
    wrote, which its neighbours then read in later steps."""
    print(f"case ghosts p_r={p_r}", flush=True)
    z, h, _ = synth.sciddicat_inputs(80, 96, 11, patch=10)
    h[h > 0] = 0.001 + (h[h > 0] - 1.0) * 2e-4     # a thin film: cells drop below eps often
    steps = 37
    _, oh, counts = native.oracle_sciddicat(z, h, p_r, 0.001, steps, True, counts=True)
    ref = run_plan(z, h, p_r, 0.001, [("run", steps)], fused=False, flows=True)
    assert same(ref[0], oh)
    for plan in ([("run", steps)], [("step", steps)], [("run", 9), ("step", 2), ("run", 26)]):
        gh, cnt, _, fl = run_plan(z, h, p_r, 0.001, plan, flows=True)
        assert same(gh, oh), f"plan={plan} max |dh| = {np.abs(gh - oh).max()}"
        assert cnt == counts[-1]
        for k in range(4):
            assert same(fl[k], ref[3][k]), f"plan={plan} flow {k}"
    # the film really left ghosts behind: some cell outside the set holds flows
    assert any(np.any(f != 0.0) for f in ref[3])


def _emptying_cases():
    """Seeded small scenarios (24 x 32, a few thin cells, r in [0.8, 4]) whose
    set empties after step 2 or later, one per emptying step."""
    rows, cols = 24, 32
    rng = np.random.default_rng(1)
    z0 = synth.dem(rows, cols, 3)
    found = {}
    for trial in range(4000):
        z = z0 if trial % 2 else np.zeros((rows, cols)) + rng.random((rows, cols)) * 0.002
        r = float(rng.uniform(0.8, 4.0))
        h = np.zeros((rows, cols))
        for _ in range(int(rng.integers(1, 6))):
            h[rng.integers(0, rows), rng.integers(0, cols)] = rng.uniform(0.0005, 0.004)
        h[rows - 3, :] = -0.0
        _, _, cnt = native.oracle_sciddicat(z, h, r, 0.001, 12, True, counts=True)
        c = cnt.tolist()
        if 0 in c and c.index(0) >= 1 and c.index(0) not in found:
            found[c.index(0)] = (z.copy(), h.copy(), r, c.index(0) + 1)
        if len(found) >= 5:
            break
    return [found[k] for k in sorted(found)]


def test_fused_active_set_empties():
    """The set empties after step e (OpenCAL's list becomes NULL): the steering
    resets every flow and each later step adds 0.0 to every thickness.  Runs
    that end before, at and after e, in launches of 4, 2 and 1 steps."""
    cases = _emptying_cases()
    assert len(cases) >= 3
    for z, h, r, e in cases:
        for steps in sorted({max(1, e - 1), e, e + 1, e + 3, 12}):
            print(f"case empties after step {e}, r={r:.3f}, steps={steps}", flush=True)
            _, oh, counts = native.oracle_sciddicat(z, h, r, 0.001, steps, True, counts=True)
            ref = run_plan(z, h, r, 0.001, [("run", steps)], fused=False, flows=True)
            assert same(ref[0], oh)
            for plan in ([("run", steps)], [("step", steps)], [("step", 1), ("run", steps - 1)]):
                plan = [p for p in plan if p[1] > 0]
                gh, cnt, _, fl = run_plan(z, h, r, 0.001, plan, flows=True)
                assert same(gh, oh), f"plan={plan}"
                assert cnt == counts[-1]
                for k in range(4):
                    assert same(fl[k], ref[3][k]), f"plan={plan} flow {k}"
            if steps > e:
                assert not (np.signbit(oh) & (oh == 0.0)).any()   # every -0.0 became +0.0


def test_fused_active_reads_between_runs():
    """Host reads of the thickness (one layer, no full download) and of the
    whole state between batched runs, then more steps."""
    from opencal_amd.models import SciddicaT
    z, h, _ = synth.sciddicat_inputs(96, 128, 12, patch=12)
    m = SciddicaT(z, h, 0.5, 0.001, 100, active=True)
    try:
        m.init()
        m.clear_stop()
        done = 0
        for n in (1, 10, 3, 17, 8):
            m.run_steps(n)
            done += n
            _, gh, cnt = m.read()
            _, oh, counts = native.oracle_sciddicat(z, h, 0.5, 0.001, done, True, counts=True)
            assert same(gh, oh), f"after {done} steps"
            assert cnt == counts[-1]
    finally:
        m.close()


def test_fused_active_refused_entry():
    """eps < 0: dry cells (h = 0 > eps) outside the set; the entry check
    refuses the mode and the steps take the two-launch path, still exact."""
    z, h, _ = synth.sciddicat_inputs(40, 64, 13, patch=8)
    _, oh, counts = native.oracle_sciddicat(z, h, 0.5, -1e-3, 20, True, counts=True)
    gh, cnt, launches, _ = run_plan(z, h, 0.5, -1e-3, [("run", 20)])
    assert same(gh, oh)
    assert cnt == counts[-1]
    assert launches >= 2 * 19
