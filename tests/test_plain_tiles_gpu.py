"""SciddicaT on the full grid (CAL_NO_OPT) in launches of up to 4 steps: the
tiles near wet cells advance in LDS, every other tile is copied (h + 0.0)
until two consecutive launches have copied it, after which both layers hold
its values and it is left alone.  Checked bit for bit against the oracle
(zero signs included) through calRun2D's loop, one calRunCAStep2D per step
(grouped four to a launch), mixtures and host edits between runs, on
tile-ragged grid shapes, with special values, and with the mode forced on
(CALHIP_SCIDDICAT_TILES_ALWAYS) where the wet area would otherwise make the
engine keep the strip sweeps."""
import os

import numpy as np
import pytest

import native
from opencal_amd import synth

pytestmark = pytest.mark.gpu


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.int64), np.asarray(b).view(np.int64))


def run_plan(z, h, p_r, p_eps, plan, env=None):
    """plan: ("run", n) for n steps of calRun2D's loop, ("step", n) for n
    calRunCAStep2D calls, ("set", [(i, j, v), ...]) for host edits of h."""
    from opencal_amd.models import SciddicaT
    total = sum(n for k, n in plan if k != "set")
    old = {k: os.environ.get(k) for k in (env or {})}
    os.environ.update(env or {})
    m = SciddicaT(z, h, p_r, p_eps, total + 8, active=False)
    try:
        m.init()
        for kind, n in plan:
            if kind == "run":
                m.run_steps(n)
            elif kind == "set":
                for i, j, v in n:
                    m.set_h(i, j, v)
            else:
                m.steps(n)
        launches = m.stats()["launches"]
        _, gh, _ = m.read()
        flows = [m.read_flow(k) for k in range(4)]
        return gh, launches, flows
    finally:
        m.close()
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


ALWAYS = {"CALHIP_SCIDDICAT_TILES_ALWAYS": "1"}


@pytest.mark.parametrize("rows,cols,steps,seed", [
    (64, 64, 61, 1),      # whole tiles: 4 steps per launch
    (37, 53, 45, 2),      # last tile row and column of 5 cells: up to 4 steps per launch
    (100, 70, 33, 3),
    (17, 17, 40, 4),      # two tile rows and columns, the second of one cell
    (48, 35, 50, 5),      # last tile column of 3 cells: up to 2 steps per launch
    (130, 260, 90, 6),    # last tile row of 2 rows
    (5, 7, 30, 7),        # one tile smaller than its halo
    (3, 3, 12, 8),
    (300, 200, 100, 9),
])
def test_plain_tiles_vs_oracle(rows, cols, steps, seed):
    print(f"case full-grid tiles {rows}x{cols} steps={steps} seed={seed}", flush=True)
    z, h, _ = synth.sciddicat_inputs(rows, cols, seed, patch=max(1, min(12, rows // 2)))
    _, oh = native.oracle_sciddicat(z, h, 0.5, 0.001, steps, False)
    for plan in ([("run", steps)], [("step", steps)], [("run", 7), ("step", 3), ("run", steps - 10)]):
        plan = [p for p in plan if p[1] > 0]
        gh, _, fl = run_plan(z, h, 0.5, 0.001, plan, ALWAYS)
        assert same(gh, oh), f"plan={plan} max |dh| = {np.abs(gh - oh).max()}"
        assert all(np.all(f == 0.0) and not np.signbit(f).any() for f in fl)   # the steering reset them


def test_plain_tiles_launches():
    """At 256 x 256 with a 12 x 12 source the mode turns on by itself and
    advances four steps per launch."""
    z, h, _ = synth.sciddicat_inputs(256, 256, 10, patch=12)
    _, oh = native.oracle_sciddicat(z, h, 0.5, 0.001, 81, False)
    gh, launches, _ = run_plan(z, h, 0.5, 0.001, [("run", 81)])
    print(f"launches {launches}", flush=True)
    assert same(gh, oh)
    assert launches <= 1 + 20 + 3   # entry list, 80 steps in 20 launches, the first step's own


@pytest.mark.parametrize("p_r,p_eps", [(0.3, 0.0), (0.9, 1e-2), (1.0, 1e-3)])
def test_plain_tiles_parameters(p_r, p_eps):
    z, h, _ = synth.sciddicat_inputs(96, 112, 12, patch=10)
    _, oh = native.oracle_sciddicat(z, h, p_r, p_eps, 57, False)
    gh, _, _ = run_plan(z, h, p_r, p_eps, [("run", 57)], ALWAYS)
    assert same(gh, oh)


def test_plain_tiles_special_values():
    """Infinite altitudes and -0.0 thickness: dry rows of -0.0 become +0.0
    (copied tiles and computed tiles alike), as in the oracle."""
    z, h, (r0, c0, p) = synth.sciddicat_inputs(72, 136, 17, patch=12)
    rows, cols = z.shape
    z[r0 + p + 2, max(c0 - 10, 0):min(c0 + p + 10, cols)] = np.inf
    z[r0 + p // 2, (c0 + p + 2) % cols] = -np.inf
    h[r0 + p + 6:r0 + p + 12, :] = -0.0
    h[:, (c0 + p + 5) % cols] = -0.0
    for steps in (1, 2, 9, 41):
        _, oh = native.oracle_sciddicat(z, h, 0.5, 0.001, steps, False)
        for plan in ([("run", steps)], [("step", steps)]):
            gh, _, _ = run_plan(z, h, 0.5, 0.001, plan, ALWAYS)
            assert same(gh, oh), f"steps={steps} plan={plan}"


def test_plain_tiles_host_edits_between_launches():
    """Dry tiles copied by two consecutive launches are left alone until they
    are listed again (both layers already hold h + 0.0).  Host edits between
    runs (-0.0, a new wet patch, a dry value) must reach both layers of
    such tiles: the result equals the strip sweeps' (which the oracle pins)
    bit for bit, zero signs included."""
    z, h, (r0, c0, p) = synth.sciddicat_inputs(160, 224, 21, patch=10)
    far = [(r, c) for r in range(120, 126) for c in range(150, 170)]
    plan = [("run", 40),
            ("set", [(r, c, -0.0) for r, c in far[:40]] + [(r, c, 3.5) for r, c in far[40:60]]),
            ("run", 31), ("step", 5),
            ("set", [(r, c, 0.25) for r, c in far[60:80]] + [(r, c, -0.0) for r, c in far[80:100]]),
            ("run", 37)]
    print("case full-grid tiles 160x224 host edits between runs", flush=True)
    want, _, _ = run_plan(z, h, 0.5, 0.001, plan, {"CALHIP_SCIDDICAT_NO_TILES": "1"})
    got, launches, _ = run_plan(z, h, 0.5, 0.001, plan, ALWAYS)
    assert same(got, want)
    assert np.signbit(got).sum() == np.signbit(want).sum()
    # without edits: the same run equals the oracle
    _, oh = native.oracle_sciddicat(z, h, 0.5, 0.001, 113, False)
    gh, _, _ = run_plan(z, h, 0.5, 0.001, [("run", 76), ("step", 37)], ALWAYS)
    assert same(gh, oh)


@pytest.mark.parametrize("parts", ["0", "1"])
@pytest.mark.parametrize("rows,cols,steps,seed", [(64, 64, 61, 1), (37, 53, 45, 2), (17, 17, 40, 4),
                                                  (130, 260, 90, 6), (3, 3, 12, 8)])
def test_plain_tiles_whole_or_quarter(parts, rows, cols, steps, seed):
    """The listed tiles computed whole (one block per tile) or in four 8 x 8
    parts (one block each), forced either way: both equal the oracle."""
    print(f"case full-grid tiles parts={parts} {rows}x{cols} steps={steps} seed={seed}", flush=True)
    z, h, _ = synth.sciddicat_inputs(rows, cols, seed, patch=max(1, min(12, rows // 2)))
    h[h.shape[0] // 2 + 1, :] = -0.0
    _, oh = native.oracle_sciddicat(z, h, 0.5, 0.001, steps, False)
    env = dict(ALWAYS, CALHIP_SCIDDICAT_TILE_PARTS=parts)
    plans = [[("run", steps)]] + ([[("run", 9), ("step", 5), ("run", steps - 14)]] if steps > 14 else [])
    for plan in plans:
        gh, _, _ = run_plan(z, h, 0.5, 0.001, plan, env)
        assert same(gh, oh), f"plan={plan}"
