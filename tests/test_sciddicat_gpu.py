"""SciddicaT on the GPU through the C ABI, checked against the oracle.

Bar: bit-exact (FP64 with the reference's operation order, no contraction).
Cases: seeded synthetic DEMs of several (ragged) sizes, full grid and active
cells, the reference's shipped 610x496x4000 result, the transient-flow
bookkeeping, and mass conservation at the 4096^2 benchmark size."""
import gzip
import os

import numpy as np
import pytest

import native
from opencal_amd import synth

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def gpu_run(z, h, p_r, p_eps, steps, active):
    from opencal_amd.models import sciddicat_run
    return sciddicat_run(z, h, p_r, p_eps, steps, active)


def test_division_equals_ieee_quotient():
    """The kernels divide the level sum by the count of remaining levels
    (sciddicaT.c:72-73) with a multiply and two FMAs; on the device, for 2^26
    seeded doubles (raw bit patterns: subnormals, infinities, NaNs, and values
    of an altitude's size) and every count 0..5, the bits equal the IEEE
    quotient's."""
    from opencal_amd import lib
    print("case division check 2^26 x 6", flush=True)
    bad = lib().calhipSciddicaTCheckDivision(1 << 26, 0x5eed)
    assert bad == 0, f"{bad} quotients differ"


@pytest.mark.parametrize("active", [False, True])
@pytest.mark.parametrize("rows,cols,steps,seed", [
    (64, 96, 300, 1), (97, 131, 200, 2), (128, 61, 250, 3), (200, 256, 150, 4), (3, 5, 20, 5),
])
def test_bit_exact_vs_oracle(rows, cols, steps, seed, active):
    print(f"case rows={rows} cols={cols} steps={steps} seed={seed} active={active}", flush=True)
    z, h, _ = synth.sciddicat_inputs(rows, cols, seed, patch=min(16, rows // 2 or 1))
    gz, gh, gcount = gpu_run(z, h, 0.5, 0.001, steps, active)
    oz, oh, ocount = native.oracle_sciddicat(z, h, 0.5, 0.001, steps, active, counts=True)
    assert np.array_equal(gz, oz)
    assert np.array_equal(gh, oh), f"max |dh| = {np.abs(gh - oh).max()}"
    if active:
        assert gcount == ocount[-1]


@pytest.mark.parametrize("rows,cols", [(7, 4), (33, 130), (40, 260), (21, 254), (50, 384), (19, 132)])
def test_strip_edges(rows, cols):
    """Even column counts run in 128-column strips whose first and last
    cells take their horizontal terms from separately loaded halo columns:
    wet cells straddle every strip boundary and the toroidal wrap, and the
    last strip holds 2, 4, 126 or 128 columns."""
    print(f"case strip edges rows={rows} cols={cols}", flush=True)
    z, h, _ = synth.sciddicat_inputs(rows, cols, 31, patch=1)
    h[:] = 0.0
    s = synth.Stream(rows * 1000 + cols)
    for j in list(range(0, cols, 128)) + [cols - 1]:
        for dj in (-3, -2, -1, 0, 1, 2):
            jj = (j + dj) % cols
            h[:, jj] = s.between(0.5, 4.0, rows)
    h[0, :] = s.between(0.0, 2.0, cols)
    for active in (False, True):
        gz, gh, _ = gpu_run(z, h, 0.5, 0.001, 60, active)
        oz, oh = native.oracle_sciddicat(z, h, 0.5, 0.001, 60, active)
        assert np.array_equal(gh, oh), f"active={active} max |dh| = {np.abs(gh - oh).max()}"


@pytest.mark.parametrize("p_r,p_eps", [(0.3, 1e-4), (0.7, 1e-2), (0.61, 0.0031)])
def test_parameters(p_r, p_eps):
    z, h, _ = synth.sciddicat_inputs(80, 72, 9, patch=12)
    for active in (False, True):
        gz, gh, _ = gpu_run(z, h, p_r, p_eps, 120, active)
        oz, oh = native.oracle_sciddicat(z, h, p_r, p_eps, 120, active)
        assert np.array_equal(gh, oh)


def test_reference_width_final_full_grid():
    """OpenCAL-OMP/examples/calomp_sciddicaT/data/width_final.txt, 4000 steps,
    written with "%.9f" exactly as the reference saves it."""
    z = synth.read_matrix_gz(os.path.join(GOLDEN, "sciddicaT_610x496_dem.txt.gz"))
    h = synth.read_matrix_gz(os.path.join(GOLDEN, "sciddicaT_610x496_source.txt.gz"))
    _, gh, _ = gpu_run(z, h, 0.5, 0.001, 4000, False)
    text = gzip.open(os.path.join(GOLDEN, "sciddicaT_610x496_width_final.txt.gz"), "rt").read()
    mine = "".join("".join("%.9f  " % v for v in row) + "\n" for row in gh)
    assert mine == text


def test_reference_width_final_active():
    """cal_sciddicaT-activecells-glut/data/width_final.txt ("%f"), 4000 steps."""
    z = synth.read_matrix_gz(os.path.join(GOLDEN, "sciddicaT_610x496_dem.txt.gz"))
    h = synth.read_matrix_gz(os.path.join(GOLDEN, "sciddicaT_610x496_source.txt.gz"))
    _, gh, _ = gpu_run(z, h, 0.5, 0.001, 4000, True)
    ref = synth.read_matrix_gz(os.path.join(GOLDEN, "sciddicaT_active_610x496_width_final.txt.gz"))
    assert np.array_equal(np.round(gh, 6), ref) or np.abs(gh - ref).max() <= 5e-7


def test_flows_observable_after_step():
    """Flows are not stored by the fused sweep; reading them writes them out
    and must give the reference's values (the steering has reset them: 0)."""
    from opencal_amd.models import SciddicaT
    z, h, _ = synth.sciddicat_inputs(64, 64, 21, patch=12)
    m = SciddicaT(z, h, 0.5, 0.001, 10, False)
    m.init()
    m.steps(10)
    f0 = m.read_flow(0)
    assert np.all(f0 == 0.0)
    _, gh, _ = m.read()
    _, oh = native.oracle_sciddicat(z, h, 0.5, 0.001, 10, False)
    assert np.array_equal(gh, oh)
    # continue after the host round trip: state re-uploaded, still exact
    m.steps(15)
    _, gh2, _ = m.read()
    _, oh2 = native.oracle_sciddicat(z, h, 0.5, 0.001, 25, False)
    assert np.array_equal(gh2, oh2)
    st = m.stats()
    assert st["uploads"] == 2 and st["downloads"] >= 2
    m.close()


@pytest.mark.parametrize("split", [1, 7, 30])
def test_active_round_trips(split):
    """Active cells stepped on the flag grid: the row-major list is rebuilt
    only when observed.  Reading the state (list, count, substates) in the
    middle of the run and continuing gives the oracle's result and counts."""
    from opencal_amd.models import SciddicaT
    z, h, _ = synth.sciddicat_inputs(90, 128, 44, patch=10)
    m = SciddicaT(z, h, 0.5, 0.001, 60, True)
    m.init()
    _, oh_all, counts = native.oracle_sciddicat(z, h, 0.5, 0.001, 60, True, counts=True)
    m.steps(split)
    _, gh, cnt = m.read()
    _, oh = native.oracle_sciddicat(z, h, 0.5, 0.001, split, True)
    assert np.array_equal(gh, oh)
    assert cnt == counts[split - 1]
    m.steps(60 - split)
    _, gh, cnt = m.read()
    assert np.array_equal(gh, oh_all)
    assert cnt == counts[-1]
    m.close()


@pytest.mark.parametrize("edit", ["dry_far_cell", "already_active_cell"])
def test_active_host_edit_mid_run(edit):
    """The host adds a cell to the active set between steps (calAddActiveCell2D
    on the host mirror).  A dry cell far from the flow (the first case) is
    visited by one step's processes, changes nothing and is removed again; a
    cell already in the set (the second) changes nothing.  Either way the run
    continues on the device from the re-uploaded state (the first case through
    the list path for one step, then the tile lists rebuilt from the flags)
    and ends with the oracle's unedited result and set size."""
    import ctypes as C
    from opencal_amd import lib
    from opencal_amd.models import SciddicaT
    rows, cols, steps, k = 90, 200, 70, 25
    z, h, (r0, c0, p) = synth.sciddicat_inputs(rows, cols, 45, patch=10)
    _, oh, counts = native.oracle_sciddicat(z, h, 0.5, 0.001, steps, True, counts=True)
    _, oh_k = native.oracle_sciddicat(z, h, 0.5, 0.001, k, True)
    if edit == "dry_far_cell":
        # the cell farthest from the source patch whose 3 x 3 block stays dry
        ii, jj = np.meshgrid(np.arange(rows), np.arange(cols), indexing="ij")
        d = np.abs(ii - (r0 + p // 2)) + np.minimum(np.abs(jj - (c0 + p // 2)), cols - np.abs(jj - (c0 + p // 2)))
        dry = np.ones_like(oh, bool)
        for di in (-1, 0, 1):
            for dj in (-1, 0, 1):
                dry &= np.roll(np.roll(oh, di, 0), dj, 1) == 0.0
        i, j = np.unravel_index(np.argmax(np.where(dry, d, -1)), d.shape)
        assert dry[i, j]
    else:
        wet = np.argwhere(oh_k > 0.001)
        i, j = wet[len(wet) // 2]
    print(f"case host edit {edit} at ({i}, {j}) after {k} steps", flush=True)
    L = lib()
    L.calAddActiveCell2D.argtypes = [C.c_void_p, C.c_int, C.c_int]
    m = SciddicaT(z, h, 0.5, 0.001, steps, True)
    m.init()
    m.steps(k)
    L.calAddActiveCell2D(m.model, int(i), int(j))
    m.steps(steps - k)
    _, gh, cnt = m.read()
    assert np.array_equal(gh, oh), f"max |dh| = {np.abs(gh - oh).max()}"
    assert cnt == counts[-1]
    assert m.stats()["uploads"] >= 2
    m.close()


@pytest.mark.parametrize("active", [False, True])
def test_special_values(active):
    """Infinite altitudes (a +inf wall below the source, a -inf sink beside
    it) and negative-zero thickness.  A wave with an infinite level takes the
    minimisation's select-and-add form instead of the FMA form, an infinite
    level sum meets the division's non-finite case, and rows of -0 that stay
    dry are written as +0 by the update (the two-step sweep's dry-row shortcut
    does the same).  Bitwise against the oracle (zero signs and infinities
    included), through calRun2D and through per-step launches."""
    from opencal_amd.models import SciddicaT
    z, h, (r0, c0, p) = synth.sciddicat_inputs(72, 136, 17, patch=12)
    rows, cols = z.shape
    z[r0 + p + 2, max(c0 - 10, 0):min(c0 + p + 10, cols)] = np.inf
    z[r0 + p // 2, (c0 + p + 2) % cols] = -np.inf
    h[r0 + p + 6:r0 + p + 12, :] = -0.0
    h[:, (c0 + p + 5) % cols] = -0.0
    print(f"case special values active={active} patch=({r0}, {c0}, {p})", flush=True)

    def same(a, b):
        return np.array_equal(a.view(np.int64), b.view(np.int64))

    for steps in (1, 2, 41):
        gz, gh, gc = gpu_run(z, h, 0.5, 0.001, steps, active)
        oz, oh, oc = native.oracle_sciddicat(z, h, 0.5, 0.001, steps, active, counts=True)
        assert same(gz, oz) and same(gh, oh), f"steps={steps}"
        assert not np.isnan(gh).any()
        if active:
            assert gc == oc[-1]
    m = SciddicaT(z, h, 0.5, 0.001, 41, active)
    try:
        m.init()
        m.steps(41)
        _, gh, _ = m.read()
    finally:
        m.close()
    _, oh = native.oracle_sciddicat(z, h, 0.5, 0.001, 41, active)
    assert same(gh, oh)
    assert (np.signbit(gh) & (gh == 0.0)).any() == active   # -0 kept where no process visits


def test_mass_conservation_4096():
    """At the benchmark size: total thickness is conserved on the torus (to
    rounding), and the first 3 steps match the oracle bit for bit."""
    from opencal_amd.models import SciddicaT
    z, h, _ = synth.sciddicat_inputs(4096, 4096, 2024)
    m = SciddicaT(z, h, 0.5, 0.001, 200, False)
    m.init()
    m.steps(3)
    _, gh3, _ = m.read()
    _, oh3 = native.oracle_sciddicat(z, h, 0.5, 0.001, 3, False)
    assert np.array_equal(gh3, oh3)
    m.steps(197)
    _, gh, _ = m.read()
    assert abs(gh.sum() - h.sum()) <= 1e-9 * h.sum()
    assert gh.min() >= 0.0
    m.close()


@pytest.mark.parametrize("rows,cols,s_rows,group", [(8, 8, 16, 4), (9, 10, 16, 1), (64, 128, 16, 4), (70, 260, 24, 3),
                                                    (130, 384, 48, 1), (41, 136, 8, 8), (300, 200, 16, 2),
                                                    (257, 512, 16, 4)])
def test_two_step_sweep(rows, cols, s_rows, group, monkeypatch):
    """calRun2D / calhipRunSteps2D with the steering twin advance the full
    grid two steps per sweep (h1 kept in registers), in big tiles of `group`
    small tiles of `s_rows` rows with the small tiles that met wet cells run
    first on their own: bit-exact against the oracle for even and odd step
    counts, tiles taller than the grid, grids of 8 columns (every halo column
    wraps) and strips of 2..128 columns, wet cells on every strip edge, and
    when mixed with single steps."""
    from opencal_amd.models import SciddicaT
    print(f"case two-step rows={rows} cols={cols} s_rows={s_rows} group={group}", flush=True)
    monkeypatch.setenv("CALHIP_SCIDDICAT_ROWS_PER_WAVE_2X", str(s_rows))
    monkeypatch.setenv("CALHIP_SCIDDICAT_2X_GROUP", str(group))
    z, h, _ = synth.sciddicat_inputs(rows, cols, rows + cols, patch=min(8, rows // 2))
    s = synth.Stream(rows * 7 + cols)
    for j in list(range(0, cols, 128)) + [cols - 1]:
        for dj in (-4, -3, -2, -1, 0, 1, 2, 3):
            h[:, (j + dj) % cols] = s.between(0.0, 3.0, rows)
    for steps in (1, 2, 37):
        gz, gh, _ = gpu_run(z, h, 0.5, 0.001, steps, False)
        oz, oh = native.oracle_sciddicat(z, h, 0.5, 0.001, steps, False)
        assert np.array_equal(gh, oh), f"steps={steps} max |dh| = {np.abs(gh - oh).max()}"
    # single steps, a batch, single steps
    m = SciddicaT(z, h, 0.5, 0.001, 40, active=False)
    try:
        m.init()
        m.steps(3)
        m.run_steps(10)
        m.steps(1)
        m.run_steps(7)
        _, gh, _ = m.read()
        st = m.stats()
    finally:
        m.close()
    _, oh = native.oracle_sciddicat(z, h, 0.5, 0.001, 21, False)
    assert np.array_equal(gh, oh), f"mixed: max |dh| = {np.abs(gh - oh).max()}"
    assert st["steps"] == 21


@pytest.mark.parametrize("rows,cols", [(1024, 1024), (300, 384)])
def test_run_steps_tuned_choice(rows, cols, monkeypatch):
    """calhipRunSteps2D with the tuner's own choice for the grid (a two-step
    tile shape, or one step per launch when that is faster): bit-exact
    against the oracle."""
    from opencal_amd.models import SciddicaT
    print(f"case run_steps tuned choice {rows}x{cols}", flush=True)
    monkeypatch.delenv("CALHIP_SCIDDICAT_ROWS_PER_WAVE_2X", raising=False)
    monkeypatch.delenv("CALHIP_SCIDDICAT_2X_GROUP", raising=False)
    z, h, _ = synth.sciddicat_inputs(rows, cols, rows + 3 * cols)
    m = SciddicaT(z, h, 0.5, 0.001, 64, active=False)
    try:
        m.init()
        m.steps(1)
        m.run_steps(63)
        _, gh, _ = m.read()
    finally:
        m.close()
    _, oh = native.oracle_sciddicat(z, h, 0.5, 0.001, 64, False)
    assert np.array_equal(gh, oh), f"max |dh| = {np.abs(gh - oh).max()}"
