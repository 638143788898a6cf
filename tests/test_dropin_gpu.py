"""The reference's own SciddicaT program (OpenCAL-OMP/examples/calomp_sciddicaT/
source/sciddicaT.c: its main, data paths and output), built against this
engine with a two-line patch (oracle/ref/Makefile target
dropin_sciddicaT), runs on the GPU and writes the reference's shipped
width_final.txt byte for byte.  Skipped where that build is absent."""
import gzip
import os
import shutil
import subprocess

import pytest

import native

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BIN = os.path.join(native.REF_DIR, "dropin_sciddicaT")


@pytest.mark.skipif(not os.path.exists(BIN), reason="dropin_sciddicaT not built (needs the reference tree)")
def test_reference_program_runs_unchanged_on_the_engine(tmp_path):
    print("case dropin sciddicaT 610x496 4000 steps", flush=True)
    data = tmp_path / "data"
    data.mkdir()
    for name, src in (("dem.txt", "sciddicaT_610x496_dem.txt.gz"), ("source.txt", "sciddicaT_610x496_source.txt.gz")):
        with gzip.open(os.path.join(GOLDEN, src), "rb") as f, open(data / name, "wb") as g:
            shutil.copyfileobj(f, g)
    r = subprocess.run([BIN], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Simulation terminated" in r.stdout
    want = gzip.open(os.path.join(GOLDEN, "sciddicaT_610x496_width_final.txt.gz"), "rt").read()
    assert (data / "width_final.txt").read_text() == want


BIN_2X = os.path.join(native.REF_DIR, "dropin_sciddicaT_2x")


@pytest.mark.skipif(not os.path.exists(BIN_2X), reason="dropin_sciddicaT_2x not built (needs the reference tree)")
def test_reference_program_two_steps_per_sweep(tmp_path):
    """The reference program with a third added line (its steering declared
    as the flow reset): calRun2D advances two steps per sweep and the shipped
    width_final.txt is reproduced byte for byte."""
    print("case dropin sciddicaT 2x 610x496 4000 steps", flush=True)
    data = tmp_path / "data"
    data.mkdir()
    for name, src in (("dem.txt", "sciddicaT_610x496_dem.txt.gz"), ("source.txt", "sciddicaT_610x496_source.txt.gz")):
        with gzip.open(os.path.join(GOLDEN, src), "rb") as f, open(data / name, "wb") as g:
            shutil.copyfileobj(f, g)
    # the tile shape fixed, so that this grid takes the two-step sweep
    # whatever the tuner would choose for it
    env = dict(os.environ, CALHIP_SCIDDICAT_ROWS_PER_WAVE_2X="12", CALHIP_SCIDDICAT_2X_GROUP="4")
    r = subprocess.run([BIN_2X], cwd=tmp_path, capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Simulation terminated" in r.stdout
    want = gzip.open(os.path.join(GOLDEN, "sciddicaT_610x496_width_final.txt.gz"), "rt").read()
    assert (data / "width_final.txt").read_text() == want


BIN_G = os.path.join(native.REF_DIR, "dropin_sciddicaT_generic")


@pytest.mark.skipif(not os.path.exists(BIN_G), reason="dropin_sciddicaT_generic not built (needs the reference tree)")
def test_reference_program_on_the_generic_device_program(tmp_path):
    """The same program with its own elementary processes compiled for the
    GPU at run time (the file's own text, five bind calls): the shipped
    width_final.txt byte for byte."""
    print("case dropin sciddicaT generic 610x496 4000 steps", flush=True)
    data = tmp_path / "data"
    data.mkdir()
    for name, src in (("dem.txt", "sciddicaT_610x496_dem.txt.gz"), ("source.txt", "sciddicaT_610x496_source.txt.gz")):
        with gzip.open(os.path.join(GOLDEN, src), "rb") as f, open(data / name, "wb") as g:
            shutil.copyfileobj(f, g)
    r = subprocess.run([BIN_G], cwd=tmp_path, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Simulation terminated" in r.stdout, r.stdout + r.stderr
    want = gzip.open(os.path.join(GOLDEN, "sciddicaT_610x496_width_final.txt.gz"), "rt").read()
    assert (data / "width_final.txt").read_text() == want


BIN_SC = os.path.join(native.REF_DIR, "dropin_sciara")


@pytest.mark.skipif(not os.path.exists(BIN_SC), reason="dropin_sciara not built (needs the reference tree)")
def test_reference_sciara_program_on_the_engine(tmp_path):
    """The reference SCIARA-fv2 program (main.cpp + its sources, one added bind
    call) on its own 2006 data: 1000 steps, full grid, saved with "%.9f".  The
    GPU's files equal, as text, the oracle run with the same correctly
    rounded elementary functions, and lie within 1e-11 of each field's scale
    (plus the 5e-10 of the text format) of the glibc run, which equals the
    reference program on the CPU: of the 780 000 saved values, 6 differ from
    the reference program's in the ninth decimal (measured 2.2e-12 in h and
    1.0e-9 K in T at most)."""
    import numpy as np
    import sciara2006
    print("case dropin sciara 2006 data 378x517 1000 steps", flush=True)
    sciara2006.write_inputs(str(tmp_path))
    r = subprocess.run([BIN_SC], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    inp = sciara2006.load_inputs()
    port = native.oracle_sciara(inp, 1000, False, crmath=True)
    libm = native.oracle_sciara(inp, 1000, False, crmath=False)
    for name, k, scale in (("Thickness", "h", 1.0), ("Temperature", "T", 1400.0), ("Morphology", "z", 3400.0),
                           ("SolidifiedLavaThickness", "Msl", 1.0)):
        path = tmp_path / "data" / "2006_SAVE" / f"{sciara2006.PREFIX}_{name}.stt"
        text = path.read_text().split()[12:]
        assert text == ["%.9f" % v for v in port[k].ravel()], name
        got = np.array(text, dtype=np.float64)
        assert np.abs(got - libm[k].ravel()).max() <= 1e-11 * scale + 5e-10, name


BIN_MB = os.path.join(native.REF_DIR, "dropin_mbusu")
BIN_MB_REF = os.path.join(native.REF_DIR, "calomp_mbusu")
MB_NAMES = ["teta", "moist_cont", "psi", "k", "h", "dqdh", "convergence", "moist_diff"]


def _floats(text):
    return [float(t) for t in text.replace(",", " ").split()]


@pytest.mark.skipif(not os.path.exists(BIN_MB), reason="dropin_mbusu not built (needs the reference tree)")
def test_reference_mbusu_3d_on_the_generic_device_program(tmp_path):
    """The reference 3D example mbusu (OpenCAL-OMP/examples/calomp_mbusu-glut,
    160x30x130 cells, 8 substates, pow-heavy Richards-equation process),
    unchanged, its process compiled for the GPU at run time, run to its own
    stop condition (ten simulated days, 8662 steps).  deltat.txt equals the
    shipped file line for line.  ris10g.txt (the moisture section, "%f")
    against the reference built from its sources here
    (make_mbusu_fixture.py): within one unit of the last printed digit (the
    GPU's pow differs from glibc's by an ulp); against the file the reference
    ships from its OpenMP build: within 4e-5, as the reference's own build
    here (3.3e-5; its shipped serial and OpenMP files differ by 2.4e-5)."""
    import time
    print("case dropin mbusu 160x30x130 to the stop condition", flush=True)
    t0 = time.perf_counter()
    r = subprocess.run([BIN_MB], cwd=tmp_path, capture_output=True, text=True, timeout=400)
    wall = time.perf_counter() - t0
    assert r.returncode == 0, r.stdout + r.stderr
    print(f"mbusu: 8662 steps in {wall:.1f} s of wall time ({1e3 * wall / 8662:.2f} ms/step, "
          f"process start and compilation included)", flush=True)
    assert "steps 8662" in r.stdout, r.stdout + r.stderr
    want_dt = gzip.open(os.path.join(GOLDEN, "mbusu_deltat-orig-opencal-omp.txt.gz"), "rt").read().splitlines()
    got_dt = (tmp_path / "deltat.txt").read_text().splitlines()
    assert len(got_dt) == 8662 and got_dt == want_dt[:8662]
    got = _floats((tmp_path / "ris10g.txt").read_text())
    for fixture, tol in (("mbusu_ris10g_reference_build.txt.gz", 1.01e-6),
                         ("mbusu_ris10g-orig-opencal-omp.txt.gz", 4e-5)):
        want = _floats(gzip.open(os.path.join(GOLDEN, fixture), "rt").read())
        assert len(got) == len(want) == 130 * 160
        diff = max(abs(a - b) for a, b in zip(got, want))
        same = sum(1 for a, b in zip(got, want) if a == b)
        print(f"mbusu ris10g vs {fixture}: max |diff| {diff:.2e}, {same}/{len(got)} printed identically", flush=True)
        assert diff <= tol, fixture


@pytest.mark.skipif(not (os.path.exists(BIN_MB) and os.path.exists(BIN_MB_REF)),
                    reason="mbusu builds not present (needs the reference tree)")
def test_reference_mbusu_3d_state_against_the_reference_engine(tmp_path):
    """All eight substates of mbusu after 25 steps: the GPU run against the
    reference OpenCAL-OMP 3D engine running the same file on the CPU.  The
    largest difference stays below 1e-12 of the substate's largest value
    (the GPU's pow against glibc's; moist_diff is a difference of two close
    values, so its cell-wise relative error is not meaningful)."""
    import numpy as np
    print("case dropin mbusu 25 steps vs reference engine", flush=True)
    (tmp_path / "gpu").mkdir()
    (tmp_path / "cpu").mkdir()
    r = subprocess.run([BIN_MB, "25", str(tmp_path / "gpu.bin")], cwd=tmp_path / "gpu", capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run([BIN_MB_REF, "25", str(tmp_path / "cpu.bin")], cwd=tmp_path / "cpu", capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    g = np.fromfile(tmp_path / "gpu.bin", np.float64).reshape(8, -1)
    c = np.fromfile(tmp_path / "cpu.bin", np.float64).reshape(8, -1)
    assert g.shape == c.shape == (8, 160 * 30 * 130)
    assert (tmp_path / "gpu" / "deltat.txt").read_text() == (tmp_path / "cpu" / "deltat.txt").read_text()
    for s, name in enumerate(MB_NAMES):
        rel = float(np.max(np.abs(g[s] - c[s])) / np.max(np.abs(c[s])))
        exact = int(np.sum(g[s] == c[s]))
        print(f"mbusu {name}: max |diff| / max |value| {rel:.2e}, {exact}/{g.shape[1]} bit-identical", flush=True)
        assert rel <= 1e-12, name


BIN_U = os.path.join(native.REF_DIR, "dropin_sciddicaT_unsafe")


def _unsafe_data(tmp_path, z=None, h=None):
    data = tmp_path / "data"
    data.mkdir()
    if z is None:
        for name, src in (("dem.txt", "sciddicaT_610x496_dem.txt.gz"),
                          ("source.txt", "sciddicaT_610x496_source.txt.gz")):
            with gzip.open(os.path.join(GOLDEN, src), "rb") as f, open(data / name, "wb") as g:
                shutil.copyfileobj(f, g)
    else:
        import numpy as np
        np.savetxt(data / "dem.txt", z, fmt="%.17g")
        np.savetxt(data / "source.txt", h, fmt="%.17g")
    return data


@pytest.mark.skipif(not os.path.exists(BIN_U), reason="dropin_sciddicaT_unsafe not built (needs the reference tree)")
def test_reference_unsafe_program_on_the_generic_device_program(tmp_path):
    """The reference example cal_sciddicaT-unsafe-glut (its model file
    unchanged: the transition writes its outflows into the neighbours' next
    layer with calSetX2Dr / calGetNextX2Dr and marks them active), its two
    processes compiled for the GPU at run time; the transition runs in colour
    phases.  On the example's data: the raw h equals the oracle in both visit
    orders bit for bit, the saved file equals what the reference OpenCAL-OMP
    engine writes for this program ("%.9f"), and printed with the serial
    engine's "%f" it is the example's shipped width_final.txt."""
    import numpy as np
    from opencal_amd import synth
    print("case dropin sciddicaT unsafe 610x496 4000 steps", flush=True)
    data = _unsafe_data(tmp_path)
    r = subprocess.run([BIN_U, "h.bin"], cwd=tmp_path, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "steps 4000 active 5507" in r.stdout, r.stdout + r.stderr
    want9 = gzip.open(os.path.join(GOLDEN, "sciddicaT_610x496_width_final.txt.gz"), "rt").read()
    assert (data / "width_final.txt").read_text() == want9
    got = np.fromfile(tmp_path / "h.bin", np.float64).reshape(610, 496)
    z = synth.read_matrix_gz(os.path.join(GOLDEN, "sciddicaT_610x496_dem.txt.gz"))
    h = synth.read_matrix_gz(os.path.join(GOLDEN, "sciddicaT_610x496_source.txt.gz"))
    for order in (1, 0):
        o, n_active = native.oracle_sciddicat_unsafe(z, h, 0.5, 0.001, 4000, order)
        assert np.array_equal(got, o) and n_active == 5507, order
    shipped = gzip.open(os.path.join(GOLDEN, "sciddicaT_active_610x496_width_final.txt.gz"), "rt").read()
    assert "".join("".join("%f  " % v for v in row) + "\n" for row in got) == shipped


@pytest.mark.skipif(not os.path.exists(BIN_U), reason="dropin_sciddicaT_unsafe not built (needs the reference tree)")
def test_reference_unsafe_program_on_synthetic_data(tmp_path):
    """The same program on a synthetic slope and source: bit for bit the
    oracle's colour-phase visit; within 1e-12 of the serial visit (the
    reference engine's, which the oracle's serial order equals bit for bit),
    with the same mass."""
    import numpy as np
    from opencal_amd import synth
    print("case dropin sciddicaT unsafe synthetic 610x496 seed 5", flush=True)
    z, h, _ = synth.sciddicat_inputs(610, 496, seed=5, patch=10)
    _unsafe_data(tmp_path, z, h)
    r = subprocess.run([BIN_U, "h.bin"], cwd=tmp_path, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(tmp_path / "h.bin", np.float64).reshape(610, 496)
    phased, _ = native.oracle_sciddicat_unsafe(z, h, 0.5, 0.001, 4000, 1)
    serial, _ = native.oracle_sciddicat_unsafe(z, h, 0.5, 0.001, 4000, 0)
    assert np.array_equal(got, phased)
    assert np.abs(got - serial).max() <= 1e-12 * h.max()
    assert abs(got.sum() - h.sum()) <= 1e-12 * h.sum()


BIN_UP = os.path.join(native.REF_DIR, "dropin_sciddicaT_unsafe_plain")


@pytest.mark.skipif(not os.path.exists(BIN_UP),
                    reason="dropin_sciddicaT_unsafe_plain not built (needs the reference tree)")
@pytest.mark.parametrize("data", ["example", "synthetic"])
def test_reference_full_grid_unsafe_program(tmp_path, data):
    """The reference example cal_sciddicaT-unsafe, its file unchanged
    (sciddicaT.c:86-87 writes each outflow into the neighbour's next layer
    with calSetX2Dr / calGetNextX2Dr; ACTIVE_CELLS is not defined, so the
    active set declared with CAL_OPT_ACTIVE_CELLS stays empty and every
    sweep covers the grid), its own main and calRun2D with only the bind
    calls added; the transition runs in colour phases.  Its "%.9f" file is
    the oracle's of the full-grid model in phase order: on the example's
    data, the file the reference engine writes for it; on a synthetic
    slope, within 1e-12 of the serial order."""
    import numpy as np
    from opencal_amd import synth
    print(f"case dropin sciddicaT unsafe full grid 610x496 4000 steps {data}", flush=True)
    if data == "example":
        _unsafe_data(tmp_path)
        z = synth.read_matrix_gz(os.path.join(GOLDEN, "sciddicaT_610x496_dem.txt.gz"))
        h = synth.read_matrix_gz(os.path.join(GOLDEN, "sciddicaT_610x496_source.txt.gz"))
    else:
        z, h, _ = synth.sciddicat_inputs(610, 496, seed=7, patch=12)
        _unsafe_data(tmp_path, z, h)
    r = subprocess.run([BIN_UP], cwd=tmp_path, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Simulation terminated." in r.stdout, r.stdout + r.stderr
    got = (tmp_path / "data" / "width_final.txt").read_text()
    if data == "example":
        assert got == gzip.open(os.path.join(GOLDEN, "sciddicaT_610x496_width_final.txt.gz"), "rt").read()
        return
    phased, n = native.oracle_sciddicat_unsafe(z, h, 0.5, 0.001, 4000, 1, track=False)
    assert n == 0
    assert got == "".join("".join("%.9f  " % v for v in row) + "\n" for row in phased)
    vals = np.array([[float(x) for x in line.split()] for line in got.splitlines()])
    assert abs(vals.sum() - h.sum()) <= 1e-6 * h.sum()


@pytest.mark.parametrize("variant,seed,steps", [("", 3, 400), ("", 8, 150), ("_implicit", 3, 400)])
def test_reference_s3hex_program_on_the_generic_device_program(tmp_path, variant, seed, steps):
    """The reference example cal_sciddicaS3hex-unsafe-explicit-glut (its
    model file unchanged: hexagonal torus, its own global transition of four
    calApplyElementaryProcess2D calls each followed by calUpdateSubstate2Dr,
    a flows process that writes its outflows into the neighbours' next
    layers of h and p, calSetCurrent2D on its single-layer substates), its
    four processes compiled for the GPU at run time; the flows run in colour
    phases.  The example ships no DEM, so the inputs are synthetic rasters of
    its 767 x 925 shape.  Every substate (h, p, z, d and the int countdown s)
    equals the oracle's colour-phase visit bit for bit; the oracle's serial
    visit is the reference program's own result on the reference engine
    (tests/test_oracle_s3hex.py).  The implicit twin
    cal_sciddicaS3hex-unsafe-glut (variant "_implicit": the same processes
    under calGlobalTransitionFunction2D with an update after each) gives the
    same bits."""
    import numpy as np
    BIN_S3 = os.path.join(native.REF_DIR, "dropin_s3hex" + variant)
    if not os.path.exists(BIN_S3):
        pytest.skip(f"dropin_s3hex{variant} not built (needs the reference tree)")
    print(f"case dropin s3hex{variant} 767x925 seed {seed}, {steps} steps", flush=True)
    z, d, s = native.s3hex_inputs(767, 925, seed)
    data = tmp_path / "data"
    data.mkdir()
    np.savetxt(data / "dem.txt", z, fmt="%.17g")
    np.savetxt(data / "regolith.txt", d, fmt="%.17g")
    np.savetxt(data / "source.txt", s, fmt="%d")
    r = subprocess.run([BIN_S3, str(tmp_path / "got"), str(steps)], cwd=tmp_path, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"steps {steps}" in r.stdout, r.stdout + r.stderr
    assert (data / "width_final.txt").exists()
    got = {k: np.fromfile(tmp_path / f"got_{k}.bin", np.float64).reshape(767, 925) for k in ("h", "p", "z", "d")}
    got["s"] = np.fromfile(tmp_path / "got_s.bin", np.int32).reshape(767, 925)
    want = native.oracle_s3hex(z, d, s, steps, 1)
    for k in ("h", "p", "z", "d", "s"):
        assert np.array_equal(got[k], want[k]), f"{k}: max diff {np.abs(got[k] - want[k]).max()}"
    # the sixteen slides started and spread (a film of debris on thousands of cells)
    assert (got["z"] < z).sum() >= 16 and (got["h"] > 0).sum() > 1000


BIN_MOD2 = os.path.join(native.REF_DIR, "dropin_mod2CA3D")


@pytest.mark.skipif(not os.path.exists(BIN_MOD2), reason="dropin_mod2CA3D not built (needs the reference tree)")
def test_reference_mod2_3d_program_on_the_generic_device_program(tmp_path):
    """The reference 3D example cal_mod2CA3D-glut (mod2CA3D.c unchanged:
    65^3 torus, Moore 3D, a byte substate, the parity of each 27-cell sum,
    one seed cell, 1000 steps), its process compiled for the GPU at run time:
    the final state is the closed form's (the outer product of the 1D
    trinomial parities), which the reference program on the reference engine
    also reaches (tests/test_engine3d.py)."""
    import numpy as np
    from test_engine3d import mod2_closed_form
    print("case dropin mod2CA3D 65^3 1000 steps", flush=True)
    out = tmp_path / "mod2.bin"
    r = subprocess.run([BIN_MOD2, str(out)], cwd=tmp_path, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "steps 1000 live 79507" in r.stdout, r.stdout + r.stderr
    got = np.fromfile(out, np.uint8).reshape(65, 65, 65)
    assert np.array_equal(got, mod2_closed_form((65, 65, 65), (12, 12, 12), 1000))
