"""The reference's own SciddicaT program (OpenCAL-OMP/examples/calomp_sciddicaT/
source/sciddicaT.c: its main, data paths and output), built against this
engine with the two-line patch of INTEGRATION.md (oracle/ref/Makefile target
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


BIN_G = os.path.join(native.REF_DIR, "dropin_sciddicaT_generic")


@pytest.mark.skipif(not os.path.exists(BIN_G), reason="dropin_sciddicaT_generic not built (needs the reference tree)")
def test_reference_program_on_the_generic_device_program(tmp_path):
    """The same program with its own elementary processes compiled for the
    GPU at run time (its globals marked CALHIP_GLOBAL, five bind calls): the
    shipped width_final.txt byte for byte."""
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
    GPU's files equal, as text, the oracle run with the same portable
    elementary functions, and lie within 1e-9 relative of the glibc run (which
    equals the reference program on the CPU)."""
    import numpy as np
    import sciara2006
    print("case dropin sciara 2006 data 378x517 1000 steps", flush=True)
    sciara2006.write_inputs(str(tmp_path))
    r = subprocess.run([BIN_SC], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    inp = sciara2006.load_inputs()
    port = native.oracle_sciara(inp, 1000, False, portable=True)
    libm = native.oracle_sciara(inp, 1000, False, portable=False)
    for name, k, scale in (("Thickness", "h", 1.0), ("Temperature", "T", 1400.0), ("Morphology", "z", 3400.0),
                           ("SolidifiedLavaThickness", "Msl", 1.0)):
        path = tmp_path / "data" / "2006_SAVE" / f"{sciara2006.PREFIX}_{name}.stt"
        text = path.read_text().split()[12:]
        assert text == ["%.9f" % v for v in port[k].ravel()], name
        got = np.array(text, dtype=np.float64)
        assert np.abs(got - libm[k].ravel()).max() <= 1e-9 * scale + 1e-9, name
