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
