"""Writes tests/golden/mbusu_ris10g_reference_build.txt.gz: the ris10g.txt
that the reference mbusu example (OpenCAL-OMP/examples/calomp_mbusu-glut,
unchanged) writes at its stop condition when built from the reference
sources on the reference OpenCAL-OMP 3D engine in this container
(oracle/_ref/calomp_mbusu, `make -C oracle/ref`), gzip-compressed.

The reference also ships ris10g-orig-opencal-omp.txt
(mbusu_ris10g-orig-opencal-omp.txt.gz here), written on another machine;
the reference build above differs from it by up to 3.3e-5 (its own serial and OpenMP files
differ by 2.4e-5), so both are kept.  Runs for a few minutes on 8 cores.

    python3 tests/golden/make_mbusu_fixture.py
"""
import gzip
import os
import shutil
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(ROOT, "oracle", "_ref", "calomp_mbusu")


def main():
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([BIN], cwd=d, capture_output=True, text=True, check=True)
        assert "steps 8662" in r.stdout, r.stdout
        with open(os.path.join(d, "ris10g.txt"), "rb") as f, \
                gzip.open(os.path.join(HERE, "mbusu_ris10g_reference_build.txt.gz"), "wb", 9) as g:
            shutil.copyfileobj(f, g)


if __name__ == "__main__":
    main()
