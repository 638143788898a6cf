"""The S3 debris-flow oracle (oracle/oracle_s3hex.c: the reference example
cal_sciddicaS3hex-unsafe-explicit-glut, sciddicaS3Hex.c, restated) against
the reference program itself built on the reference OpenCAL-OMP engine
(oracle/_ref/calomp_s3hex, one thread: its sweeps visit the cells in
row-major order), and its implicit-update twin
cal_sciddicaS3hex-unsafe-glut (calomp_s3hex_implicit: the same processes
under the engine's transition, which gives the same result).  The example ships its regolith and source rasters but no
DEM, so the inputs are synthetic rasters of its 767 x 925 shape
(native.s3hex_inputs), written as text that reads back to the same doubles.

Pins: order 0 equals the reference program bit for bit in every substate
(h, p, z, d, s); order 1 (the flows in the device's colour phases) differs
from it only by the rounding of the sums a cell receives, and conserves the
debris volume on the torus."""
import os
import subprocess

import numpy as np
import pytest

import native

# the explicit example (its own global transition) and the implicit one (the
# same processes, the engine's transition with implicit updates)
BINS = {v: os.path.join(native.REF_DIR, "calomp_s3hex" + v) for v in ("", "_implicit")}
ROWS, COLS = 767, 925


def write_inputs(root, z, d, s):
    os.makedirs(os.path.join(root, "data"), exist_ok=True)
    for name, a, fmt in (("dem", z, "%.17g"), ("regolith", d, "%.17g"), ("source", s, "%d")):
        np.savetxt(os.path.join(root, "data", name + ".txt"), a, fmt=fmt, delimiter=" ")


def read_dump(prefix):
    out = {k: np.fromfile(f"{prefix}_{k}.bin", np.float64).reshape(ROWS, COLS) for k in ("h", "p", "z", "d")}
    out["s"] = np.fromfile(f"{prefix}_s.bin", np.int32).reshape(ROWS, COLS)
    return out


def test_s3hex_oracle_volume_and_orders():
    z, d, s = native.s3hex_inputs(ROWS, COLS, 3)
    a = native.oracle_s3hex(z, d, s, 120, 0)
    b = native.oracle_s3hex(z, d, s, 120, 1)
    # the landslides started and moved; every triggered cell gave its regolith
    assert (a["h"] > 0.01).sum() > 16 and (a["z"] < z).sum() >= 16
    eroded = (d - a["d"])[a["d"] < d].sum()
    assert abs(a["h"].sum() - eroded) <= 1e-9 * eroded
    for k in ("h", "p"):
        scale = max(np.abs(a[k]).max(), 1e-300)
        assert np.abs(a[k] - b[k]).max() <= 1e-12 * scale, k
    for k in ("z", "d", "s"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("variant,seed,steps", [("", 3, 150), ("", 8, 60), ("_implicit", 3, 150)])
def test_s3hex_oracle_equals_reference_program(tmp_path, variant, seed, steps):
    BIN = BINS[variant]
    if not os.path.exists(BIN):
        pytest.skip(f"calomp_s3hex{variant} not built (needs the reference tree)")
    z, d, s = native.s3hex_inputs(ROWS, COLS, seed)
    write_inputs(str(tmp_path), z, d, s)
    env = dict(os.environ, OMP_NUM_THREADS="1")
    r = subprocess.run([BIN, str(tmp_path / "ref"), str(steps)], cwd=str(tmp_path), env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"steps {steps}" in r.stdout
    ref = read_dump(str(tmp_path / "ref"))
    # the inputs read back exactly
    assert np.array_equal(native.oracle_s3hex(z, d, s, 0, 0)["z"], z)
    ora = native.oracle_s3hex(z, d, s, steps, 0)
    for k in ("h", "p", "z", "d", "s"):
        assert np.array_equal(ora[k], ref[k]), f"{k}: {np.abs(ora[k] - ref[k]).max()}"
