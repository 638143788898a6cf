"""Readers for the reference SCIARA-fv2 example data kept under
tests/golden/sciara2006 (OpenCAL-OMP/examples/calomp_sciara-fv2/data/2006):
the parameter file, the ESRI-style .stt matrices and the emission-rate file,
turned into the input dict of native.oracle_sciara.  The formats follow the
reference loaders (io.cpp loadParameters / loadMorphology / loadVents,
vent.cpp initVents / loadEmissionRates)."""
from __future__ import annotations

import gzip
import os

import numpy as np

DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sciara2006")
PREFIX = "2006_000000000000"
PARAM_KEYS = ("Pclock", "PTsol", "PTvent", "Pr(Tsol)", "Pr(Tvent)", "Phc(Tsol)", "Phc(Tvent)", "Pcool", "Prho",
              "Pepsilon", "Psigma", "Pcv")


def _open(name):
    p = os.path.join(DIR, name)
    return gzip.open(p + ".gz", "rt") if os.path.exists(p + ".gz") else open(p)


def read_stt(f):
    """(header dict, matrix) of an ESRI-style text grid (6 header lines)."""
    head = {}
    for _ in range(6):
        k, v = f.readline().split()
        head[k] = float(v)
    vals = np.array(f.read().split(), dtype=np.float64)
    rows, cols = int(head["nrows"]), int(head["ncols"])
    return head, vals[: rows * cols].reshape(rows, cols)


def write_inputs(dst_dir):
    """The data files, uncompressed, as the reference program expects them
    (dst_dir/data/2006/...)."""
    out = os.path.join(dst_dir, "data", "2006")
    os.makedirs(out, exist_ok=True)
    for name in os.listdir(DIR):
        src = os.path.join(DIR, name)
        dst = os.path.join(out, name[:-3] if name.endswith(".gz") else name)
        opener = gzip.open if name.endswith(".gz") else open
        with opener(src, "rb") as f, open(dst, "wb") as g:
            g.write(f.read())
    os.makedirs(os.path.join(dst_dir, "data", "2006_SAVE"), exist_ok=True)


def load_inputs():
    tok = _open(PREFIX + ".cfg").read().split()
    params = np.array([float(tok[tok.index(k) + 1]) for k in PARAM_KEYS], dtype=np.float64)
    with _open(PREFIX + "_Morphology.stt") as f:
        head, z0 = read_stt(f)
    with _open(PREFIX + "_Vents.stt") as f:
        _, mv = read_stt(f)
    ids = mv.astype(np.int64).ravel()
    cells = np.nonzero(ids > 0)[0]                       # initVents: row-major scan
    vents = np.stack([cells // z0.shape[1], cells % z0.shape[1]], axis=1).astype(np.int32)
    tok = _open(PREFIX + "_EmissionRate.txt").read().split()
    emission_time = int(tok[tok.index("[TIME]") + 1])
    per_id, k = {}, tok.index("[TIME]") + 2
    while tok[k] != "[END]":
        assert tok[k] == "[VENT_ID]" and tok[k + 2] == "[EMISSION_RATE]"
        vid, k = int(tok[k + 1]), k + 3
        vals = []
        while tok[k] not in ("[VENT_ID]", "[END]"):
            vals.append(float(tok[k]))
            k += 1
        per_id[vid] = vals
    n_rates = max(len(v) for v in per_id.values())
    rates = np.zeros((len(cells), n_rates))
    for i, c in enumerate(cells):
        r = per_id[int(ids[c])]
        rates[i, : len(r)] = r
    return dict(z0=np.ascontiguousarray(z0), cell_size=float(head["cellsize"]), params=params,
                vents=np.ascontiguousarray(vents), rates=np.ascontiguousarray(rates),
                emission_time=emission_time)


def read_output(run_dir, name):
    """A matrix saved by the reference program's saveConfiguration."""
    with open(os.path.join(run_dir, "data", "2006_SAVE", f"{PREFIX}_{name}.stt")) as f:
        return read_stt(f)[1]
