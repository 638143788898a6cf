"""The C-ABI boundary, on the CPU (no GPU needed).

1. libopencal_hip.so loads and exports every function declared in include/*.h.
2. Host-phase semantics of the OpenCAL API (model definition, neighbourhoods,
   substates, init / set / get / update, active-cell add / remove / rebuild,
   text IO) are checked differentially against the real OpenCAL-OMP engine
   (its functions are exported by the oracle/_ref build), reading the model
   through the same struct layout on both sides.
3. Without a GPU, or with processes not bound to a device program, the
   global transition reports an error instead of computing anything on the
   CPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for fn in sorted(os.listdir(INCLUDE)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
        for m in re.finditer(r"\b(cal[A-Za-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from opencal_amd import LIB_PATH
    lib = C.CDLL(LIB_PATH)
    names = declared_functions()
    assert len(names) > 100
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


# ------------------------------------------------------------ struct layout

class CALCell2D(C.Structure):
    _fields_ = [("i", C.c_int), ("j", C.c_int)]


class CALActiveCells2D(C.Structure):
    _fields_ = [("flags", C.POINTER(C.c_byte)), ("size_next", C.POINTER(C.c_int)), ("cells", C.POINTER(CALCell2D)),
                ("size_current", C.c_int), ("num_threads", C.c_int)]


class CALModel2D(C.Structure):
    _fields_ = [("rows", C.c_int), ("columns", C.c_int), ("T", C.c_int), ("is_safe", C.c_int),
                ("OPTIMIZATION", C.c_int), ("A", CALActiveCells2D), ("X", C.POINTER(CALCell2D)),
                ("sizeof_X", C.c_int), ("X_id", C.c_int), ("pQb_array", C.c_void_p), ("pQi_array", C.c_void_p),
                ("pQr_array", C.c_void_p), ("sizeof_pQb_array", C.c_int), ("sizeof_pQi_array", C.c_int),
                ("sizeof_pQr_array", C.c_int), ("elementary_processes", C.c_void_p),
                ("num_of_elementary_processes", C.c_int), ("locks", C.c_void_p)]


class Sub(C.Structure):
    _fields_ = [("current", C.c_void_p), ("next", C.c_void_p)]


def _sig(L):
    P = C.POINTER(CALModel2D)
    L.calCADef2D.restype = P
    L.calCADef2D.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    for suf, t in (("b", C.c_byte), ("i", C.c_int), ("r", C.c_double)):
        for name in (f"calAddSubstate2D{suf}", f"calAddSingleLayerSubstate2D{suf}"):
            getattr(L, name).restype = C.POINTER(Sub)
            getattr(L, name).argtypes = [P]
        getattr(L, f"calInitSubstate2D{suf}").argtypes = [P, C.POINTER(Sub), t]
        getattr(L, f"calInitSubstateNext2D{suf}").argtypes = [P, C.POINTER(Sub), t]
        getattr(L, f"calUpdateSubstate2D{suf}").argtypes = [P, C.POINTER(Sub)]
        getattr(L, f"calInit2D{suf}").argtypes = [P, C.POINTER(Sub), C.c_int, C.c_int, t]
        getattr(L, f"calSet2D{suf}").argtypes = [P, C.POINTER(Sub), C.c_int, C.c_int, t]
        getattr(L, f"calSetCurrent2D{suf}").argtypes = [P, C.POINTER(Sub), C.c_int, C.c_int, t]
        getattr(L, f"calGet2D{suf}").restype = t
        getattr(L, f"calGet2D{suf}").argtypes = [P, C.POINTER(Sub), C.c_int, C.c_int]
        getattr(L, f"calGetX2D{suf}").restype = t
        getattr(L, f"calGetX2D{suf}").argtypes = [P, C.POINTER(Sub), C.c_int, C.c_int, C.c_int]
        getattr(L, f"calSaveSubstate2D{suf}").argtypes = [P, C.POINTER(Sub), C.c_char_p]
        getattr(L, f"calLoadSubstate2D{suf}").argtypes = [P, C.POINTER(Sub), C.c_char_p]
    for name in ("calUpdate2D", "calUpdateActiveCells2D", "calFinalize2D", "calGlobalTransitionFunction2D"):
        getattr(L, name).argtypes = [P]
    for name in ("calAddActiveCell2D", "calRemoveActiveCell2D"):
        getattr(L, name).argtypes = [P, C.c_int, C.c_int]
    L.calAddActiveCellX2D.argtypes = [P, C.c_int, C.c_int, C.c_int]
    L.calAddNeighbor2D.restype = C.POINTER(CALCell2D)
    L.calAddNeighbor2D.argtypes = [P, C.c_int, C.c_int]
    return L


@pytest.fixture(scope="module")
def libs():
    from opencal_amd import LIB_PATH
    mine = _sig(C.CDLL(LIB_PATH))
    refp = os.path.join(native.REF_DIR, "libref_sciddicat.so")
    ref = _sig(C.CDLL(refp)) if os.path.exists(refp) else None
    return mine, ref


def _arr(ptr, n, ctype):
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(n,)).copy()


NBS = {1: "VN", 2: "MOORE", 3: "HEX", 4: "HEX_ALT"}


@pytest.mark.parametrize("nb", [1, 2, 3, 4])
def test_neighbourhoods_match_reference(libs, nb):
    mine, ref = libs
    if ref is None:
        pytest.skip("reference build not present")
    a, b = mine.calCADef2D(7, 9, nb, 1, 0), ref.calCADef2D(7, 9, nb, 1, 0)
    assert a.contents.sizeof_X == b.contents.sizeof_X
    n_tab = 14 if nb >= 3 else a.contents.sizeof_X
    for k in range(n_tab):
        assert (a.contents.X[k].i, a.contents.X[k].j) == (b.contents.X[k].i, b.contents.X[k].j)
    # calAddNeighbor2D appends
    mine.calAddNeighbor2D(a, 2, -2)
    ref.calAddNeighbor2D(b, 2, -2)
    assert a.contents.sizeof_X == b.contents.sizeof_X
    mine.calFinalize2D(a)
    ref.calFinalize2D(b)


def _scenario(L, flat, opt, rng_seed):
    """The same sequence of host-phase API calls; returns the observable state."""
    rows, cols = 9, 11
    rng = np.random.default_rng(rng_seed)
    m = L.calCADef2D(rows, cols, 2, 0 if flat else 1, opt)
    qr = L.calAddSubstate2Dr(m)
    qi = L.calAddSubstate2Di(m)
    qb = L.calAddSubstate2Db(m)
    sr = L.calAddSingleLayerSubstate2Dr(m)
    L.calInitSubstate2Dr(m, qr, 1.5)
    L.calInitSubstate2Di(m, qi, 1)      # full grid: the reference fills bytes (memset)
    L.calInitSubstate2Db(m, qb, 3)
    L.calInitSubstate2Dr(m, sr, -2.0)
    out = []
    for _ in range(30):
        i, j = int(rng.integers(1, rows - 1)), int(rng.integers(1, cols - 1))
        L.calSet2Dr(m, qr, i, j, float(rng.random()))
        L.calSet2Di(m, qi, i, j, int(rng.integers(0, 100)))
        L.calInit2Db(m, qb, i, j, int(rng.integers(0, 5)))
        L.calSetCurrent2Dr(m, sr, i, j, float(rng.random()))
        if opt:
            L.calAddActiveCell2D(m, i, j)
            L.calAddActiveCellX2D(m, i, j, int(rng.integers(0, 9)))
            if rng.random() < 0.3:
                L.calRemoveActiveCell2D(m, i, j)
        out.append(L.calGetX2Dr(m, qr, i, j, int(rng.integers(0, 9))))
        out.append(L.calGet2Di(m, qi, i, j))
    L.calUpdate2D(m)
    if opt:
        L.calInitSubstateNext2Dr(m, qr, 7.0)   # over the active cells only
        L.calUpdateSubstate2Dr(m, qr)
    n = rows * cols
    state = {
        "gets": np.array(out),
        "r_cur": _arr(qr.contents.current, n, C.c_double), "r_next": _arr(qr.contents.next, n, C.c_double),
        "i_cur": _arr(qi.contents.current, n, C.c_int), "i_next": _arr(qi.contents.next, n, C.c_int),
        "b_cur": _arr(qb.contents.current, n, C.c_byte), "s_cur": _arr(sr.contents.current, n, C.c_double),
        "size": m.contents.A.size_current,
        "cells": [(m.contents.A.cells[k].i, m.contents.A.cells[k].j) for k in range(m.contents.A.size_current)]
        if m.contents.A.cells else None,
        "nq": (m.contents.sizeof_pQb_array, m.contents.sizeof_pQi_array, m.contents.sizeof_pQr_array),
    }
    L.calFinalize2D(m)
    return state


@pytest.mark.parametrize("flat,opt", [(False, 0), (False, 1), (True, 0), (True, 1)])
def test_host_api_semantics_match_reference(libs, flat, opt):
    mine, ref = libs
    if ref is None:
        pytest.skip("reference build not present")
    a, b = _scenario(mine, flat, opt, 3), _scenario(ref, flat, opt, 3)
    for k in a:
        if isinstance(a[k], np.ndarray):
            assert np.array_equal(a[k], b[k]), k
        else:
            assert a[k] == b[k], k


def test_int_fill_is_bytewise(libs):
    """calInitSubstate2Di over the full grid goes through memset in the
    reference (cal2DBuffer.c:227-230): value 1 becomes 0x01010101."""
    mine, _ = libs
    m = mine.calCADef2D(4, 4, 1, 1, 0)
    q = mine.calAddSubstate2Di(m)
    mine.calInitSubstate2Di(m, q, 1)
    assert mine.calGet2Di(m, q, 2, 3) == 0x01010101
    mine.calFinalize2D(m)


def test_text_io_roundtrip(libs, tmp_path):
    mine, ref = libs
    m = mine.calCADef2D(5, 6, 1, 1, 0)
    q = mine.calAddSubstate2Dr(m)
    vals = np.arange(30, dtype=np.float64) / 7.0
    for k, v in enumerate(vals):
        mine.calInit2Dr(m, q, k // 6, k % 6, float(v))
    p = str(tmp_path / "q.txt").encode()
    mine.calSaveSubstate2Dr(m, q, p)
    text = open(p, "r").read()
    assert text.splitlines()[0].startswith("%.9f  %.9f" % (vals[0], vals[1]))
    q2 = mine.calAddSubstate2Dr(m)
    mine.calLoadSubstate2Dr.restype = C.c_byte   # CALbyte
    assert mine.calLoadSubstate2Dr(m, q2, p) == 1
    assert abs(mine.calGet2Dr(m, q2, 4, 5) - vals[29]) < 1e-9
    if ref is not None:
        r = ref.calCADef2D(5, 6, 1, 1, 0)
        rq = ref.calAddSubstate2Dr(r)
        for k, v in enumerate(vals):
            ref.calInit2Dr(r, rq, k // 6, k % 6, float(v))
        p2 = str(tmp_path / "r.txt").encode()
        ref.calSaveSubstate2Dr(r, rq, p2)
        assert open(p2).read() == text
        ref.calFinalize2D(r)
    mine.calFinalize2D(m)


def test_unbound_transition_is_an_error(libs):
    """No device program bound: the global transition reports an error and
    computes nothing (there is no CPU execution path)."""
    mine, _ = libs
    from opencal_amd import lib
    L = lib()
    L.calhipClearError()
    m = mine.calCADef2D(8, 8, 1, 1, 0)
    q = mine.calAddSubstate2Dr(m)
    mine.calInitSubstate2Dr(m, q, 2.0)
    cb = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int)(lambda *a: None)
    L.calAddElementaryProcess2D.argtypes = [C.c_void_p, C.c_void_p]
    L.calAddElementaryProcess2D(C.cast(m, C.c_void_p), C.cast(cb, C.c_void_p))
    mine.calGlobalTransitionFunction2D(m)
    assert L.calhipGetLastError() == 2
    assert mine.calGet2Dr(m, q, 3, 3) == 2.0
    L.calhipClearError()
    mine.calFinalize2D(m)


def test_sciara_binding_validates_and_fails_loudly_without_gpu():
    """The SCIARA example binds on the host (the binding's checks pass for a
    model built as the reference builds it); wrong model shapes are refused;
    with no GPU visible the run reports an error instead of computing."""
    import torch
    from opencal_amd import lib
    from opencal_amd.models import Sciara
    from opencal_amd import synth
    inp = synth.sciara_inputs(32, 40, 1, steps_per_rate=5, n_rates=2)
    m = Sciara(inp, 10, active=True)   # binds; no device call yet
    L = lib()
    assert L.calhipGetLastError() == 0
    if not torch.cuda.is_available():
        assert m.E.sciara_run(m.handle) == -1
        assert L.calhipGetLastError() != 0   # no device: a HIP error, nothing computed
    m.close()
    L.calhipClearError()
    # a von Neumann model is refused
    P = C.c_void_p
    L.calCADef2D.restype = P
    vn = L.calCADef2D(8, 8, 1, 0, 0)
    class Binding(C.Structure):
        _fields_ = [("ptrs", C.c_void_p * 64)]
    b = Binding()
    L.calhipBindSciara2D.argtypes = [P, C.POINTER(Binding)]
    assert L.calhipBindSciara2D(vn, C.byref(b)) == -1
    assert L.calhipGetLastError() == 31
    L.calhipClearError()
    L.calFinalize2D.argtypes = [P]
    L.calFinalize2D(vn)
