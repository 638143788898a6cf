"""The 3D engine (include/cal3D*.h, calhip3D.h) and the global reductions
(cal2DReduction.h, cal3DReduction.h).

CPU: the 3D process sources compile for gfx950; the 3D oracle
(oracle/engine3d.py) is pinned by the mod-2 automaton's closed form, by heat
conservation, and by the reference OpenCAL-OMP 3D engine built from the
reference sources running the same model file (oracle/_ref); the reduction
fold is pinned by known answers.  The engine's neighbourhood tables equal
the reference's.

GPU: every 3D model of opencal_amd/examples/generic_models3d.c bit-exact
against the oracle (mod-2 CA, heat with reductions in the steering and the
stop condition, an active-cells model on both neighbourhoods and
boundaries, the unsafe operations), and every reduction on every type,
before the model is on the GPU (host arrays staged) and after."""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

import native
from opencal_amd import examples, lib
from oracle import engine3d as O

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_bp = C.POINTER(C.c_ubyte)
_sbp = C.POINTER(C.c_byte)
U = 2.0 ** -53

REF3 = os.path.join(native.REF_DIR, "libref_models3d.so")


def _ref3():
    if not os.path.exists(REF3):
        return None
    L = native._load(REF3)
    vp = C.c_void_p
    L.gmod2_create.restype = vp
    L.gmod2_create.argtypes = [C.c_int, C.c_int, C.c_int, _bp, C.c_int]
    L.gmod2_run.argtypes = [vp, _bp]
    L.gmod2_destroy.argtypes = [vp]
    L.gheat_create.restype = vp
    L.gheat_create.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_double, C.c_double, C.c_double, C.c_int, _dp]
    L.gheat_run.argtypes = [vp, _dp, _bp, _ip]
    L.gheat_destroy.argtypes = [vp]
    L.ginf_create.restype = vp
    L.ginf_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _ip, C.c_int, C.c_int]
    L.ginf_steps.argtypes = [vp, C.c_int]
    L.ginf_apply_steps.argtypes = [vp, C.c_int]
    L.ginf_read.argtypes = [vp, _bp, _bp, _ip, _ip, _ip, C.c_int]
    L.ginf_destroy.argtypes = [vp]
    L.ggift3_create.restype = vp
    L.ggift3_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    L.ggift3_steps.argtypes = [vp, C.c_int]
    L.ggift3_read.argtypes = [vp, _ip, _ip, _dp]
    L.ggift3_table.argtypes = [vp, _ip, C.c_int]
    L.ggift3_destroy.argtypes = [vp]
    return L


# ----------------------------------------------------------- the models

def oracle_mod2(init, steps):
    s, r, c = init.shape
    m = O.Model3D(r, c, s, "moore", torus=True)
    m.add("q", "b")
    for k in range(s):
        for i in range(r):
            for j in range(c):
                m.init_cell("q", i, j, k, int(init[k, i, j]))

    def ep(M, i, j, k):
        M.set("q", i, j, k, sum(M.getx("q", i, j, k, n) for n in range(len(M.X))) % 2)

    for _ in range(steps):
        m.transition([ep])
    return m.array("q").astype(np.uint8)


def mod2_closed_form(shape, seed, steps):
    """Single seed on a torus: the automaton is linear and separable, so the
    state is the outer product of the 1D trinomial parities
    (x^-1 + 1 + x)^t mod 2, each wrapped onto its axis."""
    def axis(n, at):
        coef = np.zeros(2 * steps + 1, np.int64)
        coef[steps] = 1
        for _ in range(steps):
            coef = (np.roll(coef, 1) + coef + np.roll(coef, -1)) % 2
        a = np.zeros(n, np.int64)
        for p, v in enumerate(coef):
            a[(at + p - steps) % n] ^= v
        return a
    s, r, c = shape
    ak, ai, aj = axis(s, seed[0]), axis(r, seed[1]), axis(c, seed[2])
    return (ak[:, None, None] * ai[None, :, None] * aj[None, None, :] % 2).astype(np.uint8)


def oracle_heat(T0, kappa, thr, steps):
    s, r, c = T0.shape
    m = O.Model3D(r, c, s, "von_neumann", torus=False)
    m.add("T", "r")
    m.add("hot", "b")
    m.add("count", "i")
    flat = T0.reshape(-1).tolist()
    m.sub["T"][1][0][:] = flat
    m.sub["T"][1][1][:] = flat

    def ep(M, i, j, k):
        t = M.get("T", i, j, k)
        d = 0.0
        for n in range(1, len(M.X)):
            di, dj, dk = M.X[n]
            ii, jj, kk = i + di, j + dj, k + dk
            if ii < 0 or jj < 0 or kk < 0 or ii >= M.rows or jj >= M.cols or kk >= M.slices:
                continue
            d += M.getx("T", i, j, k, n) - t
        t = t + kappa * d
        M.set("T", i, j, k, t)
        M.set("hot", i, j, k, 1 if t > thr else 0)
        if t > thr:
            M.set("count", i, j, k, M.get("count", i, j, k) + 1)

    rec = []
    for _ in range(steps):
        m.transition([ep])
        rec.append(m.array("T").copy())
    return m.array("T"), m.array("hot").astype(np.uint8), m.array("count"), rec


def oracle_inf(shape, nb, torus, seeds, span, steps):
    s, r, c = shape
    m = O.Model3D(r, c, s, nb, torus=torus, active=True)
    m.add("s", "b")
    m.add("age", "i")
    for (i, j, k) in seeds:
        m.init_cell("s", i, j, k, 1)
        m.add_active(i, j, k)
    m.update()

    def inside(M, i, j, k, n):
        di, dj, dk = M.X[n]
        return torus or (0 <= i + di < M.rows and 0 <= j + dj < M.cols and 0 <= k + dk < M.slices)

    def spread(M, i, j, k):
        st = M.get("s", i, j, k)
        if st == 1:
            for n in range(1, len(M.X)):
                if inside(M, i, j, k, n):
                    M.add_active_x(i, j, k, n)
            M.set("age", i, j, k, M.get("age", i, j, k) + 1)
        elif st == 0:
            for n in range(1, len(M.X)):
                if inside(M, i, j, k, n) and M.getx("s", i, j, k, n) == 1:
                    M.set("s", i, j, k, 1)
                    break

    def retire(M, i, j, k):
        if M.get("age", i, j, k) >= span:
            M.remove_active(i, j, k)
            M.set("s", i, j, k, 2)

    for _ in range(steps):
        m.transition([spread, retire])
    cells = m.cells or []
    return (m.array("s", 0).astype(np.uint8), m.array("s", 1).astype(np.uint8), m.array("age", 0),
            m.array("age", 1), np.array(cells, np.int32).reshape(-1, 3))


def oracle_gift(shape, nb, torus, steps, apply_only=False):
    s, r, c = shape
    m = O.Model3D(r, c, s, nb, torus=torus)
    m.add("q", "i")
    m.add("seen", "i")
    m.add("acc", "r")

    def give(M, i, j, k):
        w = 1 + (i * 7 + j * 3 + k * 5) % 5
        for n in range(1, len(M.X)):
            M.add_next_x("q", i, j, k, n, w)
            M.add_next_x("acc", i, j, k, n, 0.25 * w)

    def look(M, i, j, k):
        M.set("seen", i, j, k, M.get_next("q", i, j, k) + M.getx("q", i, j, k, 1))

    for _ in range(steps):
        m.transition([give] if apply_only else [give, look])
    return m.array("q"), m.array("seen"), m.array("acc")


NB = {"von_neumann": 1, "moore": 2}


# ------------------------------------------------------------------ CPU

def test_3d_sources_compile():
    log = C.create_string_buffer(65536)
    assert examples().generic3d_check_sources(log, 65536) == 0, log.value.decode()


def test_oracle_mod2_closed_form():
    for shape, seed, steps in [((5, 6, 7), (2, 3, 1), 4), ((4, 9, 5), (0, 8, 4), 6)]:
        init = np.zeros(shape, np.uint8)
        init[seed] = 1
        assert np.array_equal(oracle_mod2(init, steps), mod2_closed_form(shape, seed, steps))


def test_oracle_heat_conserves():
    rng = np.random.default_rng(3)
    T0 = rng.uniform(0.0, 100.0, (4, 5, 6))
    T, hot, count, rec = oracle_heat(T0, 0.1, 50.0, 8)
    s0 = math.fsum(T0.reshape(-1).tolist())
    for Tt in rec:   # zero-flux box: the total is kept up to rounding
        assert abs(math.fsum(Tt.reshape(-1).tolist()) - s0) <= 1e-12 * s0
    assert T.min() >= T0.min() - 1e-12 and T.max() <= T0.max() + 1e-12   # maximum principle (kappa <= 1/6)
    assert np.array_equal(hot, (T > 50.0).astype(np.uint8))


@pytest.mark.parametrize("nb", ["von_neumann", "moore"])
def test_engine_tables_equal_the_reference(nb):
    E = examples()
    hd = E.ggift3_create(3, 4, 5, NB[nb], 1)
    assert hd
    try:
        tab = np.zeros(3 * 32, np.int32)
        n = E.ggift3_table(hd, tab.ctypes.data_as(_ip), 32)
        ours = [tuple(int(v) for v in tab[3 * q:3 * q + 3]) for q in range(n)]
    finally:
        E.ggift3_destroy(hd)
    assert ours == (O.VON_NEUMANN if nb == "von_neumann" else O.MOORE)
    R = _ref3()
    if R is None:
        pytest.skip("reference build (oracle/_ref) not present")
    hd = R.ggift3_create(3, 4, 5, NB[nb], 1)
    try:
        n = R.ggift3_table(hd, tab.ctypes.data_as(_ip), 32)
        theirs = [tuple(int(v) for v in tab[3 * q:3 * q + 3]) for q in range(n)]
    finally:
        R.ggift3_destroy(hd)
    assert theirs == ours


@pytest.mark.skipif(not os.path.exists(REF3), reason="reference build (oracle/_ref) not present")
def test_oracle_vs_reference_engine_3d():
    R = _ref3()
    rng = np.random.default_rng(11)
    # mod-2 automaton
    init = (rng.random((5, 6, 4)) < 0.3).astype(np.uint8)
    hd = R.gmod2_create(6, 4, 5, init.ctypes.data_as(_bp), 5)
    out = np.zeros_like(init)
    R.gmod2_run(hd, out.ctypes.data_as(_bp))
    R.gmod2_destroy(hd)
    assert np.array_equal(out, oracle_mod2(init, 5))
    # heat (the reference's own reductions are not used: fixed step count)
    T0 = rng.uniform(0.0, 100.0, (4, 5, 6))
    rec = np.zeros(4 * 16)
    hd = R.gheat_create(5, 6, 4, T0.ctypes.data_as(_dp), 0.1, 50.0, -1e300, 6, rec.ctypes.data_as(_dp))
    T = np.zeros_like(T0)
    hot = np.zeros(T0.shape, np.uint8)
    cnt = np.zeros(T0.shape, np.int32)
    assert R.gheat_run(hd, T.ctypes.data_as(_dp), hot.ctypes.data_as(_bp), cnt.ctypes.data_as(_ip)) == 6
    R.gheat_destroy(hd)
    eT, ehot, ecnt, _ = oracle_heat(T0, 0.1, 50.0, 6)
    assert np.array_equal(T, eT) and np.array_equal(hot, ehot) and np.array_equal(cnt, ecnt)
    # active cells, both neighbourhoods and boundaries
    for nb in ("von_neumann", "moore"):
        for torus in (0, 1):
            shape, seeds = (5, 6, 7), [(0, 0, 0), (3, 2, 4)]
            exp = oracle_inf(shape, nb, bool(torus), seeds, 3, 7)
            for apply in (False, True):
                got = _run_inf(R, shape, nb, torus, seeds, 3, 7, apply)
                for a, b in zip(got, exp):
                    assert np.array_equal(a, b), (nb, torus, apply)
    # unsafe operations (torus: on a flat box the reference's additions
    # leave its arrays at the border)
    for nb in ("von_neumann", "moore"):
        got = _run_gift(R, (4, 5, 6), nb, 1, 2)
        exp = oracle_gift((4, 5, 6), nb, True, 2)
        for a, b in zip(got, exp):
            assert np.array_equal(a, b), nb


def test_reduction_fold_known_answers():
    r = O.reduce_serial
    n = np.arange(1, 101, dtype=np.int32)
    assert r(n, "i", 3) == 5050
    assert r(np.full(300, 1, np.int8), "b", 3) == 300 - 256      # byte sum wraps
    assert r(np.array([2 ** 20, 2 ** 12], np.int32), "i", 4) == 0   # 2^32 wraps to 0
    assert r(np.array([3, 1, 4, 1, 5], np.int32), "i", 1) == 5
    assert r(np.array([3, 1, 4, 1, 5], np.int32), "i", 2) == 1
    z = r(np.array([-0.0, 0.0, -1.0]), "r", 1)
    assert z == 0.0 and math.copysign(1.0, z) < 0      # the first of equal maxima is kept
    assert math.isnan(r(np.array([math.nan, 1.0, 2.0]), "r", 1))
    assert r(np.array([1.0, math.nan, 2.0]), "r", 1) == 2.0
    assert r(np.array([1, 1, 0, 1], np.int32), "i", 9) == 1        # xor of truth values: parity
    assert r(np.array([6, 3], np.int32), "i", 6) == 2
    assert r(np.array([3e9, -1.0]), "r", 6) == float(-2147483648 & -1)   # out of range -> INT_MIN
    assert r(np.array([7.0]), "r", 3) == 7.0                      # one cell: its value
    x = np.random.default_rng(5).uniform(-1, 1, 1000)
    assert abs(r(x, "r", 3) - math.fsum(x.tolist())) <= O.serial_sum_bound(x)


def test_reduction_vectorised_equals_the_loop():
    rng = np.random.default_rng(17)
    arrays = {
        "b": [rng.integers(-128, 128, 999).astype(np.int8), np.full(700, 3, np.int8)],
        "i": [rng.integers(-2 ** 31, 2 ** 31, 999, dtype=np.int64).astype(np.int32),
              rng.integers(1, 4, 999).astype(np.int32)],
        "r": [rng.uniform(-3, 3, 999), np.array([0.0, -0.0, math.nan, 3e9, -7.5, 2.0, -0.0, 1e10, -2.5e9])],
    }
    for t, lst in arrays.items():
        for v in lst:
            for op in range(11):
                a, b = O.reduce_loop(v, t, op), O.reduce_vectorised(v, t, op)
                same = (a == b and math.copysign(1, a) == math.copysign(1, b)) or (
                    isinstance(a, float) and math.isnan(a) and math.isnan(b))
                assert same, (t, op, a, b)


def _run_inf(L, shape, nb, torus, seeds, span, steps, apply=False):
    s, r, c = shape
    sd = np.array([[i, j, k] for (i, j, k) in seeds], np.int32).reshape(-1)
    hd = L.ginf_create(r, c, s, NB[nb], torus, sd.ctypes.data_as(_ip), len(seeds), span)
    assert hd, lib().calhipGetErrorString()
    try:
        run = L.ginf_apply_steps if apply else L.ginf_steps
        assert run(hd, steps) == 0, lib().calhipGetErrorString()
        sc = np.zeros(shape, np.uint8)
        sn = np.zeros(shape, np.uint8)
        ac = np.zeros(shape, np.int32)
        an = np.zeros(shape, np.int32)
        cells = np.zeros(3 * s * r * c, np.int32)
        cnt = L.ginf_read(hd, sc.ctypes.data_as(_bp), sn.ctypes.data_as(_bp), ac.ctypes.data_as(_ip),
                          an.ctypes.data_as(_ip), cells.ctypes.data_as(_ip), s * r * c)
        assert cnt >= 0
        return sc, sn, ac, an, cells[:3 * cnt].reshape(-1, 3)
    finally:
        L.ginf_destroy(hd)


def _run_gift(L, shape, nb, torus, steps):
    s, r, c = shape
    hd = L.ggift3_create(r, c, s, NB[nb], torus)
    assert hd, lib().calhipGetErrorString()
    try:
        assert L.ggift3_steps(hd, steps) == 0, lib().calhipGetErrorString()
        q = np.zeros(shape, np.int32)
        seen = np.zeros(shape, np.int32)
        acc = np.zeros(shape, np.float64)
        assert L.ggift3_read(hd, q.ctypes.data_as(_ip), seen.ctypes.data_as(_ip), acc.ctypes.data_as(_dp)) == 0
        return q, seen, acc
    finally:
        L.ggift3_destroy(hd)


# ------------------------------------------------------------------ GPU

@pytest.mark.gpu
def test_mod2_gpu():
    print("case 3d mod2", flush=True)
    E = examples()
    rng = np.random.default_rng(1)
    cases = [((7, 5, 6), None, 6), ((9, 12, 10), (4, 6, 5), 9)]
    for shape, seed, steps in cases:
        s, r, c = shape
        if seed is None:
            init = (rng.random(shape) < 0.4).astype(np.uint8)
            exp = oracle_mod2(init, steps)
        else:
            init = np.zeros(shape, np.uint8)
            init[seed] = 1
            exp = mod2_closed_form(shape, seed, steps)
        lib().calhipClearError()
        hd = E.gmod2_create(r, c, s, init.ctypes.data_as(_bp), steps)
        assert hd, lib().calhipGetErrorString()
        out = np.zeros_like(init)
        try:
            assert E.gmod2_run(hd, out.ctypes.data_as(_bp)) == 0, lib().calhipGetErrorString()
        finally:
            E.gmod2_destroy(hd)
        assert np.array_equal(out, exp), shape


@pytest.mark.gpu
def test_heat_with_reductions_gpu():
    print("case 3d heat", flush=True)
    E = examples()
    rng = np.random.default_rng(2)
    T0 = rng.uniform(0.0, 100.0, (6, 7, 8))
    T0[3, 3, 3] = 400.0
    rec = np.zeros(4 * 64)
    lib().calhipClearError()
    hd = E.gheat_create(7, 8, 6, T0.ctypes.data_as(_dp), 0.12, 60.0, 97.0, 60, rec.ctypes.data_as(_dp))
    assert hd, lib().calhipGetErrorString()
    try:
        T = np.zeros_like(T0)
        hot = np.zeros(T0.shape, np.uint8)
        cnt = np.zeros(T0.shape, np.int32)
        steps = E.gheat_run(hd, T.ctypes.data_as(_dp), hot.ctypes.data_as(_bp), cnt.ctypes.data_as(_ip))
        assert steps > 0, lib().calhipGetErrorString()
    finally:
        E.gheat_destroy(hd)
    eT, ehot, ecnt, erec = oracle_heat(T0, 0.12, 60.0, steps)
    assert np.array_equal(T, eT) and np.array_equal(hot, ehot) and np.array_equal(cnt, ecnt)
    # the stop condition (max T <= 97) held first after `steps` steps
    maxes = [float(x.max()) for x in erec]
    assert maxes[-1] <= 97.0 and all(m > 97.0 for m in maxes[:-1]) and steps < 60
    # the steering's reductions, step by step, on the GPU-resident state
    for t in range(steps):
        x = erec[t].reshape(-1)
        exact = math.fsum(x.tolist())
        assert abs(rec[4 * t] - exact) <= 2 * U * abs(exact) + x.size * U * U * float(np.abs(x).sum())
        assert rec[4 * t + 1] == x.max()
    # counters and hot flags of the final step, and their running totals
    assert rec[4 * (steps - 1) + 2] == float(ecnt.sum())
    assert rec[4 * (steps - 1) + 3] == (1.0 if ehot.any() else 0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("nb", ["von_neumann", "moore"])
@pytest.mark.parametrize("torus", [0, 1], ids=["flat", "toroidal"])
def test_active_cells_3d_gpu(nb, torus):
    print(f"case 3d active nb={nb} torus={torus}", flush=True)
    E = examples()
    shape, seeds = (8, 9, 10), [(0, 0, 0), (4, 5, 3), (8, 2, 7)]
    lib().calhipClearError()
    got = _run_inf(E, shape, nb, torus, seeds, 3, 9)
    exp = oracle_inf(shape, nb, bool(torus), seeds, 3, 9)
    names = ["s current", "s next", "age current", "age next", "active list"]
    for a, b, n in zip(got, exp, names):
        assert np.array_equal(a, b), n


@pytest.mark.gpu
@pytest.mark.parametrize("nb", ["von_neumann", "moore"])
def test_active_cells_3d_apply_then_update_gpu(nb):
    """Each process through calApplyElementaryProcess3D followed by
    calUpdate3D: the retiring process writes a cell's next layer and removes
    the cell, so the update (at the cells of the rebuilt set, cal3D.c:677-715;
    the 2D engine's update is the same code) leaves that cell's current layer
    as it was, and both layers equal the oracle's (the reference engine's
    own sequence, pinned against the reference build above)."""
    print(f"case 3d active apply+update nb={nb}", flush=True)
    E = examples()
    shape, seeds = (8, 9, 10), [(0, 0, 0), (4, 5, 3), (8, 2, 7)]
    lib().calhipClearError()
    got = _run_inf(E, shape, nb, 1, seeds, 3, 9, apply=True)
    exp = oracle_inf(shape, nb, True, seeds, 3, 9)
    names = ["s current", "s next", "age current", "age next", "active list"]
    for a, b, n in zip(got, exp, names):
        assert np.array_equal(a, b), n
    assert (got[0] != got[1]).any()   # retired cells: next 2, current kept


@pytest.mark.gpu
@pytest.mark.parametrize("nb", ["von_neumann", "moore"])
@pytest.mark.parametrize("torus", [0, 1], ids=["flat", "toroidal"])
def test_unsafe_operations_3d_gpu(nb, torus):
    print(f"case 3d unsafe nb={nb} torus={torus}", flush=True)
    E = examples()
    shape = (5, 6, 7)
    lib().calhipClearError()
    s, r, c = shape
    hd = E.ggift3_create(r, c, s, NB[nb], torus)
    assert hd, lib().calhipGetErrorString()
    try:
        assert E.ggift3_steps(hd, 3) == 0, lib().calhipGetErrorString()
        q = np.zeros(shape, np.int32)
        seen = np.zeros(shape, np.int32)
        acc = np.zeros(shape, np.float64)
        assert E.ggift3_read(hd, q.ctypes.data_as(_ip), seen.ctypes.data_as(_ip), acc.ctypes.data_as(_dp)) == 0
        eq, eseen, eacc = oracle_gift(shape, nb, bool(torus), 3)
        assert np.array_equal(q, eq) and np.array_equal(seen, eseen)
        assert np.array_equal(acc, eacc)   # multiples of 0.25: exact in any order
        assert E.ggift3_apply_give(hd) == 0, lib().calhipGetErrorString()
        assert E.ggift3_read(hd, q.ctypes.data_as(_ip), seen.ctypes.data_as(_ip), acc.ctypes.data_as(_dp)) == 0
        eq1, _, _ = oracle_gift(shape, nb, bool(torus), 1, apply_only=True)
        assert np.array_equal(q, eq + eq1)
    finally:
        E.ggift3_destroy(hd)


# the operation calReductionComputeProd{2D,3D}{b,i,r} reduces with: the
# reference's wrappers use PROD for bytes and SUM for ints and reals
# (OpenCAL-OMP/source/cal2DReduction.c:563-571, cal3DReduction.c:563-571)
PROD_WRAPPER_OP = {"b": O.OPS.index("PROD"), "i": O.OPS.index("SUM"), "r": O.OPS.index("SUM")}


def test_reference_compute_prod_wrappers():
    """The reference build's calReductionComputeProd3D{b,i,r} on the heat
    model's fields equal its own calReductionOperation3D with PROD for the
    byte variant and SUM for the int and real variants (the behaviour this
    engine reproduces; the reference's reductions themselves cover only part
    of the grid, so they are compared with each other, not with numpy)."""
    R = _ref3()
    if R is None:
        pytest.skip("reference build oracle/_ref/libref_models3d.so not present")
    R.gheat_compute_prod.restype = C.c_double
    R.gheat_compute_prod.argtypes = [C.c_void_p, C.c_int]
    R.gheat_load.argtypes = [C.c_void_p, _dp, _bp, _ip]
    rng = np.random.default_rng(4)
    shape = (3, 4, 5)
    T0 = rng.uniform(0.5, 2.0, shape)
    hot = np.ones(shape, np.uint8)
    cnt = rng.integers(-50, 50, shape).astype(np.int32)
    rec = np.zeros(8)
    hd = R.gheat_create(4, 5, 3, T0.ctypes.data_as(_dp), 0.1, 50.0, -1e300, 1, rec.ctypes.data_as(_dp))
    try:
        assert R.gheat_load(hd, T0.ctypes.data_as(_dp), hot.ctypes.data_as(_bp), cnt.ctypes.data_as(_ip)) == 0
        R.gheat_reduce.restype = C.c_double
        R.gheat_reduce.argtypes = [C.c_void_p, C.c_int, C.c_int]
        SUM, PROD = O.OPS.index("SUM"), O.OPS.index("PROD")
        for which in (0, 2):   # real T, int count
            assert R.gheat_compute_prod(hd, which) == R.gheat_reduce(hd, which, SUM)
        assert R.gheat_reduce(hd, 2, SUM) != R.gheat_reduce(hd, 2, PROD)
        assert R.gheat_compute_prod(hd, 1) == R.gheat_reduce(hd, 1, PROD)
    finally:
        R.gheat_destroy(hd)


def _check_reduction(got, vals, t, op):
    exp = O.reduce_serial(vals, t, op)
    name = O.OPS[op]
    if t == "r" and name == "SUM":
        x = vals.reshape(-1).astype(np.float64)
        if np.all(np.isfinite(x)):
            exact = math.fsum(x.tolist())
            assert abs(got - exact) <= 2 * U * abs(exact) + x.size * U * U * float(np.abs(x).sum()), (got, exact)
            assert abs(got - exp) <= O.serial_sum_bound(x) + 2 * U * abs(exact)
            return
    if t == "r" and isinstance(exp, float) and math.isnan(exp):
        assert math.isnan(got)
        return
    if t == "r" and name == "PROD":
        # a product that overflows or meets a zero depends on the order of the
        # factors (the reference's own order depends on its thread count)
        if math.isfinite(exp) and exp != 0.0:
            assert abs(got - exp) <= vals.size * 2 * U * abs(exp), (got, exp)
        return
    assert got == exp and (t != "r" or math.copysign(1, got) == math.copysign(1, exp)), (name, t, got, exp)


def _reduction_arrays(rng, shape):
    b = rng.integers(-128, 128, shape).astype(np.int8)
    iv = rng.integers(-2 ** 31, 2 ** 31, shape, dtype=np.int64).astype(np.int32)
    r = rng.uniform(-3.0, 3.0, shape)
    return b, iv, r


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(1, 1), (257, 300), (1100, 1030)], ids=["one_cell", "small", "large"])
def test_reductions_2d_gpu(shape):
    print(f"case reductions 2d {shape}", flush=True)
    E = examples()
    rng = np.random.default_rng(shape[0])
    b, iv, r = _reduction_arrays(rng, shape)
    small = rng.integers(1, 3, shape).astype(np.int32)   # products that do not collapse to 0 at once
    edge = r.copy()
    if edge.size >= 4:
        flat = edge.reshape(-1)
        flat[:4] = [-0.0, 0.0, 3e9, math.nan]
    near_one = rng.uniform(0.999, 1.001, shape)   # a product that stays finite
    for arrays in [(b, iv, r), (b, small, edge), (b, small, near_one)]:
        lib().calhipClearError()
        hd = E.gred_create(shape[0], shape[1], arrays[0].ctypes.data_as(_sbp), arrays[1].ctypes.data_as(_ip),
                           arrays[2].ctypes.data_as(_dp))
        assert hd, lib().calhipGetErrorString()
        try:
            for phase in ("host", "device"):
                if phase == "device":
                    assert E.gred_step(hd) == 0, lib().calhipGetErrorString()
                    assert E.gred_on_device(hd) == 1
                else:
                    assert E.gred_on_device(hd) == 0
                for which, t in enumerate("bir"):
                    for op in range(0, 11):
                        got = E.gred_reduce(hd, which, op)
                        assert lib().calhipGetLastError() == 0, lib().calhipGetErrorString()
                        _check_reduction(got, arrays[which], t, op)
                    # calReductionComputeProd2D*: PROD for bytes, SUM for int
                    # and real, as the reference's wrappers
                    _check_reduction(E.gred_compute_prod(hd, which), arrays[which], t, PROD_WRAPPER_OP[t])
        finally:
            E.gred_destroy(hd)


@pytest.mark.gpu
def test_reductions_3d_gpu():
    print("case reductions 3d", flush=True)
    E = examples()
    rng = np.random.default_rng(9)
    shape = (33, 40, 50)
    T0 = rng.uniform(0.0, 100.0, shape)
    rec = np.zeros(4 * 8)
    lib().calhipClearError()
    hd = E.gheat_create(40, 50, 33, T0.ctypes.data_as(_dp), 0.1, 50.0, -1e300, 2, rec.ctypes.data_as(_dp))
    assert hd, lib().calhipGetErrorString()
    try:
        b = rng.integers(0, 2, shape).astype(np.uint8)
        iv = rng.integers(-1000, 1000, shape).astype(np.int32)
        assert E.gheat_load(hd, T0.ctypes.data_as(_dp), b.ctypes.data_as(_bp), iv.ctypes.data_as(_ip)) == 0
        for phase in ("host", "device"):
            if phase == "device":
                assert E.gheat_step(hd) == 0, lib().calhipGetErrorString()
                assert E.gheat_on_device(hd) == 1
                T = np.zeros_like(T0)
                hot = np.zeros(shape, np.uint8)
                cnt = np.zeros(shape, np.int32)
                eT, ehot, ecnt, _ = oracle_heat_from(T0, b, iv, 0.1, 50.0)
                vals = [eT, ehot.astype(np.int8), ecnt]
            else:
                vals = [T0, b.astype(np.int8), iv]
            for which, t in enumerate("rbi"):
                for op in range(0, 11):
                    got = E.gheat_reduce(hd, which, op)
                    assert lib().calhipGetLastError() == 0, lib().calhipGetErrorString()
                    _check_reduction(got, vals[which], t, op)
                _check_reduction(E.gheat_compute_prod(hd, which), vals[which], t, PROD_WRAPPER_OP[t])
    finally:
        E.gheat_destroy(hd)


def oracle_heat_from(T0, hot0, count0, kappa, thr):
    """One heat step from given T, hot and count (numpy form of gheat_diffuse:
    the neighbour terms are added in von Neumann order)."""
    s, r, c = T0.shape
    d = np.zeros_like(T0)
    for (di, dj, dk) in O.VON_NEUMANN[1:]:
        nbv = np.zeros_like(T0)
        valid = np.zeros(T0.shape, bool)
        ks, is_, js = np.meshgrid(np.arange(s), np.arange(r), np.arange(c), indexing="ij")
        kk, ii, jj = ks + dk, is_ + di, js + dj
        valid = (kk >= 0) & (kk < s) & (ii >= 0) & (ii < r) & (jj >= 0) & (jj < c)
        nbv[valid] = T0[kk[valid], ii[valid], jj[valid]]
        d = np.where(valid, d + (nbv - T0), d)
    T = T0 + kappa * d
    hot = (T > thr).astype(np.uint8)
    count = np.where(T > thr, count0 + 1, count0).astype(np.int32)
    return T, hot, count, None


MOD2_REF = os.path.join(native.REF_DIR, "calomp_mod2CA3D")


@pytest.mark.skipif(not os.path.exists(MOD2_REF), reason="calomp_mod2CA3D not built (needs the reference tree)")
def test_reference_mod2_program_matches_closed_form(tmp_path):
    """The reference example cal_mod2CA3D-glut (mod2CA3D.c unchanged, 65^3
    torus, seed at (12, 12, 12), 1000 steps) on the reference 3D engine ends
    in the closed form's state, cell for cell."""
    out = tmp_path / "mod2.bin"
    r = subprocess.run([MOD2_REF, str(out)], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, OMP_NUM_THREADS=os.environ.get("OMP_NUM_THREADS", "8")))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "steps 1000" in r.stdout
    got = np.fromfile(out, np.uint8).reshape(65, 65, 65)
    assert np.array_equal(got, mod2_closed_form((65, 65, 65), (12, 12, 12), 1000))
