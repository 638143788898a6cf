"""The generic device program: a model's own elementary processes, written
against OpenCAL's cell API, compiled for the GPU at run time and stepped
with OpenCAL's per-process update (opencal_amd/examples/generic_models.c).

CPU: the device API and every example source compile for gfx950; a bad
source is reported with the compiler's message.  GPU: SciddicaT (full grid
and active cells) and Life bit-exact against the oracle; the unsafe
operations (calAddNextX2D*, calGetNext2D*) and calApplyElementaryProcess2D
exact against a restatement in numpy on von Neumann, Moore and both
hexagonal neighbourhoods, flat and toroidal."""
import ctypes as C

import numpy as np
import pytest

import native
from opencal_amd import examples, lib, synth

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_bp = C.POINTER(C.c_ubyte)


def test_example_sources_compile():
    E = examples()
    log = C.create_string_buffer(16384)
    assert E.generic_check_sources(log, 16384) == 0, log.value.decode()


def test_bad_source_reports_the_compiler():
    L = lib()
    L.calhipCheckProcessSource2D.argtypes = [C.c_char_p, C.c_void_p, C.c_int, C.c_char_p, C.c_size_t]
    log = C.create_string_buffer(8192)
    names = (C.c_char_p * 1)(b"f")
    src = b"void f(struct CALModel2D* ca, int i, int j) { calGet2Dr(ca, 0, i); }"
    L.calhipClearError()
    assert L.calhipCheckProcessSource2D(src, names, 1, log, 8192) != 0
    assert b"calGet2Dr" in log.value
    L.calhipClearError()
    bad_name = (C.c_char_p * 1)(b"f; int x")
    assert L.calhipCheckProcessSource2D(b"void f(struct CALModel2D* c, int i, int j) {}", bad_name, 1, log,
                                        8192) != 0
    L.calhipClearError()


def test_processes_that_reach_other_cells_run_in_phases():
    """The generic program runs a process in colour phases when it, or a
    function of the source it calls, uses an operation that reaches another
    cell (calSetX*, calSetCurrentX*, calInitX*, calAddNextX*, calGetNextX*);
    mentions in comments and strings do not count; a use outside every
    function (a macro) marks every process."""
    L = lib()
    L.calhipProcessRunsInPhases.argtypes = [C.c_char_p, C.c_char_p]
    src = b"""
    #include <cal2DUnsafe.h>
    /* calSetX2Dr in a comment */
    static void give(struct CALModel2D* ca, int i, int j, int n, double f)
    {
        calSetX2Dr(ca, Q.h, i, j, n, calGetNextX2Dr(ca, Q.h, i, j, n) + f);
    }
    static double half(double x) { return 0.5 * x; }
    void flows(struct CALModel2D* ca, int i, int j)
    {
        int n;
        for (n = 1; n < ca->sizeof_X; n++) give(ca, i, j, n, half(calGet2Dr(ca, Q.h, i, j)));
    }
    void mark(struct CALModel2D* ca, int i, int j)
    {
        // calAddNextX2Dr(ca, Q.h, i, j, 1, 1.0);
        const char* s = "calInitX2Dr";
        if (calGet2Dr(ca, Q.h, i, j) > 0) calAddActiveCellX2D(ca, i, j, 1);
    }
    void self_only(struct CALModel2D* ca, int i, int j) { calSet2Dr(ca, Q.h, i, j, half(calGetNext2Dr(ca, Q.h, i, j))); }
    void gift3(struct CALModel3D* ca, int i, int j, int k) { calAddNextX3Di(ca, Q.q, i, j, k, 1, 1); }
    """
    assert L.calhipProcessRunsInPhases(src, b"flows") == 1
    assert L.calhipProcessRunsInPhases(src, b"mark") == 0
    assert L.calhipProcessRunsInPhases(src, b"self_only") == 0
    assert L.calhipProcessRunsInPhases(src, b"gift3") == 1
    # a name the source does not define plainly: it may reach any use
    assert L.calhipProcessRunsInPhases(src, b"elsewhere") == 1
    assert L.calhipProcessRunsInPhases(b"void a(struct CALModel2D* c, int i, int j) {}", b"elsewhere") == 0
    macro = src + b"#define GIVE(ca, i, j, n, f) calSetCurrentX2Dr(ca, Q.h, i, j, n, f)\n"
    assert L.calhipProcessRunsInPhases(macro, b"self_only") == 1
    L.calhipClearError()
    assert L.calhipProcessRunsInPhases(src, b"not a name") == -1
    L.calhipClearError()


# ------------------------------------------------------------------- GPU

def _gsct(z, h, p_r, p_eps, steps, active):
    E = examples()
    lib().calhipClearError()
    hd = E.gsct_create(z.shape[0], z.shape[1], z.ctypes.data_as(_dp), h.ctypes.data_as(_dp), p_r, p_eps, steps,
                       int(active))
    assert hd, lib().calhipGetErrorString()
    try:
        assert E.gsct_run(hd) == 0, lib().calhipGetErrorString()
        oz = np.empty_like(z)
        oh = np.empty_like(h)
        cnt = C.c_int(0)
        assert E.gsct_read(hd, oz.ctypes.data_as(_dp), oh.ctypes.data_as(_dp), C.byref(cnt)) == 0
        return oz, oh, cnt.value
    finally:
        E.gsct_destroy(hd)


@pytest.mark.gpu
@pytest.mark.parametrize("active", [False, True])
@pytest.mark.parametrize("rows,cols,steps", [(48, 64, 120), (37, 29, 80)])
def test_generic_sciddicat_bit_exact(rows, cols, steps, active):
    print(f"case generic sciddicaT {rows}x{cols} steps={steps} active={active}", flush=True)
    z, h, _ = synth.sciddicat_inputs(rows, cols, rows + cols, patch=8)
    gz, gh, cnt = _gsct(z, h, 0.5, 0.001, steps, active)
    oz, oh, counts = native.oracle_sciddicat(z, h, 0.5, 0.001, steps, active, counts=True)
    assert np.array_equal(gz, oz)
    assert np.array_equal(gh, oh), f"max |dh| = {np.abs(gh - oh).max()}"
    if active:
        assert cnt == counts[-1]


@pytest.mark.gpu
@pytest.mark.parametrize("rows,cols,steps", [(64, 96, 150), (33, 47, 61)])
def test_generic_life_bit_exact(rows, cols, steps):
    print(f"case generic life {rows}x{cols} steps={steps}", flush=True)
    E = examples()
    a0 = synth.life_inputs(rows, cols, seed=rows + 3 * cols)
    lib().calhipClearError()
    hd = E.glife_create(rows, cols, a0.ctypes.data_as(_bp), steps)
    assert hd, lib().calhipGetErrorString()
    try:
        assert E.glife_run(hd) == 0, lib().calhipGetErrorString()
        got = np.empty_like(a0)
        assert E.glife_read(hd, got.ctypes.data_as(_bp)) == 0
    finally:
        E.glife_destroy(hd)
    assert np.array_equal(got, native.oracle_life(a0, steps))


def _gift_expected(rows, cols, table, nb, torus, steps, apply_only=False):
    """numpy restatement: q_next = q + sum over cells c of w(c) at each in-grid
    neighbour of c; seen = q_next (after give) + q_after_give at neighbour 1;
    acc += 0.25 * gifts.  Hexagonal: odd columns (hex) / odd rows (alt) use
    the second table (n + 7)."""
    q = np.zeros((rows, cols), np.int64)
    seen = np.zeros((rows, cols), np.int64)
    acc = np.zeros((rows, cols), np.float64)
    n_x = len(table)
    size_x = {1: 5, 2: 9, 3: 7, 4: 7}[nb]

    def nbr(i, j, n):
        if (nb == 3 and j % 2 == 1) or (nb == 4 and i % 2 == 1):
            n += 7
        ii, jj = i + table[n][0], j + table[n][1]
        if torus:
            ii = ii + rows if ii < 0 else (ii - rows if ii > rows - 1 else ii)
            jj = jj + cols if jj < 0 else (jj - cols if jj > cols - 1 else jj)
        lin = ii * cols + jj
        return lin if 0 <= lin < rows * cols else None

    assert n_x >= size_x
    for _ in range(steps):
        gift = np.zeros(rows * cols, np.int64)
        for i in range(rows):
            for j in range(cols):
                w = 1 + (i * 7 + j * 3) % 5
                for n in range(1, size_x):
                    t = nbr(i, j, n)
                    if t is not None:
                        gift[t] += w
        q = q + gift.reshape(rows, cols)
        acc = acc + 0.25 * gift.reshape(rows, cols)
        if apply_only:
            continue
        qf = q.reshape(-1)
        new_seen = np.zeros(rows * cols, np.int64)
        for i in range(rows):
            for j in range(cols):
                t = nbr(i, j, 1)
                new_seen[i * cols + j] = qf[i * cols + j] + (qf[t] if t is not None else 0)
        seen = new_seen.reshape(rows, cols)
    return q, seen, acc


@pytest.mark.gpu
@pytest.mark.parametrize("nb", [1, 2, 3, 4], ids=["von_neumann", "moore", "hex", "hex_alt"])
@pytest.mark.parametrize("torus", [0, 1], ids=["flat", "toroidal"])
def test_generic_unsafe_operations(nb, torus):
    print(f"case generic unsafe nb={nb} torus={torus}", flush=True)
    E = examples()
    rows, cols, steps = 13, 18, 3
    lib().calhipClearError()
    hd = E.ggift_create(rows, cols, nb, torus)
    assert hd, lib().calhipGetErrorString()
    try:
        tab = np.zeros(64, np.int32)
        nx = E.ggift_table(hd, tab.ctypes.data_as(_ip), 32)
        table = [(int(tab[2 * k]), int(tab[2 * k + 1])) for k in range(nx)]
        assert E.ggift_steps(hd, steps) == 0, lib().calhipGetErrorString()
        q = np.zeros((rows, cols), np.int32)
        seen = np.zeros((rows, cols), np.int32)
        acc = np.zeros((rows, cols), np.float64)
        assert E.ggift_read(hd, q.ctypes.data_as(_ip), seen.ctypes.data_as(_ip), acc.ctypes.data_as(_dp)) == 0
        eq, eseen, eacc = _gift_expected(rows, cols, table, nb, torus, steps)
        assert np.array_equal(q, eq)
        assert np.array_equal(seen, eseen)
        assert np.array_equal(acc, eacc)   # sums of multiples of 0.25: exact in any order
        # a single process through calApplyElementaryProcess2D + calUpdate2D
        assert E.ggift_apply_give(hd) == 0, lib().calhipGetErrorString()
        assert E.ggift_read(hd, q.ctypes.data_as(_ip), seen.ctypes.data_as(_ip), acc.ctypes.data_as(_dp)) == 0
        eq2, _, eacc2 = _gift_expected(rows, cols, table, nb, torus, steps + 1)
        eq_apply, _, _ = _gift_expected(rows, cols, table, nb, torus, 1, apply_only=True)
        assert np.array_equal(q, eq + eq_apply)
    finally:
        E.ggift_destroy(hd)


@pytest.mark.gpu
@pytest.mark.parametrize("rows,cols,density,seed", [
    (1, 1, 1.0, 1), (3, 5, 0.5, 2), (37, 53, 0.3, 3), (64, 64, 0.0, 4), (64, 64, 1.0, 5),
    (257, 129, 0.001, 6), (1000, 1000, 0.3, 7), (2048, 3001, 0.7, 8), (4096, 4096, 0.02, 9),
])
def test_active_rebuild_row_major(rows, cols, density, seed):
    """The device rebuild of the active list (one launch: ballot ranks per
    wave, look-back prefix across chunks) gives OpenCAL's row-major A.cells
    and A.size_current for seeded flag patterns from empty to full, over
    grids of one chunk to 4096 chunks, on three rebuilds in a row."""
    print(f"case rebuild {rows}x{cols} density={density} seed={seed}", flush=True)
    n = rows * cols
    mark = (synth.uniform(seed, n) < density).astype(np.int32).reshape(rows, cols)
    want = np.flatnonzero(mark)
    E = examples()
    h = E.gmark_create(rows, cols, mark.ctypes.data_as(_ip))
    assert h
    try:
        for rep in range(3):
            assert E.gmark_steps(h, 1) == 0
            got = np.zeros(n + 1, np.int32)
            cnt = C.c_int(0)
            assert E.gmark_read(h, got.ctypes.data_as(_ip), n, C.byref(cnt)) == 0
            assert cnt.value == len(want), f"rebuild {rep}"
            assert np.array_equal(got[:cnt.value], want), f"rebuild {rep}"
    finally:
        E.gmark_destroy(h)
