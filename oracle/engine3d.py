"""CPU restatement of OpenCAL's 3D engine loop and of the substate
reductions, in plain Python/numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing else) as the
oracle of the GPU engine's 3D models and reductions; the product never
imports, links or runs anything under oracle/.

Follows OpenCAL-OMP:
  * neighbourhoods: von Neumann cal3D.c:29-52, Moore cal3D.c:58-101;
  * calGetX3D*: single toroidal wrap per axis (calCommon.h:62) or the raw
    offset for a flat space (cal3D.c:776-799); the cell index is
    (k*rows + i)*columns + j (cal3DBuffer.h:131-136).  The reference does not
    check the flat boundary; like the engine's device API, a neighbour whose
    linear index falls outside the array reads 0 and takes no write;
  * calApplyElementaryProcess3D (cal3D.c:658-675): every cell in (i, j, k)
    order, or the active list;
  * calUpdate3D (cal3D.c:697-715): active set rebuilt from the flags in
    (i, j, k) order (cal3D.c:319-378), then next -> current for every
    double-layer substate at the listed cells (the whole grid while the list
    is empty, cal3D.c:635-654);
  * the reductions fold `op` over current[0..n) from cell 0 as the serial
    build of cal2DReduction.c:18-520 / cal3DReduction.c:18-530 does, over the
    whole grid (the reference's loop bound `end = rows` is its defect).

Every element type keeps its C width: bytes wrap as int8, ints as int32.
Parity of this restatement is pinned in tests/test_engine3d.py: the mod-2
automaton against its closed form (a product of 1D trinomial parities), heat
diffusion against its conservation law, and every model against the
reference engine itself built from the reference sources (oracle/_ref).
"""
from __future__ import annotations

import math

import numpy as np

VON_NEUMANN = [(0, 0, 0), (-1, 0, 0), (0, -1, 0), (0, 1, 0), (1, 0, 0), (0, 0, -1), (0, 0, 1)]
_MOORE2 = [(0, 0), (-1, 0), (0, -1), (0, 1), (1, 0), (-1, -1), (1, -1), (1, 1), (-1, 1)]
MOORE = [(i, j, k) for k in (0, -1, 1) for (i, j) in _MOORE2]

_DT = {"b": np.int8, "i": np.int32, "r": np.float64}


def _wrap(t: str, v):
    if t == "r":
        return float(v)
    bits = 8 if t == "b" else 32
    v = int(v) & ((1 << bits) - 1)
    return v - (1 << bits) if v >= 1 << (bits - 1) else v


class Model3D:
    """A 3D cellular automaton with the reference engine's semantics."""

    def __init__(self, rows, cols, slices, neighborhood="moore", torus=True, active=False):
        self.rows, self.cols, self.slices = rows, cols, slices
        self.X = list(VON_NEUMANN if neighborhood == "von_neumann" else MOORE)
        self.torus = torus
        self.active = active
        self.n = rows * cols * slices
        self.sub = {}          # name -> (type, [current, next] flat lists)
        self.flags = [0] * self.n
        self.cells = None      # active list of (i, j, k), None when empty

    # ---- set-up
    def add(self, name, t, fill=0):
        v = _wrap(t, fill)
        self.sub[name] = (t, [[v] * self.n, [v] * self.n])

    def lin(self, i, j, k):
        return (k * self.rows + i) * self.cols + j

    def nb(self, i, j, k, n):
        di, dj, dk = self.X[n]
        ii, jj, kk = i + di, j + dj, k + dk
        if self.torus:
            ii = ii + self.rows if ii < 0 else (ii - self.rows if ii > self.rows - 1 else ii)
            jj = jj + self.cols if jj < 0 else (jj - self.cols if jj > self.cols - 1 else jj)
            kk = kk + self.slices if kk < 0 else (kk - self.slices if kk > self.slices - 1 else kk)
        l = (kk * self.rows + ii) * self.cols + jj
        return l if 0 <= l < self.n else None

    # ---- cell API (calGet3D*, calGetX3D*, calSet3D*, ...)
    def get(self, q, i, j, k):
        return self.sub[q][1][0][self.lin(i, j, k)]

    def getx(self, q, i, j, k, n):
        l = self.nb(i, j, k, n)
        return 0 if l is None else self.sub[q][1][0][l]

    def get_next(self, q, i, j, k):
        return self.sub[q][1][1][self.lin(i, j, k)]

    def set(self, q, i, j, k, v):
        t, L = self.sub[q]
        L[1][self.lin(i, j, k)] = _wrap(t, v)

    def add_next_x(self, q, i, j, k, n, v):
        t, L = self.sub[q]
        l = self.nb(i, j, k, n)
        if l is not None:
            L[1][l] = _wrap(t, L[1][l] + v)

    def init_cell(self, q, i, j, k, v):
        t, L = self.sub[q]
        l = self.lin(i, j, k)
        L[0][l] = L[1][l] = _wrap(t, v)

    def add_active(self, i, j, k):
        self.flags[self.lin(i, j, k)] = 1

    def add_active_x(self, i, j, k, n):
        l = self.nb(i, j, k, n)
        if l is not None:
            self.flags[l] = 1

    def remove_active(self, i, j, k):
        self.flags[self.lin(i, j, k)] = 0

    # ---- engine
    def _order(self):
        return [(i, j, k) for i in range(self.rows) for j in range(self.cols) for k in range(self.slices)]

    def apply(self, ep):
        for (i, j, k) in (self.cells if self.cells else self._order()):
            ep(self, i, j, k)

    def update(self):
        if self.active:
            listed = [c for c in self._order() if self.flags[self.lin(*c)]]
            self.cells = listed or None
        for t, L in self.sub.values():
            if self.cells:
                for c in self.cells:
                    l = self.lin(*c)
                    L[0][l] = L[1][l]
            else:
                L[0][:] = L[1][:]

    def transition(self, eps):
        for ep in eps:
            self.apply(ep)
            self.update()

    def array(self, q, layer=0):
        t, L = self.sub[q]
        return np.array(L[layer], dtype=_DT[t]).reshape(self.slices, self.rows, self.cols)


# ------------------------------------------------------------ reductions

OPS = ["NONE", "MAX", "MIN", "SUM", "PROD", "LOGICAL_AND", "BINARY_AND", "LOGICAL_OR", "BINARY_OR",
       "LOGICAL_XOR", "BINARY_XOR"]


def to_cint(x: float) -> int:
    """(CALint)x as x86's truncating conversion: INT_MIN for NaN/out of range."""
    if not (-2147483649.0 < x < 2147483648.0):
        return -2147483648
    return int(x)


def reduce_serial(values, t: str, op: int):
    """The serial fold of calReductionOperation{2,3}D{b,i,r}: acc = v[0];
    acc = op(acc, v[c]) for c = 1..n-1, in C arithmetic of type t.  Large
    inputs take an equivalent vectorised form (reduce_vectorised)."""
    if np.asarray(values).size > 4096:
        return reduce_vectorised(values, t, op)
    return reduce_loop(values, t, op)


def reduce_vectorised(values, t: str, op: int):
    """The serial fold in numpy: modular integer sums/products, order-free
    integer and logical folds, sequential accumulation (ufunc.accumulate) for
    real SUM/PROD, and the first index of the extreme value for MAX/MIN."""
    v = np.asarray(values).reshape(-1)
    name = OPS[op] if 0 <= op < len(OPS) else "NONE"
    if name == "NONE" or v.size == 1:
        return v[0].item()
    if t == "r":
        x = v.astype(np.float64)
        if name in ("MAX", "MIN"):
            if math.isnan(x[0]):
                return float(x[0])
            ok = x[~np.isnan(x)]
            m = ok.max() if name == "MAX" else ok.min()
            return float(x[int(np.argmax(x == m))])
        if name == "SUM":
            return float(np.add.accumulate(x)[-1])
        if name == "PROD":
            with np.errstate(over="ignore", under="ignore", invalid="ignore"):
                return float(np.multiply.accumulate(x)[-1])
        nz = x != 0.0
        if name == "LOGICAL_AND":
            return 1.0 if nz.all() else 0.0
        if name == "LOGICAL_OR":
            return 1.0 if nz.any() else 0.0
        if name == "LOGICAL_XOR":
            return float(int(nz.sum()) & 1)
        ci = np.where((x > -2147483649.0) & (x < 2147483648.0), np.trunc(np.nan_to_num(x)), -2147483648.0)
        ci = ci.astype(np.int64)
        f = {"BINARY_AND": np.bitwise_and, "BINARY_OR": np.bitwise_or, "BINARY_XOR": np.bitwise_xor}[name]
        return float(_wrap("i", int(f.reduce(ci))))
    x = v.astype(np.int64)
    if name == "MAX":
        return int(x.max())
    if name == "MIN":
        return int(x.min())
    if name == "SUM":
        return _wrap(t, int(x.sum()))
    if name == "PROD":
        with np.errstate(over="ignore"):
            return _wrap(t, int(np.multiply.reduce(x.astype(np.uint64)) & np.uint64(0xffffffff)))
    nz = x != 0
    if name == "LOGICAL_AND":
        return 1 if nz.all() else 0
    if name == "LOGICAL_OR":
        return 1 if nz.any() else 0
    if name == "LOGICAL_XOR":
        return int(nz.sum()) & 1
    f = {"BINARY_AND": np.bitwise_and, "BINARY_OR": np.bitwise_or, "BINARY_XOR": np.bitwise_xor}[name]
    return _wrap(t, int(f.reduce(x)))


def reduce_loop(values, t: str, op: int):
    """The serial fold, one cell at a time."""
    v = list(np.asarray(values).reshape(-1).tolist())
    acc = v[0]
    name = OPS[op] if 0 <= op < len(OPS) else "NONE"
    if name == "NONE":
        return acc
    for x in v[1:]:
        if name == "MAX":
            if acc < x:
                acc = x
        elif name == "MIN":
            if acc > x:
                acc = x
        elif name == "SUM":
            acc = _wrap(t, acc + x)
        elif name == "PROD":
            acc = _wrap(t, acc * x)
        elif name == "LOGICAL_AND":
            acc = _wrap(t, 1 if (acc != 0 and x != 0) else 0)
        elif name == "LOGICAL_OR":
            acc = _wrap(t, 1 if (acc != 0 or x != 0) else 0)
        elif name == "LOGICAL_XOR":
            acc = _wrap(t, 1 if ((acc != 0) != (x != 0)) else 0)
        else:
            if t == "r":
                a, b = to_cint(acc), to_cint(x)
            else:
                a, b = int(acc), int(x)
            r = a & b if name == "BINARY_AND" else (a | b if name == "BINARY_OR" else a ^ b)
            acc = float(_wrap("i", r)) if t == "r" else _wrap(t, r)
    return acc


def serial_sum_bound(values) -> float:
    """Error bound of the serial left-to-right sum of n doubles:
    (n-1) u sum|x| (u = 2^-53)."""
    a = np.abs(np.asarray(values, dtype=np.float64)).reshape(-1)
    return (a.size - 1) * 2.0 ** -53 * math.fsum(a.tolist())
