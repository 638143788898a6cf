"""Host-side drivers of the OpenCAL example models built against the engine
(opencal_amd/examples/*.c), used by tests and benchmarks.  Every call goes
through the C ABI of libopencal_hip.so."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import check, examples, lib, stats

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_bp = C.POINTER(C.c_ubyte)


def _d(a):
    return a.ctypes.data_as(_dp)


class SciddicaT:
    """One SciddicaT model (full grid or active cells) on the GPU.

    z, h: altitude and thickness before the model's init process (the init
    lowers z by h where h > 0, as the reference example's init does)."""

    def __init__(self, z, h, p_r, p_eps, steps, active=False):
        self.z = np.ascontiguousarray(z, np.float64)
        self.h = np.ascontiguousarray(h, np.float64)
        self.rows, self.cols = self.z.shape
        self.E = examples()
        lib().calhipClearError()
        self.handle = self.E.sciddicat_create(self.rows, self.cols, _d(self.z), _d(self.h), float(p_r),
                                              float(p_eps), int(steps), int(bool(active)))
        if not self.handle:
            check(-1, "sciddicat_create")

    @property
    def model(self):
        return self.E.sciddicat_model(self.handle)

    def run(self):
        check(self.E.sciddicat_run(self.handle), "sciddicat_run")
        return self

    def init(self):
        check(self.E.sciddicat_init(self.handle), "sciddicat_init")

    def steps(self, n):
        check(self.E.sciddicat_steps(self.handle, int(n)), "sciddicat_steps")

    def run_steps(self, n):
        """n steps of calRun2D's loop (calhipRunSteps2D): the model with its
        steering twin advances several steps per launch."""
        check(self.E.sciddicat_run_steps(self.handle, int(n)), "sciddicat_run_steps")

    def clear_stop(self):
        """Remove the active variant's stop condition (a host callback), so
        that run_steps hands its steps to the device program together."""
        check(self.E.sciddicat_clear_stop(self.handle), "sciddicat_clear_stop")

    def set_explicit(self, mode):
        """Run the processes one at a time from a global transition of the
        model's own: mode 1 with calUpdate2D after each, mode 2 (full grid)
        with calUpdateSubstate2Dr of the written substates."""
        check(self.E.sciddicat_set_explicit(self.handle, int(mode)), "sciddicat_set_explicit")

    def sync(self):
        check(self.E.sciddicat_sync(self.handle), "sciddicat_sync")

    def read(self):
        oz = np.empty((self.rows, self.cols), np.float64)
        oh = np.empty_like(oz)
        cnt = C.c_int(0)
        check(self.E.sciddicat_read(self.handle, _d(oz), _d(oh), C.byref(cnt)), "sciddicat_read")
        return oz, oh, cnt.value

    def read_flow(self, k):
        out = np.empty((self.rows, self.cols), np.float64)
        check(self.E.sciddicat_read_flow(self.handle, int(k), _d(out)), "sciddicat_read_flow")
        return out

    def stats(self):
        return stats(self.model)

    def set_h(self, i, j, v):
        """A host edit between steps: h = v at (i, j), both layers."""
        check(self.E.sciddicat_set_h(self.handle, int(i), int(j), C.c_double(v)), "sciddicat_set_h")

    def timing(self, enable):
        """1: events around each sweep launch; 2: one event pair around the
        whole region (launches back to back); 0: off."""
        lib().calhipEnableKernelTiming2D(self.model, int(enable))

    def kernel_time(self):
        ms = C.c_double()
        n = C.c_longlong()
        lib().calhipGetKernelTime2D(self.model, C.byref(ms), C.byref(n))
        return ms.value, n.value

    def close(self):
        if self.handle:
            self.E.sciddicat_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sciddicat_run(z, h, p_r, p_eps, steps, active=False):
    """Run `steps` steps; returns (z, h, active_count) after the run."""
    m = SciddicaT(z, h, p_r, p_eps, steps, active)
    try:
        m.run()
        return m.read()
    finally:
        m.close()


class Life:
    def __init__(self, alive0, steps):
        self.a0 = np.ascontiguousarray(alive0, np.uint8)
        self.rows, self.cols = self.a0.shape
        self.E = examples()
        lib().calhipClearError()
        self.handle = self.E.life_create(self.rows, self.cols, self.a0.ctypes.data_as(_bp), int(steps))
        if not self.handle:
            check(-1, "life_create")

    @property
    def model(self):
        return self.E.life_model(self.handle)

    def run(self):
        check(self.E.life_run(self.handle), "life_run")
        return self

    def steps(self, n):
        check(self.E.life_steps(self.handle, int(n)), "life_steps")

    def set_explicit(self, mode):
        """The transition applied alone from the model's own global
        transition, then calUpdateSubstate2Db (mode 1) or calUpdate2D (2)."""
        check(self.E.life_set_explicit(self.handle, int(mode)), "life_set_explicit")

    def read(self):
        out = np.empty((self.rows, self.cols), np.uint8)
        check(self.E.life_read(self.handle, out.ctypes.data_as(_bp)), "life_read")
        return out

    def timing(self, enable):
        lib().calhipEnableKernelTiming2D(self.model, int(enable))

    def kernel_time(self):
        ms = C.c_double()
        n = C.c_longlong()
        lib().calhipGetKernelTime2D(self.model, C.byref(ms), C.byref(n))
        return ms.value, n.value

    def close(self):
        if self.handle:
            self.E.life_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def life_run(alive0, steps):
    m = Life(alive0, steps)
    try:
        m.run()
        return m.read()
    finally:
        m.close()


def sciddicat_batch(instances, steps):
    """Batched ensemble through calhipSciddicaTBatchRun.

    instances: list of (z, h, p_r, p_eps) with z, h before the init process.
    Returns (list of final h, GPU milliseconds of the stepping launches)."""
    from . import lib
    n = len(instances)
    zs = [np.ascontiguousarray(i[0], np.float64) for i in instances]
    hs = [np.ascontiguousarray(i[1], np.float64) for i in instances]
    outs = [np.empty_like(h) for h in hs]
    rows = (C.c_int * n)(*[z.shape[0] for z in zs])
    cols = (C.c_int * n)(*[z.shape[1] for z in zs])
    pr = (C.c_double * n)(*[float(i[2]) for i in instances])
    pe = (C.c_double * n)(*[float(i[3]) for i in instances])
    zp = (_dp * n)(*[_d(z) for z in zs])
    hp = (_dp * n)(*[_d(h) for h in hs])
    op = (_dp * n)(*[_d(o) for o in outs])
    ms = C.c_double(0.0)
    L = lib()
    L.calhipClearError()
    rc = L.calhipSciddicaTBatchRun(n, rows, cols, pr, pe, zp, hp, int(steps), op, C.byref(ms))
    check(rc, "calhipSciddicaTBatchRun")
    return outs, ms.value


class Sciara:
    """One SCIARA-fv2 model (opencal_amd/examples/sciara.c) on the GPU.

    inp: the dict of synth.sciara_inputs (z0, cell_size, params, vents, rates,
    emission_time).  twin=False leaves the steering to the model's host
    callback (every step then reads the border mask on the host)."""

    def __init__(self, inp, steps, active=True, twin=True):
        self.z0 = np.ascontiguousarray(inp["z0"], np.float64)
        self.rows, self.cols = self.z0.shape
        self.params = np.ascontiguousarray(inp["params"], np.float64)
        self.vents = np.ascontiguousarray(inp["vents"], np.int32)
        self.rates = np.ascontiguousarray(inp["rates"], np.float64)
        self.E = examples()
        lib().calhipClearError()
        self.handle = self.E.sciara_create(self.rows, self.cols, _d(self.z0), float(inp["cell_size"]),
                                           _d(self.params), int(self.vents.shape[0]),
                                           self.vents.ctypes.data_as(_ip), int(self.rates.shape[1]),
                                           _d(self.rates), int(inp["emission_time"]), int(steps),
                                           int(bool(active)), 1 if twin else 2)
        if not self.handle:
            check(-1, "sciara_create")
        self.steps_run = None

    @property
    def model(self):
        return self.E.sciara_model(self.handle)

    def run(self):
        r = self.E.sciara_run(self.handle)
        check(-1 if r < 0 else 0, "sciara_run")
        self.steps_run = r
        return self

    def init(self):
        check(self.E.sciara_init(self.handle), "sciara_init")

    def steps(self, n):
        """Up to n steps (calRunCAStep2D each); returns the number run."""
        r = self.E.sciara_steps(self.handle, int(n))
        check(-1 if r < 0 else 0, "sciara_steps")
        return r

    def sync(self):
        check(self.E.sciara_sync(self.handle), "sciara_sync")

    def read(self):
        out = {}
        for k, name in enumerate(("z", "h", "T", "Msl")):
            a = np.empty((self.rows, self.cols), np.float64)
            check(self.E.sciara_read(self.handle, k, _d(a)), "sciara_read")
            out[name] = a
        out["steps"] = self.steps_run
        out["active"] = self.E.sciara_active_count(self.handle)
        out["elapsed"] = self.E.sciara_elapsed(self.handle)
        return out

    def stats(self):
        return stats(self.model)

    def timing(self, enable):
        lib().calhipEnableKernelTiming2D(self.model, int(enable))

    def kernel_time(self):
        ms = C.c_double()
        n = C.c_longlong()
        lib().calhipGetKernelTime2D(self.model, C.byref(ms), C.byref(n))
        return ms.value, n.value

    def close(self):
        if self.handle:
            self.E.sciara_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sciara_run(inp, steps, active=True, twin=True):
    """calRun2D of a SCIARA-fv2 model; returns the dict of Sciara.read()."""
    m = Sciara(inp, steps, active, twin)
    try:
        m.run()
        return m.read()
    finally:
        m.close()
