"""opencal_amd: an MI355X engine for OpenCAL's 2D cellular-automata hot path.

The product is the C-ABI library ``libopencal_hip.so`` (headers in
``include/``): OpenCAL's 2D API with substates resident on the GPU and
hand-written HIP device programs for SciddicaT, Life and SCIARA.  This package
loads it with ctypes for tests, benchmarks and the torch.autograd wrapper of
the SciddicaT adjoint.  It never falls back to a CPU path: when the library is
missing, loading fails.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libopencal_hip.so")
EXAMPLES_PATH = os.path.join(HERE, "libopencal_examples.so")

_lib = None
_ex = None


def _import_torch_first():
    # torch ships its own HIP runtime (same soname as ROCm's); loading torch
    # before the engine makes both use that one runtime in this process.
    try:
        import torch  # noqa: F401
    except Exception:
        pass


def lib() -> C.CDLL:
    """The engine library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
        _import_torch_first()
        _lib = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        _declare(_lib)
    return _lib


def examples() -> C.CDLL:
    """The OpenCAL example models built against the engine."""
    global _ex
    if _ex is None:
        lib()
        if not os.path.exists(EXAMPLES_PATH):
            raise RuntimeError(f"{EXAMPLES_PATH} is not built")
        _ex = C.CDLL(EXAMPLES_PATH)
        _declare_examples(_ex)
    return _ex


class CalhipStats(C.Structure):
    _fields_ = [("steps", C.c_longlong), ("launches", C.c_longlong), ("uploads", C.c_longlong),
                ("downloads", C.c_longlong), ("rebuilds", C.c_longlong), ("materialisations", C.c_longlong),
                ("reruns", C.c_longlong)]


def _declare(L):
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.calhipGetLastError.restype = C.c_int
    L.calhipGetErrorString.restype = C.c_char_p
    L.calCADef2D.restype = vp
    L.calCADef2D.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    L.calFinalize2D.argtypes = [vp]
    L.calhipGetStats2D.argtypes = [vp, C.POINTER(CalhipStats)]
    L.calhipEnableKernelTiming2D.argtypes = [vp, C.c_int]
    L.calhipGetKernelTime2D.argtypes = [vp, dp, C.POINTER(C.c_longlong)]
    L.calhipSynchronize2D.argtypes = [vp]
    L.calhipSyncHost2D.argtypes = [vp]
    L.calhipIsOnDevice2D.argtypes = [vp]
    L.calhipRunSteps2D.argtypes = [vp, C.c_int]
    L.calhipBindSciddicaTSteering2D.argtypes = [vp, vp]
    L.calhipSciddicaTCheckDivision.restype = C.c_longlong
    L.calhipSciddicaTCheckDivision.argtypes = [C.c_longlong, C.c_ulonglong]
    L.calhipSciddicaTBatchRun.argtypes = [C.c_int, ip, ip, dp, dp, C.POINTER(dp), C.POINTER(dp), C.c_int,
                                          C.POINTER(dp), dp]
    L.calhipSciddicaTForwardDev.argtypes = [C.c_int, C.c_int, vp, vp, C.c_double, C.c_double, C.c_int, C.c_int,
                                            vp, vp, dp]
    L.calhipSciddicaTForwardNoAllocDev.argtypes = [C.c_int, C.c_int, vp, vp, C.c_double, C.c_double, C.c_int,
                                                   C.c_int, vp, vp]
    L.calhipSciddicaTAdjointDev.argtypes = [C.c_int, C.c_int, vp, vp, C.c_double, C.c_double, C.c_int, C.c_int,
                                            vp, vp, dp, dp, dp]
    L.calhipSciddicaTAdjointBoxDev.argtypes = [C.c_int, C.c_int, vp, vp, C.c_double, C.c_double, C.c_int, C.c_int,
                                               vp, vp, dp, dp, C.c_int, C.c_int, C.c_int, C.c_int, dp]


def _declare_examples(E):
    vp, dp, ip, bp = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_ubyte)
    E.sciddicat_create.restype = vp
    E.sciddicat_create.argtypes = [C.c_int, C.c_int, dp, dp, C.c_double, C.c_double, C.c_int, C.c_int]
    for name in ("sciddicat_run", "sciddicat_init", "sciddicat_sync"):
        getattr(E, name).argtypes = [vp]
    E.sciddicat_steps.argtypes = [vp, C.c_int]
    E.sciddicat_run_steps.argtypes = [vp, C.c_int]
    E.sciddicat_clear_stop.argtypes = [vp]
    E.sciddicat_read.argtypes = [vp, dp, dp, ip]
    E.sciddicat_set_explicit.argtypes = [vp, C.c_int]
    E.sciddicat_read_flow.argtypes = [vp, C.c_int, dp]
    E.sciddicat_set_h.argtypes = [vp, C.c_int, C.c_int, C.c_double]
    E.sciddicat_model.restype = vp
    E.sciddicat_model.argtypes = [vp]
    E.sciddicat_destroy.argtypes = [vp]
    E.life_create.restype = vp
    E.life_create.argtypes = [C.c_int, C.c_int, bp, C.c_int]
    E.life_run.argtypes = [vp]
    E.life_steps.argtypes = [vp, C.c_int]
    E.life_read.argtypes = [vp, bp]
    E.life_set_explicit.argtypes = [vp, C.c_int]
    E.life_model.restype = vp
    E.life_model.argtypes = [vp]
    E.life_destroy.argtypes = [vp]
    E.sciara_create.restype = vp
    E.sciara_create.argtypes = [C.c_int, C.c_int, dp, C.c_double, dp, C.c_int, ip, C.c_int, dp, C.c_uint,
                                C.c_int, C.c_int, C.c_int]
    E.sciara_run.argtypes = [vp]
    E.sciara_init.argtypes = [vp]
    E.sciara_steps.argtypes = [vp, C.c_int]
    E.sciara_sync.argtypes = [vp]
    E.sciara_read.argtypes = [vp, C.c_int, dp]
    E.sciara_active_count.argtypes = [vp]
    E.sciara_elapsed.restype = C.c_double
    E.sciara_elapsed.argtypes = [vp]
    E.sciara_model.restype = vp
    E.sciara_model.argtypes = [vp]
    E.sciara_destroy.argtypes = [vp]
    # generic device program examples (examples/generic_models.c)
    E.gsct_create.restype = vp
    E.gsct_create.argtypes = [C.c_int, C.c_int, dp, dp, C.c_double, C.c_double, C.c_int, C.c_int]
    E.gsct_run.argtypes = [vp]
    E.gsct_read.argtypes = [vp, dp, dp, ip]
    E.gsct_destroy.argtypes = [vp]
    E.glife_create.restype = vp
    E.glife_create.argtypes = [C.c_int, C.c_int, bp, C.c_int]
    E.glife_run.argtypes = [vp]
    E.glife_read.argtypes = [vp, bp]
    E.glife_destroy.argtypes = [vp]
    E.ggift_create.restype = vp
    E.ggift_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    E.ggift_steps.argtypes = [vp, C.c_int]
    E.ggift_apply_give.argtypes = [vp]
    E.ggift_read.argtypes = [vp, ip, ip, dp]
    E.ggift_table.argtypes = [vp, ip, C.c_int]
    E.ggift_destroy.argtypes = [vp]
    E.gmark_create.restype = vp
    E.gmark_create.argtypes = [C.c_int, C.c_int, ip]
    E.gmark_steps.argtypes = [vp, C.c_int]
    E.gmark_read.argtypes = [vp, ip, C.c_int, ip]
    E.gmark_destroy.argtypes = [vp]
    E.generic_check_sources.argtypes = [C.c_char_p, C.c_size_t]
    sbp = C.POINTER(C.c_byte)
    E.gred_create.restype = vp
    E.gred_create.argtypes = [C.c_int, C.c_int, sbp, ip, dp]
    E.gred_reduce.restype = C.c_double
    E.gred_reduce.argtypes = [vp, C.c_int, C.c_int]
    E.gred_compute_prod.restype = C.c_double
    E.gred_compute_prod.argtypes = [vp, C.c_int]
    E.gred_step.argtypes = [vp]
    E.gred_on_device.argtypes = [vp]
    E.gred_destroy.argtypes = [vp]
    # 3D models (examples/generic_models3d.c)
    E.gmod2_create.restype = vp
    E.gmod2_create.argtypes = [C.c_int, C.c_int, C.c_int, bp, C.c_int]
    E.gmod2_run.argtypes = [vp, bp]
    E.gmod2_destroy.argtypes = [vp]
    E.gheat_create.restype = vp
    E.gheat_create.argtypes = [C.c_int, C.c_int, C.c_int, dp, C.c_double, C.c_double, C.c_double, C.c_int, dp]
    E.gheat_run.argtypes = [vp, dp, bp, ip]
    E.gheat_reduce.restype = C.c_double
    E.gheat_reduce.argtypes = [vp, C.c_int, C.c_int]
    E.gheat_compute_prod.restype = C.c_double
    E.gheat_compute_prod.argtypes = [vp, C.c_int]
    E.gheat_load.argtypes = [vp, dp, bp, ip]
    E.gheat_on_device.argtypes = [vp]
    E.gheat_step.argtypes = [vp]
    E.gheat_destroy.argtypes = [vp]
    E.ginf_create.restype = vp
    E.ginf_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, ip, C.c_int, C.c_int]
    E.ginf_steps.argtypes = [vp, C.c_int]
    E.ginf_apply_steps.argtypes = [vp, C.c_int]
    E.ginf_read.argtypes = [vp, bp, bp, ip, ip, ip, C.c_int]
    E.ginf_destroy.argtypes = [vp]
    E.ggift3_create.restype = vp
    E.ggift3_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    E.ggift3_steps.argtypes = [vp, C.c_int]
    E.ggift3_apply_give.argtypes = [vp]
    E.ggift3_read.argtypes = [vp, ip, ip, dp]
    E.ggift3_table.argtypes = [vp, ip, C.c_int]
    E.ggift3_destroy.argtypes = [vp]
    E.generic3d_check_sources.argtypes = [C.c_char_p, C.c_size_t]


def last_error() -> tuple[int, str]:
    L = lib()
    return L.calhipGetLastError(), L.calhipGetErrorString().decode()


def check(rc=0, what="opencal_amd"):
    code, msg = last_error()
    if rc or code:
        raise RuntimeError(f"{what}: rc={rc} error={code}: {msg}")


def stats(model_ptr) -> dict:
    s = CalhipStats()
    lib().calhipGetStats2D(model_ptr, C.byref(s))
    return {k: getattr(s, k) for k, _ in s._fields_}
