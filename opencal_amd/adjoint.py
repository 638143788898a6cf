"""SciddicaT adjoint behind torch.autograd.Function.

forward: initial thickness h0 = the source patch placed in a zero grid, the
reference init lowers the altitude under it (z = z_dem - h0 where h0 > 0,
sciddicaT.c:104-130), then `steps` full-grid steps on the GPU, with h written
to a checkpoint every `ck` steps.
backward: the hand-written adjoint kernels (libopencal_hip.so) recompute each
segment from its checkpoint and run the reverse steps, giving dL/dp_r,
dL/dp_eps and dL/dpatch (through both h0 and the lowered altitude).

torch provides device memory and tensors only; every step and every adjoint
step runs in the engine's HIP kernels.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import check, lib


def _ptr(t: torch.Tensor) -> C.c_void_p:
    assert t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
    return C.c_void_p(t.data_ptr())


def _initial_state(z_dem, patch, box):
    r0, c0 = box
    p = patch.shape[0]
    h0 = torch.zeros_like(z_dem)
    h0[r0:r0 + p, c0:c0 + patch.shape[1]] = patch
    z = torch.where(h0 > 0, z_dem - h0, z_dem)
    return z.contiguous(), h0.contiguous()


class SciddicaTRun(torch.autograd.Function):
    """h_T = SciddicaT^steps(h0(patch)); differentiable in p_r, p_eps, patch."""

    @staticmethod
    def forward(ctx, p_r, p_eps, patch, z_dem, box, steps, ck):
        z, h0 = _initial_state(z_dem, patch, box)
        rows, cols = z.shape
        n_ck = steps // ck + 1
        ckpt = torch.empty((n_ck, rows, cols), dtype=torch.float64, device=z.device)
        h_T = torch.empty_like(z)
        torch.cuda.current_stream().synchronize()
        ms = C.c_double(0.0)
        L = lib()
        L.calhipClearError()
        rc = L.calhipSciddicaTForwardDev(rows, cols, _ptr(z), _ptr(h0), float(p_r), float(p_eps), int(steps), int(ck),
                                         _ptr(ckpt), _ptr(h_T), C.byref(ms))
        check(rc, "calhipSciddicaTForwardDev")
        ctx.save_for_backward(z, ckpt, patch)
        ctx.box, ctx.steps, ctx.ck = box, int(steps), int(ck)
        ctx.p = (float(p_r), float(p_eps))
        ctx.forward_ms = ms.value
        return h_T

    @staticmethod
    def backward(ctx, grad_hT):
        z, ckpt, patch = ctx.saved_tensors
        rows, cols = z.shape
        r0, c0 = ctx.box
        pr, pc = patch.shape
        lam = grad_hT.detach().to(torch.float64).contiguous().clone()
        lam_z = torch.zeros_like(lam)
        torch.cuda.current_stream().synchronize()
        d_r, d_e, ms = C.c_double(), C.c_double(), C.c_double()
        L = lib()
        L.calhipClearError()
        rc = L.calhipSciddicaTAdjointBoxDev(rows, cols, _ptr(z), _ptr(ckpt), ctx.p[0], ctx.p[1], ctx.steps, ctx.ck,
                                            _ptr(lam), _ptr(lam_z), C.byref(d_r), C.byref(d_e), r0, c0, pr, pc,
                                            C.byref(ms))
        check(rc, "calhipSciddicaTAdjointBoxDev")
        ctx.backward_ms = ms.value
        lb = lam[r0:r0 + pr, c0:c0 + pc]
        lzb = lam_z[r0:r0 + pr, c0:c0 + pc]
        d_patch = torch.where(patch > 0, lb - lzb, lb)
        dev = z.device
        return (torch.tensor(d_r.value, dtype=torch.float64, device=dev),
                torch.tensor(d_e.value, dtype=torch.float64, device=dev), d_patch, None, None, None, None)


def sciddicat_run(p_r, p_eps, patch, z_dem, box, steps, ck):
    return SciddicaTRun.apply(p_r, p_eps, patch, z_dem, tuple(box), int(steps), int(ck))


def forward_only(z_dem, patch, box, p_r, p_eps, steps):
    """h_T without autograd bookkeeping (one checkpoint slot)."""
    with torch.no_grad():
        return SciddicaTRun.apply(torch.tensor(p_r, dtype=torch.float64), torch.tensor(p_eps, dtype=torch.float64),
                                  patch, z_dem, tuple(box), int(steps), max(int(steps), 1))


def adjoint_raw(z, h0, p_r, p_eps, steps, ck, lam_T):
    """Device adjoint on whole grids (dL/dz for every cell), for checks:
    returns (dL/dh0, dL/dz, dL/dr, dL/deps) for L with dL/dh_T = lam_T."""
    rows, cols = z.shape
    ckpt = torch.empty((steps // ck + 1, rows, cols), dtype=torch.float64, device=z.device)
    h_T = torch.empty_like(z)
    L = lib()
    L.calhipClearError()
    torch.cuda.current_stream().synchronize()
    check(L.calhipSciddicaTForwardDev(rows, cols, _ptr(z), _ptr(h0), float(p_r), float(p_eps), int(steps), int(ck),
                                      _ptr(ckpt), _ptr(h_T), None), "forward")
    lam = lam_T.contiguous().clone()
    lam_z = torch.zeros_like(lam)
    d_r, d_e = C.c_double(), C.c_double()
    check(L.calhipSciddicaTAdjointDev(rows, cols, _ptr(z), _ptr(ckpt), float(p_r), float(p_eps), int(steps), int(ck),
                                      _ptr(lam), _ptr(lam_z), C.byref(d_r), C.byref(d_e), None), "adjoint")
    return lam, lam_z, d_r.value, d_e.value, h_T
