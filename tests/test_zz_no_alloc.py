"""The allocation-free SciddicaT forward (calhipSciddicaTForwardNoAllocDev)
against calhipSciddicaTForwardDev, through the C ABI.

Pins: two calls in a row, the second on a smaller grid (odd column count,
so the other sweep kernel), give the existing entry point's final thickness
and checkpoints bit for bit, including when the step count is odd (the last
step lands in the caller's output from the scratch grid) and when d_h_out is
d_h0; a grid beyond the scratch's 4 Mi cells is refused before any GPU work
with error 51."""
import ctypes as C

import numpy as np
import pytest

from opencal_amd import synth

pytestmark = pytest.mark.gpu


def _lib():
    from opencal_amd import lib
    return lib()


def _p(t):
    return C.c_void_p(t.data_ptr())


def _inputs(rows, cols, seed):
    import torch
    z, h, _ = synth.sciddicat_inputs(rows, cols, seed=seed, patch=min(40, rows // 3, cols // 3))
    z = np.where(h > 0, z - h, z)
    return (torch.tensor(z, dtype=torch.float64, device="cuda"),
            torch.tensor(h, dtype=torch.float64, device="cuda"))


def _both(L, rows, cols, seed, steps, ck, p_r, p_eps):
    import torch
    z, h0 = _inputs(rows, cols, seed)
    n_ck = steps // ck + 1
    ck_a = torch.full((n_ck, rows, cols), -1.0, dtype=torch.float64, device="cuda")
    ck_b = torch.full_like(ck_a, -2.0)
    out_a = torch.empty_like(z)
    out_b = torch.full_like(z, -3.0)
    ms = C.c_double(0.0)
    torch.cuda.synchronize()
    L.calhipClearError()
    assert L.calhipSciddicaTForwardDev(rows, cols, _p(z), _p(h0), p_r, p_eps, steps, ck, _p(ck_a), _p(out_a),
                                       C.byref(ms)) == 0
    assert L.calhipSciddicaTForwardNoAllocDev(rows, cols, _p(z), _p(h0), p_r, p_eps, steps, ck, _p(ck_b),
                                              _p(out_b)) == 0, L.calhipGetErrorString()
    assert L.calhipGetLastError() == 0
    a, b = out_a.cpu().numpy(), out_b.cpu().numpy()
    assert np.array_equal(a, b), f"{rows}x{cols}: max diff {np.abs(a - b).max()}"
    # every checkpoint slot the schedule writes (t % ck == 0, t <= steps)
    used = [k for k in range(n_ck) if k * ck <= steps]
    assert np.array_equal(ck_a[used].cpu().numpy(), ck_b[used].cpu().numpy())
    assert not np.array_equal(a, h0.cpu().numpy())   # the run moved debris
    return a


def test_no_alloc_forward_equals_forward_two_calls():
    L = _lib()
    print("case no-alloc forward 384x512 odd steps, then 97x131", flush=True)
    _both(L, 384, 512, 3, 61, 7, 0.5, 0.001)      # aligned sweep, odd step count
    _both(L, 97, 131, 8, 40, 10, 0.45, 0.003)     # smaller grid, odd columns


def test_no_alloc_forward_in_place_and_limit():
    import torch
    L = _lib()
    print("case no-alloc forward in place 200x256 and size limit", flush=True)
    z, h0 = _inputs(200, 256, 5)
    ms = C.c_double(0.0)
    for steps in (30, 31):   # even: h0's buffer is the first ping-pong buffer; odd: the scratch is
        ref = torch.empty_like(h0)
        ckpt = torch.empty((2, 200, 256), dtype=torch.float64, device="cuda")
        h = h0.clone()
        torch.cuda.synchronize()
        assert L.calhipSciddicaTForwardDev(200, 256, _p(z), _p(h0), 0.5, 0.001, steps, 30, _p(ckpt), _p(ref),
                                           C.byref(ms)) == 0
        assert L.calhipSciddicaTForwardNoAllocDev(200, 256, _p(z), _p(h), 0.5, 0.001, steps, 100, None, _p(h)) == 0
        assert torch.equal(h, ref), steps
    L.calhipClearError()
    assert L.calhipSciddicaTForwardNoAllocDev(4097, 1024, _p(z), _p(h0), 0.5, 0.001, 1, 1, None, _p(h0)) == -1
    assert L.calhipGetLastError() == 51
    L.calhipClearError()
