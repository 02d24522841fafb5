"""fd2d_energy_noalloc: the energy reduction without a caller workspace.

Called through the C-ABI twice in a row, the second time on a smaller
input, and compared with fd2d_energy (which takes the caller's work area)
on the same fields.  Both run the same kernels in the same order, so the
results must be equal bit for bit; they are also checked against the
oracle's FP64 sum.
"""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _fields(n, m, seed, dev):
    g = torch.Generator().manual_seed(seed)
    return [torch.randn((n, m), generator=g, dtype=torch.float32).to(dev) for _ in range(3)]


def _with_work(L, nx, ny, win, ez, hx, hy, dev):
    work = torch.empty(1025, dtype=torch.float64, device=dev)
    out = C.c_double(0.0)
    rc = L.fd2d_energy(nx, ny, *win, ez.data_ptr(), hx.data_ptr(), hy.data_ptr(), work.data_ptr(),
                       C.byref(out), torch.cuda.current_stream().cuda_stream)
    assert rc == 0, L.fdtd_last_error().decode()
    return out.value


def _noalloc(L, nx, ny, win, ez, hx, hy, res):
    rc = L.fd2d_energy_noalloc(nx, ny, *win, ez.data_ptr(), hx.data_ptr(), hy.data_ptr(),
                               res.data_ptr(), torch.cuda.current_stream().cuda_stream)
    assert rc == 0, L.fdtd_last_error().decode()
    torch.cuda.synchronize()
    return float(res.cpu()[0])


def test_energy_noalloc_matches_work_variant(gpu, oracle):
    from simulation_amd import _lib
    L = _lib.lib()
    res = torch.full((1,), -1.0, dtype=torch.float64, device=gpu)
    # first call: a 1031 x 777 grid, whole-grid window; second: a smaller
    # 300 x 211 grid with an interior window
    cases = [((1031, 777), (0, 1031, 0, 777), 1), ((300, 211), (7, 290, 3, 200), 2)]
    for (nx, ny), win, seed in cases:
        print(f"[noalloc] {nx}x{ny} window {win}", flush=True)
        ez, hx, hy = _fields(nx, ny, seed, gpu)
        a = _noalloc(L, nx, ny, win, ez, hx, hy, res)
        b = _with_work(L, nx, ny, win, ez, hx, hy, gpu)
        assert a == b, (a, b)
        i0, i1, j0, j1 = win
        ref = oracle.fd2d_energy(*[np.ascontiguousarray(x.cpu().numpy()[i0:i1, j0:j1])
                                  for x in (ez, hx, hy)])
        assert abs(a - ref) <= 1e-12 * ref


def test_energy_noalloc_rejects_bad_window(gpu):
    from simulation_amd import _lib
    L = _lib.lib()
    ez, hx, hy = _fields(16, 16, 3, gpu)
    res = torch.zeros(1, dtype=torch.float64, device=gpu)
    rc = L.fd2d_energy_noalloc(16, 16, 0, 17, 0, 16, ez.data_ptr(), hx.data_ptr(), hy.data_ptr(),
                               res.data_ptr(), None)
    assert rc == _lib.FDTD_EINVAL
    assert "window" in L.fdtd_last_error().decode()
    rc = L.fd2d_energy_noalloc(16, 16, 0, 16, 0, 16, ez.data_ptr(), hx.data_ptr(), hy.data_ptr(),
                               None, None)
    assert rc == _lib.FDTD_EINVAL
