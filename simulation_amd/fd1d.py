"""1-D Ex/Hy FDTD on MI355X -- host mirror of the PyCUDA 1-D programs.

Names, argument order and meaning follow fd1d/pycuda/fd1d_1_x.py: the same
``exfield``/``hyfield`` kernels (here HIP kernels in libfdtd_hip.so called
through ctypes), ``dielectric`` for the lossy ca/cb coefficients, host-side
``gaussian``/``sinusoidal`` waveforms, and fields as 1-D device arrays of
``nx`` cells.  ``run`` replaces the per-step launch loop of ``main()``
(fd1d/pycuda/fd1d_1_2.py:87-89) by one launch that keeps the grid in LDS.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._dev import ptr, require_gpu, stream
from ._lib import SRC_ADD, SRC_HARD, SRC_NONE, call

__all__ = ["gaussian", "sinusoidal", "waveform", "dielectric", "exfield", "hyfield", "run",
           "Fd1dConfig", "simulate", "SRC_NONE", "SRC_HARD", "SRC_ADD"]


def _tdt(dtype):
    return torch.float64 if np.dtype(dtype) == np.float64 else torch.float32


def gaussian(t: int, t0: int, sigma: float, dtype=np.float32):
    """fd1d/pycuda/fd1d_1_2.py:37-40 (device expf there; host libm here)."""
    L = _lib.lib()
    if np.dtype(dtype) == np.float64:
        return np.float64(L.fdtd_gaussian_f64(int(t), int(t0), float(sigma)))
    return np.float32(L.fdtd_gaussian_f32(int(t), int(t0), float(sigma)))


def sinusoidal(t: int, ds: float, freq: float, dtype=np.float32):
    """fd1d/pycuda/fd1d_1_5.py:40-44."""
    L = _lib.lib()
    if np.dtype(dtype) == np.float64:
        return np.float64(L.fdtd_sinusoidal_f64(int(t), float(ds), float(freq)))
    return np.float32(L.fdtd_sinusoidal_f32(int(t), float(ds), float(freq)))


def waveform(kind: str, nsteps: int, t0: int = 0, dtype=np.float32, **kw) -> np.ndarray:
    """Source values of steps t0+1 .. t0+nsteps (host array)."""
    out = np.empty(nsteps, dtype=dtype)
    for s in range(nsteps):
        t = t0 + 1 + s
        if kind == "gaussian":
            out[s] = gaussian(t, kw["center"], kw["spread"], dtype)
        elif kind == "sinusoidal":
            out[s] = sinusoidal(t, kw.get("ds", 0.01), kw["freq"], dtype)
        else:
            raise ValueError(f"unknown waveform {kind!r}")
    return out


def dielectric(nx: int, dt: float, epsr: float, sigma: float, device=None, dtype=np.float32):
    """ca, cb for a lossy half space i >= nx/2 (fd1d/pycuda/fd1d_1_5.py:71-78),
    computed on the device in FP64 and stored in ``dtype``."""
    dev = device or require_gpu()
    td = _tdt(dtype)
    ca = torch.empty(nx, dtype=td, device=dev)
    cb = torch.empty(nx, dtype=td, device=dev)
    name = "fd1d_dielectric_f64" if td == torch.float64 else "fd1d_dielectric"
    call(name, nx, float(dt), float(epsr), float(sigma), ptr(ca, td), ptr(cb, td), stream())
    return ca, cb


def exfield(t: int, nx: int, ca, cb, ex: torch.Tensor, hy: torch.Tensor,
            src_mode: int = SRC_NONE, src_pos: int = 0, src_val: float = 0.0) -> None:
    """One E update + source (fd1d/pycuda/fd1d_1_5.py:49-56; ca None = 1,
    cb None = 0.5 as fd1d/pycuda/fd1d_1_2.py:43-50)."""
    dt = ex.dtype
    name = "fd1d_exfield_f64" if dt == torch.float64 else "fd1d_exfield"
    call(name, int(t), int(nx), ptr(ca, dt), ptr(cb, dt), ptr(ex, dt), ptr(hy, dt),
         int(src_mode), int(src_pos), float(src_val), stream())


def hyfield(nx: int, ex: torch.Tensor, hy: torch.Tensor, bc: torch.Tensor | None) -> None:
    """Two-step boundary + H update (fd1d/pycuda/fd1d_1_2.py:53-61)."""
    dt = ex.dtype
    name = "fd1d_hyfield_f64" if dt == torch.float64 else "fd1d_hyfield"
    call(name, int(nx), ptr(ex, dt), ptr(hy, dt), ptr(bc, dt), stream())


def run(nx: int, nsteps: int, ex, hy, bc=None, ca=None, cb=None, src_mode=SRC_NONE,
        src_pos=0, src_wave: torch.Tensor | None = None, probes=(), t0: int = 0):
    """nsteps steps in one launch; returns probe traces (nprobe, nsteps) on the device."""
    dt = ex.dtype
    pk = np.ascontiguousarray(np.asarray(probes, dtype=np.int32).reshape(-1))
    pout = torch.zeros((len(pk), max(nsteps, 0)), dtype=dt, device=ex.device)
    name = "fd1d_run_f64" if dt == torch.float64 else "fd1d_run"
    pk_c = pk.ctypes.data_as(_lib._IP) if len(pk) else None
    call(name, int(nx), int(t0), int(nsteps), ptr(ca, dt), ptr(cb, dt), ptr(ex, dt), ptr(hy, dt),
         ptr(bc, dt), int(src_mode), int(src_pos), ptr(src_wave, dt), len(pk), pk_c,
         ptr(pout, dt) if len(pk) else None, stream())
    return pout


@dataclass
class Fd1dConfig:
    """A 1-D run: defaults are the ke=200 free-space pulse (hard Gaussian at
    kc=100, t0=40, spread 12, two-step boundary, probe at 150, 500 steps)."""
    nx: int = 200
    ns: int = 500
    dtype: type = np.float32
    source: str = "gaussian"          # or "sinusoidal"
    src_mode: int = SRC_HARD
    src_pos: int | None = None        # default nx//2 (fd1d/pycuda/fd1d_1_2.py:49)
    center: int = 40
    spread: float = 12.0
    freq: float = 700e6
    abc: bool = True
    epsr: float | None = None         # lossy half space (fd1d/pycuda/fd1d_1_5.py)
    sigma: float = 0.0
    probes: tuple = (150,)
    snap_every: int = 100


def simulate(cfg: Fd1dConfig, device=None) -> dict:
    """Run cfg on the GPU; snapshots of ex/hy every snap_every steps."""
    dev = device or require_gpu()
    td = _tdt(cfg.dtype)
    ex = torch.zeros(cfg.nx, dtype=td, device=dev)
    hy = torch.zeros(cfg.nx, dtype=td, device=dev)
    bc = torch.zeros(4, dtype=td, device=dev) if cfg.abc else None
    ca = cb = None
    if cfg.epsr is not None:
        # dt = ds/6e8 as a Python float (fd1d/pycuda/fd1d_1_5.py:92)
        ca, cb = dielectric(cfg.nx, 0.01 / 6e8, cfg.epsr, cfg.sigma, dev, cfg.dtype)
    pos = cfg.nx // 2 if cfg.src_pos is None else cfg.src_pos
    if cfg.source == "gaussian":
        w = waveform("gaussian", cfg.ns, 0, cfg.dtype, center=cfg.center, spread=cfg.spread)
    else:
        w = waveform("sinusoidal", cfg.ns, 0, cfg.dtype, freq=cfg.freq)
    wave = torch.from_numpy(w).to(dev)
    probes = []
    snaps_ex, snaps_hy, snap_t = [], [], []
    fits = cfg.nx <= _lib.lib().fd1d_run_max_nx(np.dtype(cfg.dtype).itemsize)
    t = 0
    seg = cfg.snap_every if cfg.snap_every else cfg.ns
    while t < cfg.ns:
        n = min(seg, cfg.ns - t)
        if fits:
            pr = run(cfg.nx, n, ex, hy, bc, ca, cb, cfg.src_mode, pos, wave[t:t + n].contiguous(),
                     cfg.probes, t0=t)
        else:
            pr = torch.zeros((len(cfg.probes), n), dtype=td, device=dev)
            for s in range(n):
                exfield(t + s + 1, cfg.nx, ca, cb, ex, hy, cfg.src_mode, pos, float(w[t + s]))
                hyfield(cfg.nx, ex, hy, bc)
                for p, k in enumerate(cfg.probes):
                    pr[p, s] = ex[k]
        probes.append(pr)
        t += n
        if cfg.snap_every and t % cfg.snap_every == 0:
            snaps_ex.append(ex.clone())
            snaps_hy.append(hy.clone())
            snap_t.append(t)
    torch.cuda.synchronize()
    out = dict(ex=ex.cpu().numpy(), hy=hy.cpu().numpy(),
               bc=None if bc is None else bc.cpu().numpy(),
               probe=torch.cat(probes, dim=1).cpu().numpy() if probes else None,
               snap_t=snap_t)
    if snaps_ex:
        out["snap_ex"] = torch.stack(snaps_ex).cpu().numpy()
        out["snap_hy"] = torch.stack(snaps_hy).cpu().numpy()
    return out
