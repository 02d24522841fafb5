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


# ------------------------------------------------ D/E form, lossy and Debye media
# fd1d/pycuda/fd1d_2_1.py .. fd1d_2_3.py

from collections import namedtuple  # noqa: E402
import ctypes as C  # noqa: E402

from ._dev import fptr  # noqa: E402

medium1d = namedtuple("medium", ("nax", "nbx", "ncx", "ndx"))
ftrans = namedtuple("ftrans", ("r_pt", "i_pt", "r_in", "i_in"))


def _md1(md: medium1d) -> _lib.FdtdMedium1d:
    return _lib.FdtdMedium1d(fptr(md.nax), fptr(md.nbx), fptr(md.ncx), fptr(md.ndx))


def _ft1(ft: ftrans) -> _lib.FdtdFtrans:
    return _lib.FdtdFtrans(fptr(ft.r_pt), fptr(ft.i_pt), fptr(ft.r_in), fptr(ft.i_in))


def dielectric_dform(nx: int, dt: float, chi: float, tau: float, epsr: float, sigma: float,
                     debye: bool = True, device=None) -> medium1d:
    """nax/nbx(/ncx/ndx) of fd1d/clang/test_2_3.c:75-90 (debye=False:
    test_2_2.c's lossy medium): free space below nx/2."""
    dev = device or require_gpu()
    z = lambda: torch.empty(nx, dtype=torch.float32, device=dev)  # noqa: E731
    md = medium1d(z(), z(), z() if debye else None, z() if debye else None)
    call("fd1d_dielectric_dform", int(nx), float(dt), float(chi), float(tau), float(epsr),
         float(sigma), C.byref(_md1(md)), stream())
    return md


def dxfield(t: int, nx: int, dx, hy, src_pos: int = 1, src_val: float | None = None) -> None:
    """D update and the additive source dx[src_pos] += src_val, default
    gaussian(t, 50, 10) at 1 (fd1d/pycuda/fd1d_2_3.py:96-102)."""
    v = gaussian(t, 50, 10.0) if src_val is None else src_val
    call("fd1d_dxfield", int(t), int(nx), ptr(dx), ptr(hy), int(src_pos), float(v), stream())


def exfield_dform(nx: int, md: medium1d, dx, ix, sx, ex) -> None:
    """E from D through the medium (fd1d/pycuda/fd1d_2_3.py:105-113)."""
    call("fd1d_exfield_dform", int(nx), C.byref(_md1(md)), ptr(dx), ptr(ix), ptr(sx), ptr(ex),
         stream())


def fourier(t: int, nf: int, nx: int, dt: float, freq, ex, ft: ftrans) -> None:
    """Running DFT of ex and of the probe ex[10] (fd1d/pycuda/fd1d_2_3.py:79-93)."""
    fr = (C.c_float * nf)(*[float(f) for f in freq])
    call("fd1d_fourier", int(t), int(nf), int(nx), float(dt), fr, ptr(ex), C.byref(_ft1(ft)),
         stream())


class DFormSolver:
    """The 1-D D/E-form programs (fd1d/pycuda/fd1d_2_2.py, fd1d_2_3.py): a
    pulse added at src_pos striking a lossy (or Debye) half space, with the
    running DFT at freq, through fd1d_dform_run.  src_wave: host values of
    the additive source for steps 1.. (default gaussian(t, 50, 10)).
    fused: the segment-blocked launches (up to 32 steps each, nf <= 3) with
    the partner state set in a work buffer; else the per-step kernels."""

    def __init__(self, nx, md: medium1d, nsteps_max, freq=(), dt=None, src_pos=1,
                 src_wave=None, probes=(), device=None, fused=True):
        dev = device or require_gpu()
        _lib.preload()
        self.nx, self.md, self.device = int(nx), md, dev
        z = lambda: torch.zeros(nx, dtype=torch.float32, device=dev)  # noqa: E731
        self.dx, self.ix, self.ex, self.hy = z(), z(), z(), z()
        self.sx = z() if md.ncx is not None else None
        self.bc = torch.zeros(4, dtype=torch.float32, device=dev)
        if src_wave is None:
            src_wave = [gaussian(t, 50, 10.0) for t in range(1, nsteps_max + 1)]
        self.src_wave = np.ascontiguousarray(np.asarray(src_wave, np.float32))
        self.src_pos = int(src_pos)
        self.freq = np.ascontiguousarray(np.asarray(freq, np.float32))
        self.dt = float(np.float32(np.float32(0.01) / np.float32(6e8)) if dt is None else dt)
        nf = len(self.freq)
        self.ft = (ftrans(torch.zeros((nf, nx), dtype=torch.float32, device=dev),
                          torch.zeros((nf, nx), dtype=torch.float32, device=dev),
                          torch.zeros(nf, dtype=torch.float32, device=dev),
                          torch.zeros(nf, dtype=torch.float32, device=dev)) if nf else None)
        self.probes = [int(k) for k in probes]
        self.probe_len = int(nsteps_max)
        self.probe_out = (torch.zeros((len(probes), self.probe_len), dtype=torch.float32,
                                      device=dev) if probes else None)
        self.t = 0
        n = _lib.lib().fd1d_dform_work_elems(int(nx)) if fused and nf <= 3 else 0
        self.work = torch.empty(n, dtype=torch.float32, device=dev) if n else None

    def desc(self) -> _lib.Fd1dDformDesc:
        d = _lib.Fd1dDformDesc()
        d.nx = self.nx
        for k in ("dx", "ix", "sx", "ex", "hy", "bc"):
            setattr(d, k, fptr(getattr(self, k)))
        d.md = _md1(self.md)
        d.src_pos = self.src_pos
        d.src_wave = self.src_wave.ctypes.data_as(C.POINTER(C.c_float))
        d.wave_len = len(self.src_wave)
        d.nf = len(self.freq)
        d.freq = self.freq.ctypes.data_as(C.POINTER(C.c_float)) if d.nf else None
        d.dt = self.dt
        if self.ft is not None:
            d.ft = _ft1(self.ft)
        d.nprobe = len(self.probes)
        for p, k in enumerate(self.probes):
            d.probe_k[p] = k
        d.probe_out = fptr(self.probe_out)
        d.probe_len = self.probe_len
        d.work = fptr(self.work)
        d.work_elems = 0 if self.work is None else self.work.numel()
        return d

    def run(self, nsteps: int) -> None:
        d = self.desc()
        call("fd1d_dform_run", C.byref(d), int(self.t), int(nsteps), stream())
        self.t += int(nsteps)

    def amplitude_phase(self):
        """(amplt, phase), (nf, nx): the reference's post-processing
        (fd1d/pycuda/fd1d_2_3.py:219-220), host NumPy on the FP32 transforms."""
        r_pt, i_pt = self.ft.r_pt.cpu().numpy(), self.ft.i_pt.cpu().numpy()
        r_in, i_in = self.ft.r_in.cpu().numpy()[:, None], self.ft.i_in.cpu().numpy()[:, None]
        return (1 / np.hypot(r_in, i_in) * np.hypot(r_pt, i_pt),
                np.arctan2(i_pt, r_pt) - np.arctan2(i_in, r_in))


__all__ += ["medium1d", "dielectric_dform", "dxfield", "exfield_dform", "fourier", "DFormSolver"]
