"""Learned absorbing boundary: network, training and the 2-D TM solver that
uses it in place of a PML.

The reference absorbs outgoing waves with a Sullivan PML
(fd2d/pycuda/fd2d_3_2.py:126-135, 219-260); here a small network (108 -> 64
-> 64 -> 3, GELU, bf16 weights, FP32 accumulation on MFMA) predicts the next
ez of the outer three cells of every edge from the last four canonical
boundary samples (27 values each: ez, normal- and tangential-coupled H over
3 depths x 3 tangential offsets; all four edges share one network through
the exact mirror/transpose symmetries of the Yee update).

Training data come from wide-PML runs of the plain solver (npml=64): the
boundary samples of an inner window are recorded every step and the network
learns the window's next-step edge ez from them.  Forward, backward and the
Adam update are HIP kernels of libfdtd_hip.so (lbc_*); PyTorch only holds the
device buffers.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, fd2d
from ._dev import ptr, require_gpu, stream
from ._lib import call

IN, HID, OUT, NSAMP = 108, 64, 3, 27


def sidx(f: int, d: int, o: int) -> int:
    """Index of field f (0 ez, 1 normal-coupled H, 2 tangential-coupled H),
    depth d and tangential offset o in a 27-value sample."""
    return (f * 3 + d) * 3 + (o + 1)


def npos(nx: int, ny: int) -> int:
    return 2 * ny + 2 * (nx - 6)


def init_params(seed: int = 0) -> np.ndarray:
    """Fan-in scaled uniform weights, zero biases, a small last layer."""
    rng = np.random.default_rng(seed)
    w1 = rng.uniform(-1, 1, (HID, IN)) * math.sqrt(3.0 / IN)
    w2 = rng.uniform(-1, 1, (HID, HID)) * math.sqrt(3.0 / HID)
    w3 = rng.uniform(-1, 1, (OUT, HID)) * math.sqrt(3.0 / HID) * 0.1
    return np.concatenate([w1.ravel(), np.zeros(HID), w2.ravel(), np.zeros(HID), w3.ravel(),
                           np.zeros(OUT)]).astype(np.float32)


class Net:
    """Network parameters on the device: FP32 master copy, Adam moments and
    the bf16 operand buffer the kernels read."""

    def __init__(self, params=None, device=None, seed: int = 0):
        dev = device or require_gpu()
        self.device = dev
        L = _lib.lib()
        self.nparam = L.lbc_param_count()
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        if p.shape != (self.nparam,):
            raise ValueError(f"expected {self.nparam} parameters, got {p.shape}")
        self.params = torch.from_numpy(p.copy()).to(dev)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.wbuf = torch.zeros(L.lbc_wbuf_elems(), dtype=torch.int16, device=dev)
        self.step = 0
        self.pack()

    def pack(self) -> None:
        call("lbc_pack", ptr(self.params), self.wbuf.data_ptr(), stream())

    def numpy(self) -> np.ndarray:
        return self.params.cpu().numpy()

    def forward(self, src: torch.Tensor, offs: torch.Tensor, tstride: int):
        """Normalised outputs y (n, 3) and scales s (n) for rows whose sample k
        starts at src[offs[r] - k * tstride]."""
        n = offs.numel()
        y = torch.empty((n, OUT), dtype=torch.float32, device=self.device)
        s = torch.empty(n, dtype=torch.float32, device=self.device)
        call("lbc_forward", n, ptr(src), offs.data_ptr(), int(tstride), ptr(self.params),
             self.wbuf.data_ptr(), ptr(y), ptr(s), stream())
        return y, s

    def grad(self, src, offs, tstride, tgt=None, norm=None, nblocks: int = 0):
        """(loss, gradient) over the rows (deterministic reduction)."""
        n = offs.numel()
        nb = nblocks or max(1, min(512, (n + 255) // 256))
        part = torch.empty((nb, self.nparam), dtype=torch.float32, device=self.device)
        ploss = torch.empty(nb, dtype=torch.float32, device=self.device)
        g = torch.empty(self.nparam, dtype=torch.float32, device=self.device)
        nrm = float(norm if norm is not None else n)
        call("lbc_train_grad", n, ptr(src), offs.data_ptr(), int(tstride),
             None if tgt is None else ptr(tgt), nrm, ptr(self.params), self.wbuf.data_ptr(), nb,
             ptr(part), ptr(ploss), stream())
        call("lbc_adam", nb, ptr(part), None, None, None, 0, 0.0, 0.9, 0.999, 1e-8, None, ptr(g),
             stream())
        return float(ploss.double().sum()) / nrm, g

    def train_step(self, src, offs, tstride, lr: float, tgt=None, nblocks: int = 0,
                   part=None, ploss=None) -> torch.Tensor:
        """One Adam step on the rows; returns the device loss partials."""
        n = offs.numel()
        nb = nblocks or max(1, min(512, (n + 255) // 256))
        if part is None or part.shape[0] < nb:
            part = torch.empty((nb, self.nparam), dtype=torch.float32, device=self.device)
        if ploss is None or ploss.numel() < nb:
            ploss = torch.empty(nb, dtype=torch.float32, device=self.device)
        self.step += 1
        call("lbc_train_grad", n, ptr(src), offs.data_ptr(), int(tstride),
             None if tgt is None else ptr(tgt), float(n), ptr(self.params), self.wbuf.data_ptr(),
             nb, ptr(part), ptr(ploss), stream())
        call("lbc_adam", nb, ptr(part), ptr(self.params), ptr(self.m), ptr(self.v), self.step,
             float(lr), 0.9, 0.999, 1e-8, self.wbuf.data_ptr(), None, stream())
        return ploss[:nb]


def sample(ez: torch.Tensor, hx: torch.Tensor, hy: torch.Tensor, out: torch.Tensor | None = None,
           window=None) -> torch.Tensor:
    """Canonical boundary samples (npos, 27) of ez/hx/hy (or of the window
    (i0, i1, j0, j1) of them)."""
    nxg, nyg = ez.shape
    i0, i1, j0, j1 = window if window is not None else (0, nxg, 0, nyg)
    nx, ny = i1 - i0, j1 - j0
    if out is None:
        out = torch.empty((npos(nx, ny), NSAMP), dtype=torch.float32, device=ez.device)
    off = i0 * nyg + j0
    call("lbc_sample", nx, ny, nyg, ptr(ez) + 4 * off, ptr(hx) + 4 * off, ptr(hy) + 4 * off,
         ptr(out), stream())
    return out


class LearnedTMSolver:
    """2-D TM time stepping with the learned boundary in place of a PML.

    The grid has no PML: the blocked TM kernel runs with PEC edges and a frame
    kernel recomputes the outer cells with the network applied every step
    (fd2d_lbc_run).  md: fd2d.medium (nbz None = lossless); sources
    [(i, j, wave)] and probes [(i, j)] at least fd2d_lbc_margin() cells from
    every edge.  hist holds the last four boundary samples of every position
    (zero for fields at rest)."""

    def __init__(self, nx, ny, net: Net, md, sources=(), probes=(), nsteps_max=None,
                 depth=0, rows=0, cols=0, device=None):
        self.tm = fd2d.TMSolver(nx, ny, None, md, sources, probes, nsteps_max, depth, rows, cols,
                                device)
        self.net = net
        n = _lib.lib().lbc_hist_elems(nx, ny)
        dev = self.tm.device
        self.hist = torch.zeros(n, dtype=torch.float32, device=dev)
        self.hist2 = torch.zeros(n, dtype=torch.float32, device=dev)

    @property
    def t(self) -> int:
        return self.tm.t

    def desc(self):
        return _lib.Fd2dLbcDesc(self.net.params.data_ptr(), self.net.wbuf.data_ptr(),
                                self.hist.data_ptr(), self.hist2.data_ptr())

    def run(self, nsteps: int) -> None:
        d = self.tm.desc()
        b = self.desc()
        call("fd2d_lbc_run", C.byref(d), C.byref(b), int(self.tm.t), int(nsteps), stream())
        self.tm.t += int(nsteps)

    def state(self) -> dict:
        return self.tm.state()

    def probe_traces(self):
        return self.tm.probe_traces()

    def energy(self, window=None) -> float:
        return self.tm.energy(window)

    def history(self) -> np.ndarray:
        """(npos, 4, 27) boundary samples, newest first."""
        torch.cuda.synchronize()
        return self.hist.view(-1, 4, NSAMP).cpu().numpy()
