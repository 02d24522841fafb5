"""Learned absorbing boundary: network, training and the 2-D TM solver that
uses it in place of a PML.

The reference absorbs outgoing waves with a Sullivan PML
(fd2d/pycuda/fd2d_3_2.py:126-135, 219-260); here a small network (108 -> 64
-> 64 -> 3, GELU, bf16 weights, FP32 accumulation on MFMA) predicts the next
ez of the outer three cells of every edge from the last four canonical
boundary samples (27 features each: ez, normal- and tangential-coupled H over
grid depths 3..5 -- just inside the predicted cells -- x 3 tangential
offsets; all four edges share one network through the exact mirror/transpose
symmetries of the Yee update), as an odd correction, y(-x) = -y(x), to the
reference's two-step boundary (fd1d/pycuda/fd1d_1_2.py:53-61) applied to the
three outer layers.

Training data come from wide-PML runs of the plain solver (npml=64): the
boundary samples of an inner window are recorded every step and the network
learns the window's next-step edge ez from them.  Forward, backward and the
Adam update are HIP kernels of libfdtd_hip.so (lbc_*); PyTorch only holds the
device buffers.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, fd2d
from ._dev import ptr, require_gpu, stream
from ._lib import call

IN, HID, OUT = 108, 64, 3
NFEAT = 27          # features per sample
NSAMP = 30          # stored sample: features, then ez of grid depths 2, 1, 0


def sidx(f: int, d: int, o: int) -> int:
    """Index of field f (0 ez, 1 normal-coupled H, 2 tangential-coupled H),
    feature depth d (grid depth d + 3) and tangential offset o in a sample."""
    return (f * 3 + d) * 3 + (o + 1)


def npos(nx: int, ny: int) -> int:
    return 2 * ny + 2 * (nx - 6)


# flat parameter layout: W1 (64x108) b1 W2 (64x64) b2 W3 (3x64) b3 L (3x108)
P_W3 = HID * IN + HID + HID * HID + HID
P_L = P_W3 + OUT * HID + OUT
NPARAM = P_L + OUT * IN


def init_params(seed: int = 0) -> np.ndarray:
    """Fan-in scaled uniform weights, zero biases, a small last layer."""
    rng = np.random.default_rng(seed)
    w1 = rng.uniform(-1, 1, (HID, IN)) * math.sqrt(3.0 / IN)
    w2 = rng.uniform(-1, 1, (HID, HID)) * math.sqrt(3.0 / HID)
    w3 = rng.uniform(-1, 1, (OUT, HID)) * math.sqrt(3.0 / HID) * 0.1
    return np.concatenate([w1.ravel(), np.zeros(HID), w2.ravel(), np.zeros(HID), w3.ravel(),
                           np.zeros(OUT), np.zeros(OUT * IN)]).astype(np.float32)


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

    def set_params(self, p) -> None:
        """New FP32 parameters (Adam moments reset, operands repacked)."""
        self.params.copy_(torch.from_numpy(np.asarray(p, np.float32)))
        self.m.zero_()
        self.v.zero_()
        self.step = 0
        self.pack()

    def gram(self, src, offs, tstride, nblocks: int = 0, round_bf16: bool = False, tgt=None,
             wpow: float = 0.0):
        """Per-block normal-equation sums of the linear path over the rows
        (lbc_gram; targets tgt (n, 3) or the next samples; row weights
        s^wpow): a (nblocks, lbc_gram_elems()) FP64 tensor for fit_linear."""
        n = offs.numel()
        nb = nblocks or max(1, min(2048, (n + 2047) // 2048))
        L = _lib.lib()
        part = torch.empty((nb, L.lbc_gram_elems()), dtype=torch.float64, device=self.device)
        call("lbc_gram", n, ptr(src), offs.data_ptr(), int(tstride),
             None if tgt is None else ptr(tgt), int(round_bf16), float(wpow), nb,
             ptr(part, torch.float64),
             stream())
        return part

    def fit_linear(self, part: torch.Tensor, rho: float) -> torch.Tensor:
        """Ridge solution of the linear path from gram() sums (FP64 Cholesky
        on the GPU), written into the parameters; returns [G | R | count]."""
        L = _lib.lib()
        g = torch.empty(L.lbc_gram_elems(), dtype=torch.float64, device=self.device)
        st = torch.zeros(1, dtype=torch.int32, device=self.device)
        call("lbc_fit_linear", part.shape[0], ptr(part, torch.float64), ptr(g, torch.float64),
             float(rho), ptr(self.params), st.data_ptr(), stream())
        bad = int(st.item())
        if bad:
            raise _lib.FdtdError("lbc_fit_linear", bad,
                                 f"normal equations numerically singular at pivot {bad - 1}; "
                                 f"raise rho")
        self.pack()
        return g

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
             ptr(part), ptr(ploss), 0.0, 0, stream())
        call("lbc_adam", nb, ptr(part), None, None, None, 0, 0.0, 0.9, 0.999, 1e-8, None, ptr(g),
             0, 0, stream())
        return float(ploss.double().sum()) / nrm, g

    def train_step(self, src, offs, tstride, lr: float, tgt=None, nblocks: int = 0,
                   part=None, ploss=None, noise: float = 0.0, only=None) -> torch.Tensor:
        """One Adam step on the rows (inputs perturbed by `noise`, a fresh
        pattern each step; only = (first, count) restricts the update to those
        parameters); returns the device loss partials."""
        n = offs.numel()
        nb = nblocks or max(1, min(512, (n + 255) // 256))
        if part is None or part.shape[0] < nb:
            part = torch.empty((nb, self.nparam), dtype=torch.float32, device=self.device)
        if ploss is None or ploss.numel() < nb:
            ploss = torch.empty(nb, dtype=torch.float32, device=self.device)
        self.step += 1
        call("lbc_train_grad", n, ptr(src), offs.data_ptr(), int(tstride),
             None if tgt is None else ptr(tgt), float(n), ptr(self.params), self.wbuf.data_ptr(),
             nb, ptr(part), ptr(ploss), float(noise), self.step, stream())
        q0, qn = only or (0, 0)
        call("lbc_adam", nb, ptr(part), ptr(self.params), ptr(self.m), ptr(self.v), self.step,
             float(lr), 0.9, 0.999, 1e-8, self.wbuf.data_ptr(), None, int(q0), int(qn), stream())
        return ploss[:nb]


def sample(ez: torch.Tensor, hx: torch.Tensor, hy: torch.Tensor, out: torch.Tensor | None = None,
           window=None) -> torch.Tensor:
    """Canonical boundary samples (npos, 30) of ez/hx/hy (or of the window
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
        wb = _lib.lib().lbc_work_bytes(nx, ny)
        self.work = torch.empty(max(1, wb), dtype=torch.uint8, device=dev)

    @property
    def t(self) -> int:
        return self.tm.t

    def desc(self):
        return _lib.Fd2dLbcDesc(self.net.params.data_ptr(), self.net.wbuf.data_ptr(),
                                self.hist.data_ptr(), self.hist2.data_ptr(),
                                self.work.data_ptr())

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
        """(npos, 4, 30) boundary samples, newest first."""
        torch.cuda.synchronize()
        return self.hist.view(-1, 4, NSAMP).cpu().numpy()


# ------------------------------------------------------------ PML fallback

@dataclass
class FallbackPolicy:
    """When a guarded learned-boundary run hands over to a PML.

    Every `every` steps the FP64 energy of the grid (fd2d_energy) is taken.
    The run hands over when that energy is not finite, or -- from step
    `quiet_after` on (when the sources have stopped injecting), once the
    largest energy seen exceeds `floor` -- when it exceeds `growth` times the
    largest energy seen at earlier checks: an absorbing boundary only removes
    energy, so growth past that factor means the boundary went unstable.
    `force_at`: hand over at the first check at or after that step whatever
    the energy (None: never).  The PML it hands over to is the reference's
    Sullivan layer (fd2d/pycuda/fd2d_3_2.py:126-135) of `npml` cells, placed
    outside the window as in the npml=32 comparison form."""
    every: int = 64
    growth: float = 4.0
    quiet_after: int = 0
    floor: float = 0.0
    npml: int = 32
    force_at: int | None = None


def fallback_reason(policy: FallbackPolicy, energy: float, emax: float, t: int) -> str | None:
    """Why a check at step t with this energy (largest earlier: emax) hands
    over, or None to carry on."""
    if policy.force_at is not None and t >= policy.force_at:
        return f"handover forced at step {t}"
    if not math.isfinite(energy):
        return f"energy {energy} at step {t}"
    if t >= policy.quiet_after and emax > policy.floor and energy > policy.growth * emax:
        return f"energy {energy:.3e} at step {t} above {policy.growth} x {emax:.3e}"
    return None


class GuardedLearnedSolver:
    """A learned-boundary run (LearnedTMSolver on an n x m window) with the
    plain PML kept as fallback.

    run() advances in chunks of policy.every steps; before each chunk the
    state (dz, iz, ez, hx, hy, the boundary-sample history, probe traces) is
    copied on the device, after it the energy is checked (fallback_reason).
    On a handover the state before the failing chunk is restored and copied
    into the centre of an (n + 2 npml) x (m + 2 npml) grid whose outer npml
    cells are a Sullivan PML (pmlparam), with the window's medium, free space
    around it, zero fields outside the window and zero PML accumulators; the
    run continues there (a TMSolver), from the same step, with the sources
    and probes shifted by npml.  `events` records each handover (step and
    reason); `pml` is the PML solver once handed over."""

    def __init__(self, nx, ny, net: Net, md, sources=(), probes=(), nsteps_max=None,
                 policy: FallbackPolicy | None = None, device=None, **kw):
        self.policy = policy or FallbackPolicy()
        self.learned = LearnedTMSolver(nx, ny, net, md, sources, probes, nsteps_max, device=device,
                                       **kw)
        self.nx, self.ny = int(nx), int(ny)
        self.md = md
        self.kw = kw
        self.pml = None
        self.events = []
        self.emax = 0.0
        self.energies = []

    @property
    def t(self) -> int:
        return self.pml.t if self.pml is not None else self.learned.t

    @property
    def active(self):
        """The solver currently advancing: the learned one, or the PML one."""
        return self.pml if self.pml is not None else self.learned

    def _snapshot(self):
        f = self.learned.tm
        keep = {k: getattr(f, k).clone() for k in ("dz", "iz", "ez", "hx", "hy")
                if getattr(f, k) is not None}
        keep["hist"] = self.learned.hist.clone()
        keep["hist2"] = self.learned.hist2.clone()
        if f.probe_out is not None:
            keep["probe_out"] = f.probe_out.clone()
        return f.t, keep

    def _restore(self, snap):
        t, keep = snap
        f = self.learned.tm
        for k, v in keep.items():
            if k in ("hist", "hist2"):
                getattr(self.learned, k).copy_(v)
            else:
                getattr(f, k).copy_(v)
        f.t = t

    def handover_state(self):
        """The learned run's current state as the PML grid's initial state:
        (dict of (N, M) arrays dz iz ez hx hy ihx ihy, step, pad)."""
        f = self.learned.tm
        npml = self.policy.npml
        N, M = self.nx + 2 * npml, self.ny + 2 * npml
        dev = f.device
        out = {}
        for k in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy"):
            src = getattr(f, k, None)
            if k == "iz" and src is None:
                out[k] = None
                continue
            big = torch.zeros((N, M), dtype=torch.float32, device=dev)
            if src is not None:
                big[npml:npml + self.nx, npml:npml + self.ny].copy_(src)
            out[k] = big
        return out, f.t, npml

    def _hand_over(self, reason: str) -> None:
        f = self.learned.tm
        state, t, npml = self.handover_state()
        N, M = self.nx + 2 * npml, self.ny + 2 * npml
        dev = f.device
        pml = fd2d.new_pmlayer(N, M, dev)
        fd2d.pmlparam(N, M, npml, pml)
        naz = torch.ones((N, M), dtype=torch.float32, device=dev)
        naz[npml:npml + self.nx, npml:npml + self.ny].copy_(self.md.naz)
        nbz = None
        if self.md.nbz is not None:
            nbz = torch.zeros((N, M), dtype=torch.float32, device=dev)
            nbz[npml:npml + self.nx, npml:npml + self.ny].copy_(self.md.nbz)
        srcs = []
        if f.sources:
            waves = f.src_wave.cpu().numpy()
            srcs = [(i + npml, j + npml, waves[k]) for k, (i, j) in enumerate(f.sources)]
        probes = [(i + npml, j + npml) for i, j in f.probes]
        s = fd2d.TMSolver(N, M, pml, fd2d.medium(naz, nbz), srcs, probes,
                          f.probe_len or None, device=dev, **self.kw)
        for k, v in state.items():
            if v is not None and getattr(s, k) is not None:
                getattr(s, k).copy_(v)
        if f.probe_out is not None:
            s.probe_out[:, :t].copy_(f.probe_out[:, :t])
        s.t = t
        self.pml = s
        self.events.append({"step": t, "reason": reason, "npml": npml})

    def check(self) -> str | None:
        """Energy check of the learned run at its current step."""
        f = self.learned.tm
        e = f.energy()
        self.energies.append((f.t, e))
        reason = fallback_reason(self.policy, e, self.emax, f.t)
        if math.isfinite(e):
            self.emax = max(self.emax, e)
        return reason

    def run(self, nsteps: int) -> None:
        left = int(nsteps)
        while left > 0 and self.pml is None:
            n = min(self.policy.every, left)
            if left - n == 1:
                n = left                  # a learned run advances at least two steps
            snap = self._snapshot()
            self.learned.run(n)
            reason = self.check()
            if reason is not None:
                self._restore(snap)
                self._hand_over(reason)
                break
            left -= n
        if self.pml is not None:
            # whatever the learned run did not finish (the failing chunk
            # included) runs on the PML form
            self.pml.run(left)

    def window(self, k: str) -> torch.Tensor:
        """Field k on the n x m window, whichever form holds it."""
        if self.pml is None:
            return getattr(self.learned.tm, k)
        p = self.policy.npml
        return getattr(self.pml, k)[p:p + self.nx, p:p + self.ny]

    def probe_traces(self):
        return self.active.probe_traces()


# ----------------------------------------------------------------- training

@dataclass
class TrainConfig:
    """Wide-PML training runs and the optimisation.

    Each run is an N x N grid, N = window + 2 (npml + gap), with an npml=64
    Sullivan PML; the learned boundary is trained for the inner window (its
    edges gap cells clear of the PML).  Every run starts from a point pulse --
    dz = exp(-r^2 / (2 w^2)), w U[1.5, 4] cells, all other fields zero -- at
    least `margin` cells inside the window, with 0-3 lossy cylinders (eps_r
    U[1,8], sigma U[0,0.05]) at least 10 cells clear of the window edge.  A
    pulse rather than the reference's hard source (fd2d/pycuda/fd2d_3_1.py:69):
    a hard source pins ez at its cell and leaves a static vortex of H around
    it, a field that never radiates and whose far part depends on how the grid
    is closed -- not what an absorbing boundary is for.  The window's boundary
    samples are recorded every step."""
    runs: int = 256
    window: int = 160
    npml: int = 64
    gap: int = 8
    steps: int = 520
    margin: int = 32
    seed: int = 0
    fit_rows: int = 6_000_000     # rows (random subset) of the linear path's normal equations
    wpow: tuple = (0.0, 1.0, 2.0)            # row-weight exponents s^wpow tried for the fit
    rho: tuple = (1e-7, 3e-7, 1e-6, 1e-5)    # ridge candidates (x weight sum)
    select_n: int = 2048          # closed-loop validation window for the fit and the hidden scale
    scales: tuple = (0.0, 0.25, 0.5, 1.0)   # hidden-layer output scales tried after training
    train_steps: int = 8000       # Adam steps of the hidden layers (linear path fixed)
    batch: int = 8192
    lr: float = 2e-3
    lr_min: float = 2e-5
    noise: float = 0.0            # input noise amplitude during training
    s_floor: float = 1e-6         # rows whose max|feature| is below this are dropped


def training_specs(cfg: TrainConfig):
    rng = np.random.default_rng(cfg.seed)
    w, o = cfg.window, cfg.npml + cfg.gap
    specs = []
    for _ in range(cfg.runs):
        width = float(rng.uniform(1.5, 4.0))
        si, sj = (float(v) for v in rng.uniform(cfg.margin, w - cfg.margin, 2))
        cyl = []
        for _ in range(int(rng.integers(0, 4))):
            r = float(rng.uniform(0.05, 0.15) * w)
            cx, cy = (float(v) for v in rng.uniform(10 + r, w - 10 - r, 2))
            cyl.append((cx + o, cy + o, r, float(rng.uniform(1.0, 8.0)), float(rng.uniform(0.0, 0.05))))
        specs.append(dict(pulse=(si + o, sj + o, width), cylinders=cyl))
    return specs


def pulse(n: int, m: int, ci: float, cj: float, width: float, device) -> torch.Tensor:
    """n x m FP32 field exp(-((i-ci)^2 + (j-cj)^2) / (2 width^2))."""
    i = np.arange(n, dtype=np.float64)[:, None] - ci
    j = np.arange(m, dtype=np.float64)[None, :] - cj
    f = np.exp(-(i * i + j * j) / (2.0 * width * width)).astype(np.float32)
    return torch.from_numpy(f).to(device)


def generate(cfg: TrainConfig, device=None, log=None) -> torch.Tensor:
    """Boundary samples of the training runs: G[run, 3 + t, position, 30] is
    the window's sample after step t + 1 (the first three slices are the
    fields at rest)."""
    dev = device or require_gpu()
    w, o = cfg.window, cfg.npml + cfg.gap
    N = w + 2 * o
    npw = npos(w, w)
    G = torch.zeros((cfg.runs, cfg.steps + 3, npw, NSAMP), dtype=torch.float32, device=dev)
    dt = fd2d.time_step()
    for r, sp in enumerate(training_specs(cfg)):
        pml = fd2d.new_pmlayer(N, N, dev)
        fd2d.pmlparam(N, N, cfg.npml, pml)
        md = fd2d.dielectric(N, N, cfg.npml, 0, dt, 1.0, 0.0, dev,
                             cylinders=sp["cylinders"] or np.zeros((0, 5)))
        s = fd2d.TMSolver(N, N, pml, md, [], [], nsteps_max=cfg.steps, device=dev)
        s.dz.copy_(pulse(N, N, *sp["pulse"], dev))
        for t in range(cfg.steps):
            s.run(1)
            sample(s.ez, s.hx, s.hy, out=G[r, 3 + t], window=(o, o + w, o, o + w))
        if log and (r + 1) % 64 == 0:
            torch.cuda.synchronize()
            log(f"generated {r + 1}/{cfg.runs} runs")
    torch.cuda.synchronize()
    return G


def training_rows(G: torch.Tensor, net: Net, s_floor: float) -> torch.Tensor:
    """Offsets (into G) of the rows with a target (slices 3 .. T+1) whose
    features exceed s_floor."""
    R, T3, npw, _ = G.shape
    ts = npw * NSAMP
    keep = []
    for r in range(R):
        t = torch.arange(3, T3 - 1, device=G.device, dtype=torch.int64)
        p = torch.arange(npw, device=G.device, dtype=torch.int64)
        offs = ((r * T3 + t)[:, None] * npw + p[None, :]).reshape(-1) * NSAMP
        _, s = net.forward(G, offs, ts)
        keep.append(offs[s > s_floor])
    return torch.cat(keep)


def validation_reflection(net: Net, n: int, device=None) -> float:
    """Closed-loop reflection ratio of the learned boundary on a validation
    case that differs from the evaluation defaults of reflection(): pulse
    (width 2.5) at (0.3 n, 0.6 n), cylinder (radius n/8, eps_r 3, sigma 0.01)
    at (0.65 n, 0.35 n)."""
    r = reflection(net, n, start=("pulse", 0.3 * n, 0.6 * n, 2.5),
                   cylinders=[(0.65 * n, 0.35 * n, n / 8, 3.0, 0.01)], methods=("learned",),
                   device=device)
    return r["ratio"]["learned"]


def train(cfg: TrainConfig, G: torch.Tensor, net: Net | None = None, log=None) -> tuple:
    """Three stages, all on the GPU:
    1. the linear path by weighted ridge least squares (lbc_gram /
       lbc_fit_linear on a random subset of cfg.fit_rows rows, row weights
       s^wpow), the weight exponent and ridge chosen among cfg.wpow x cfg.rho
       by the closed-loop reflection of a validation case
       (validation_reflection at cfg.select_n) -- the one-step loss does not
       tell a stable boundary from an unstable one;
    2. the hidden layers by Adam (lbc_train_grad / lbc_adam) on random
       batches, cosine learning rate, the linear path fixed: they fit what
       it leaves;
    3. the hidden layers' output scale chosen among cfg.scales by the same
       closed-loop validation (0 keeps the linear path alone).
    Returns (net, history: dicts of the stages' figures)."""
    dev = G.device
    if net is None:
        p = init_params(cfg.seed)
        p[P_W3:P_L] = 0.0
        p[P_L:] = 0.0
        net = Net(p, dev)
    ts = G.shape[2] * NSAMP
    rows = training_rows(G, net, cfg.s_floor)
    gen = torch.Generator(device=dev)
    gen.manual_seed(cfg.seed + 1)
    perm = torch.randperm(rows.numel(), device=dev, generator=gen)
    nval = min(65536, rows.numel() // 10)
    val, trn = rows[perm[:nval]], rows[perm[nval:]]
    hist = []
    v0, _ = net.grad(G, val, ts)
    hist.append({"stage": "start", "validation": v0})
    if log:
        log(f"{rows.numel()} rows ({trn.numel()} train, {nval} validation); "
            f"validation loss at start (the two-step boundary): {v0:.3e}")
    # 1. linear path
    fit = trn[torch.randperm(trn.numel(), device=dev, generator=gen)[:cfg.fit_rows]]
    best = None
    for wp in cfg.wpow:
        part = net.gram(G, fit, ts, wpow=wp)
        for rho in cfg.rho:
            try:
                net.fit_linear(part, rho)
            except _lib.FdtdError as e:
                hist.append({"stage": "linear", "wpow": wp, "rho": rho, "error": str(e)})
                continue
            vl, _ = net.grad(G, val, ts)
            rr = validation_reflection(net, cfg.select_n, dev)
            hist.append({"stage": "linear", "wpow": wp, "rho": rho, "validation": vl,
                         "reflection": rr})
            if log:
                log(f"linear path wpow {wp} rho {rho:.0e}: validation {vl:.3e} "
                    f"closed-loop reflection {rr:.2e}")
            if math.isfinite(rr) and (best is None or rr < best[0]):
                best = (rr, wp, rho, net.numpy()[P_L:].copy())
        del part
    if best is None:
        raise RuntimeError("no linear fit gave a finite closed-loop reflection")
    _, wp, rho, lbest = best
    p = net.numpy()
    p[P_L:] = lbest
    net.set_params(p)
    # 2. hidden layers
    nb = max(1, min(512, (cfg.batch + 255) // 256))
    gpart = torch.empty((nb, net.nparam), dtype=torch.float32, device=dev)
    ploss = torch.empty(nb, dtype=torch.float32, device=dev)
    p = net.numpy()
    p[P_W3:P_L] = init_params(cfg.seed)[P_W3:P_L]
    net.set_params(p)
    for k in range(cfg.train_steps):
        lr = cfg.lr_min + 0.5 * (cfg.lr - cfg.lr_min) * (1 + math.cos(math.pi * k / cfg.train_steps))
        idx = torch.randint(0, trn.numel(), (cfg.batch,), device=dev, generator=gen)
        pl = net.train_step(G, trn[idx], ts, lr, nblocks=nb, part=gpart, ploss=ploss,
                            noise=cfg.noise, only=(0, P_L))
        if (k + 1) % 1000 == 0 or k == cfg.train_steps - 1:
            tl = float(pl.double().sum()) / cfg.batch
            vl, _ = net.grad(G, val, ts)
            hist.append({"stage": "hidden", "step": k + 1, "train": tl, "validation": vl})
            if log:
                log(f"hidden step {k + 1}: train {tl:.3e} validation {vl:.3e} lr {lr:.2e}")
    # 3. hidden-layer scale
    trained = net.numpy()
    best = None
    for lam in cfg.scales:
        cand = scaled(net, lam)
        vl, _ = cand.grad(G, val, ts)
        rr = validation_reflection(cand, cfg.select_n, dev)
        hist.append({"stage": "scale", "scale": lam, "validation": vl, "reflection": rr})
        if log:
            log(f"hidden scale {lam}: validation {vl:.3e} closed-loop reflection {rr:.2e}")
        if math.isfinite(rr) and (best is None or rr < best[0]):
            best = (rr, lam)
    p = trained.copy()
    p[P_W3:P_L] *= best[1] if best else 0.0
    net.set_params(p)
    hist.append({"stage": "chosen", "wpow": wp, "rho": rho, "scale": best[1] if best else 0.0})
    return net, hist


def scaled(net: Net, lam: float) -> Net:
    """A copy whose hidden-layer output (W3, b3) is scaled by lam; lam = 0
    leaves the linear path alone."""
    p = net.numpy().copy()
    p[P_W3:P_L] *= lam
    return Net(p, net.device)


# --------------------------------------------------------------- evaluation

def _solver(method: str, n: int, net, cylinders, start, steps, dev, **kw):
    """(solver, window offset) of a boundary method around an n x n window.
    start: ("pulse", ci, cj, width) -- initial dz pulse -- or ("source", i, j,
    t0, spread) -- hard Gaussian source (fd2d/pycuda/fd2d_3_1.py:69);
    coordinates relative to the window."""
    from .fd1d import waveform
    pad = {"learned": 0, "abc": 0, "pec": 0, "pml8": 8, "pml32": 32, "ref": 64 + 8}[method]
    npml = {"pml8": 8, "pml32": 32, "ref": 64}.get(method, 0)
    N = n + 2 * pad
    dt = fd2d.time_step()
    cyl = [(cx + pad, cy + pad, r, e, sg) for cx, cy, r, e, sg in cylinders]
    md = fd2d.dielectric(N, N, npml, 0, dt, 1.0, 0.0, dev, cylinders=cyl or np.zeros((0, 5)))
    src = []
    if start[0] == "source":
        _, si, sj, t0, spread = start
        wave = waveform("gaussian", steps, 0, np.float32, center=t0, spread=spread)
        src = [(si + pad, sj + pad, wave)]
    if method in ("learned", "abc"):
        use = net if method == "learned" else Net(np.zeros(_lib.lib().lbc_param_count(), np.float32), dev)
        s = LearnedTMSolver(N, N, use, md, src, [], nsteps_max=steps, device=dev, **kw)
        f = s.tm
    else:
        pml = None
        if npml:
            pml = fd2d.new_pmlayer(N, N, dev)
            fd2d.pmlparam(N, N, npml, pml)
        s = fd2d.TMSolver(N, N, pml, md, src, [], nsteps_max=steps, device=dev, **kw)
        f = s
    if start[0] == "pulse":
        _, ci, cj, width = start
        f.dz.copy_(pulse(N, N, ci + pad, cj + pad, width, dev))
    return s, pad


def _fields(s):
    return (s.tm if isinstance(s, LearnedTMSolver) else s)


def reflection(net: Net | None, n: int = 2048, steps: int | None = None, chunk: int = 100,
               methods=("learned", "abc", "pml32", "pml8", "pec"), start=None, cylinders=None,
               device=None, log=None) -> dict:
    """Reflected energy of each boundary method on an n x n window.

    Methods: learned (the network), abc (zero network: the two-step boundary
    it corrects), pml32 / pml8 (Sullivan PML of that thickness outside the
    window), pec (the bare grid).  Reference: the same window inside an
    npml=64 grid with an 8-cell gap.  Every `chunk` steps the FP64 energy of
    the difference between a method's window and the reference window is
    taken (fd2d_energy_diff) over the window less its outer three cells (the
    cells the learned boundary writes; the grid's last row and column also
    hold no H there); the reflection ratio is its maximum over the run
    divided by the peak energy of the reference on the same cells.  Defaults: a point
    pulse (width 3 cells) at (n/2 + n/16, n/2 - n/32), one cylinder (radius
    n/10, eps_r 4, sigma 0.02) centred at (n/2 - n/8, n/2 + n/8), and steps
    long enough for the wave to leave through the corners."""
    dev = device or require_gpu()
    c = n // 2
    source = start or ("pulse", c + n / 16, c - n / 32, 3.0)
    cylinders = cylinders if cylinders is not None else [(c - n / 8, c + n / 8, n / 10, 4.0, 0.02)]
    steps = steps or int(1.6 * n) + 400
    steps -= steps % chunk
    ref, rp = _solver("ref", n, net, cylinders, source, steps, dev)
    sol = {m: _solver(m, n, net, cylinders, source, steps, dev) for m in methods}
    work = torch.empty(_lib.ENERGY_WORK, dtype=torch.float64, device=dev)
    err = {m: [] for m in methods}
    eref = []
    F = 3                  # measured window: the n x n window less its outer 3 cells
    wn = n - 2 * F
    for t in range(0, steps, chunk):
        ref.run(chunk)
        rf = _fields(ref)
        eref.append(fd2d.energy(rf.ez, rf.hx, rf.hy, (rp + F, rp + F + wn, rp + F, rp + F + wn),
                                work))
        for m, (s, p) in sol.items():
            s.run(chunk)
            f = _fields(s)
            out = C.c_double()
            call("fd2d_energy_diff", wn, wn, f.nx, f.ny, p + F, p + F, ptr(f.ez), ptr(f.hx),
                 ptr(f.hy), rf.nx, rf.ny, rp + F, rp + F, ptr(rf.ez), ptr(rf.hx), ptr(rf.hy),
                 ptr(work, torch.float64), C.byref(out), stream())
            err[m].append(out.value)
    peak = max(eref)
    res = {"n": n, "steps": steps, "peak_energy": peak, "final_ref_energy": eref[-1],
           "ratio": {m: max(e) / peak for m, e in err.items()},
           "final_ratio": {m: e[-1] / peak for m, e in err.items()},
           "curve_t": list(range(chunk, steps + 1, chunk)),
           "curve": {m: [x / peak for x in e] for m, e in err.items()},
           "ref_curve": [x / peak for x in eref]}
    if log:
        log("reflection " + " ".join(f"{m}={r:.2e}" for m, r in res["ratio"].items()))
    return res


def speed(net: Net, n: int = 2048, steps: int = 2000, methods=("learned", "pml32"), device=None,
          **kw) -> dict:
    """Wall time of `steps` steps for each method around an n x n window:
    raw cell-updates/s of its grid and window (physical) cell-updates/s."""
    dev = device or require_gpu()
    c = n // 2
    src = ("pulse", c, c, 3.0)
    cyl = [(c - n / 8, c + n / 8, n / 10, 4.0, 0.02)]
    out = {}
    for m in methods:
        s, p = _solver(m, n, net, cyl, src, steps + 16, dev, **kw)
        s.run(16)
        torch.cuda.synchronize()
        import time
        t0 = time.perf_counter()
        s.run(steps)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        N = n + 2 * p
        out[m] = {"grid": N, "seconds": dt, "mcups_grid": N * N * steps / dt / 1e6,
                  "mcups_window": n * n * steps / dt / 1e6}
        del s
        torch.cuda.empty_cache()
    return out


def save(net: Net, path: str, meta: dict | None = None) -> None:
    """Parameters (FP32) and metadata as .npz (no pickled objects)."""
    import json
    np.savez(path, params=net.numpy(), meta=np.array(json.dumps(meta or {})))


# the network trained by `python -m simulation_amd.learned` (shipped with the
# package: parameters and training record, no pickled objects)
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "lbc_weights.npz")


def load(path: str = WEIGHTS, device=None) -> Net:
    d = np.load(path, allow_pickle=False)
    return Net(d["params"], device)


def main(argv=None) -> int:
    """python -m simulation_amd.learned [--runs R] [--window W] [--steps T]
    [--train-steps K] [--eval N] [--out weights.npz]: generate, train,
    evaluate, save."""
    import argparse
    import json
    import time
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=256)
    ap.add_argument("--window", type=int, default=160)
    ap.add_argument("--steps", type=int, default=520)
    ap.add_argument("--train-steps", type=int, default=8000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--noise", type=float, default=0.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--eval", type=int, nargs="*", default=[512])
    ap.add_argument("--long", type=int, nargs=2, default=[0, 0], metavar=("N", "STEPS"),
                    help="long-run check: an N x N window for STEPS steps")
    ap.add_argument("--out", default="")
    ap.add_argument("--report", default="")
    a = ap.parse_args(argv)
    t0 = time.perf_counter()

    def log(m):
        print(f"[{time.perf_counter() - t0:7.1f}s] {m}", flush=True)
    cfg = TrainConfig(runs=a.runs, window=a.window, steps=a.steps, train_steps=a.train_steps,
                      batch=a.batch, lr=a.lr, seed=a.seed, noise=a.noise)
    dev = require_gpu()
    G = generate(cfg, dev, log)
    log(f"samples {tuple(G.shape)} ({G.numel() * 4 / 2**30:.1f} GiB)")
    net, hist = train(cfg, G, None, log)
    del G
    torch.cuda.empty_cache()
    rep = {"config": cfg.__dict__, "train": hist, "reflection": {}}
    for n in a.eval:
        r = reflection(net, n, device=dev, log=log)
        rep["reflection"][str(n)] = {k: r[k] for k in ("n", "steps", "ratio", "final_ratio")}
    if a.long[0]:
        r = reflection(net, a.long[0], steps=a.long[1], chunk=500,
                       methods=("learned", "abc", "pml32"), device=dev, log=log)
        rep["long"] = {k: r[k] for k in ("n", "steps", "ratio", "final_ratio", "curve_t", "curve")}
        log("long run final " + " ".join(f"{m}={v:.2e}" for m, v in r["final_ratio"].items()))
    if a.out:
        save(net, a.out, {"config": cfg.__dict__, "train": hist})
        log(f"saved {a.out}")
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
