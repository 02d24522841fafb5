"""Batched uneven 2-D TM grids on MI355X (libfdtd_hip.so fd2d_batch_*).

Hundreds of independent grids of different sizes, each with its own lossy
cylinders, PML profiles, hard Gaussian source and probe, advance together: a
persistent kernel walks a work list of (grid, strip, row block) units, each
launch fuses ``depth`` leapfrog steps for every grid, and a whole run can be
captured into one HIP graph and replayed as a single launch.  Every grid
evolves exactly as it would alone through ``fd2d.TMSolver`` (bit for bit).

No reference counterpart: the reference runs one grid per program
(fd2d/pycuda/fd2d_3_4.py:269-389); the per-grid physics is the same.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib, fd2d
from ._dev import fptr, require_gpu, stream
from ._lib import call
from .fd1d import waveform

INT_MAX = 2**31 - 1


@dataclass
class GridSpec:
    """One grid: size, npml, cylinders [(cx, cy, radius, epsr, sigma)], hard
    Gaussian source (i, j, t0, spread) and an optional ez probe (i, j)."""
    nx: int
    ny: int
    npml: int = 8
    cylinders: list = field(default_factory=list)
    source: tuple | None = None
    probe: tuple | None = None


def c3_specs(seed: int = 0, count: int = 256, lo: int = 128, hi: int = 1024, npml: int = 8):
    """The batched configuration: `count` square grids with sides uniform in
    [lo, hi], 1-5 lossy cylinders each (eps_r U[1,8], sigma U[0,0.05], radius
    U[4%, 20%] of the side, whole disc at least 2 cells clear of the PML),
    npml=8, and a
    hard Gaussian source of spread U[4,12] (t0 = 20*spread/6, the ratio of
    fd2d/pycuda/fd2d_3_1.py:69) at a random interior point, probed 10 cells
    away.  Deterministic in `seed`."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(count):
        n = int(rng.integers(lo, hi + 1))
        inner = n - 2 * npml
        cyl = []
        for _ in range(int(rng.integers(1, 6))):
            r = float(rng.uniform(0.04, 0.2) * n)
            cx = float(rng.uniform(npml + r + 2, n - npml - r - 2))
            cy = float(rng.uniform(npml + r + 2, n - npml - r - 2))
            cyl.append((cx, cy, r, float(rng.uniform(1.0, 8.0)), float(rng.uniform(0.0, 0.05))))
        spread = float(rng.uniform(4.0, 12.0))
        si = int(rng.integers(npml + 2, npml + inner - 12))
        sj = int(rng.integers(npml + 2, npml + inner - 12))
        specs.append(GridSpec(n, n, npml, cyl, (si, sj, int(round(20 * spread / 6)), spread),
                              (si + 10, sj)))
    return specs


class BatchSolver:
    """Allocates and advances a batch of grids (lossy PML form)."""

    def __init__(self, specs, nsteps_max: int, depth: int = 0, rows: int = 0, cols: int = 0,
                 waves: int = 0, device=None, unit_flags: bool = True):
        dev = device or require_gpu()
        _lib.preload()
        self.device = dev
        self.specs = list(specs)
        self.nsteps_max = int(nsteps_max)
        self.t = 0
        self._keep = []
        # free-space units: the first launch of each run records which
        # interior units read only naz = 1, nbz = 0, later launches skip both
        # arrays there (fd2d_batch_desc.unit_work, one int per unit)
        self.unit_flags = bool(unit_flags)
        self._unit_work = None
        n = len(self.specs)
        self.grids = (_lib.Fd2dGrid * n)()
        self.state = []
        dt = fd2d.time_step()
        for k, sp in enumerate(self.specs):
            z = lambda: torch.zeros((sp.nx, sp.ny), dtype=torch.float32, device=dev)  # noqa: E731
            st = {name: z() for name in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy",
                                         "dz2", "iz2", "hx2", "hy2", "ihx2", "ihy2")}
            pml = fd2d.new_pmlayer(sp.nx, sp.ny, dev)
            fd2d.pmlparam(sp.nx, sp.ny, sp.npml, pml)
            md = fd2d.dielectric(sp.nx, sp.ny, sp.npml, 0, dt, 1.0, 0.0, dev,
                                 cylinders=sp.cylinders or np.zeros((0, 5)))
            g = self.grids[k]
            g.nx, g.ny = sp.nx, sp.ny
            for name, t in st.items():
                setattr(g, name, fptr(t))
            g.md = _lib.FdtdMedium(fptr(md.naz), fptr(md.nbz))
            g.pml = _lib.FdtdPml(*[fptr(getattr(pml, f)) for f in fd2d.pmlayer._fields])
            # profile values are the defaults on [npml, n-1-npml) (fd2d_pmlparam)
            g.band_rows[0], g.band_rows[1] = sp.npml, sp.nx - 1 - sp.npml
            g.band_cols[0], g.band_cols[1] = sp.npml, sp.ny - 1 - sp.npml
            wave = probe = None
            if sp.source is not None:
                si, sj, t0, spread = sp.source
                wave = torch.from_numpy(waveform("gaussian", self.nsteps_max, 0, np.float32,
                                                 center=t0, spread=spread)).to(dev)
                g.src_i, g.src_j, g.src_wave, g.wave_len = si, sj, fptr(wave), self.nsteps_max
            else:
                g.src_i, g.src_j, g.wave_len = -1, 0, 0
            if sp.probe is not None:
                probe = torch.zeros(self.nsteps_max, dtype=torch.float32, device=dev)
                g.probe_i, g.probe_j = sp.probe
                g.probe_out, g.probe_len = fptr(probe), self.nsteps_max
            else:
                g.probe_i, g.probe_j, g.probe_len = -1, 0, 0
            st.update(pml=pml, md=md, wave=wave, probe=probe)
            self.state.append(st)
        self.grids_dev = torch.empty(n * C.sizeof(_lib.Fd2dGrid), dtype=torch.uint8, device=dev)
        self.wave_len = self.nsteps_max if any(s.source for s in self.specs) else INT_MAX
        self.probe_len = self.nsteps_max if any(s.probe for s in self.specs) else INT_MAX
        self._graphs = {}
        self.plan(depth, rows, cols, waves)

    def plan(self, depth: int = 0, rows: int = 0, cols: int = 0, waves: int = 0) -> None:
        """(Re)build the work list for a launch shape (0 = library default).
        Captured graphs of the previous shape are destroyed."""
        self.close()
        n = len(self.specs)
        nunits = C.c_int()
        call("fd2d_batch_units", self.grids, n, depth, rows, cols, C.byref(nunits))
        self.units = torch.empty(4 * nunits.value, dtype=torch.int32, device=self.device)
        self.desc = _lib.Fd2dBatchDesc()
        call("fd2d_batch_upload", self.grids, n, depth, rows, cols, waves,
             self.grids_dev.data_ptr(), self.units.data_ptr(), nunits.value,
             C.byref(self.desc), stream())
        self._unit_work = None
        if self.unit_flags:
            self._unit_work = torch.zeros(nunits.value, dtype=torch.int32, device=self.device)
            self.desc.unit_work = self._unit_work.data_ptr()

    def reset(self) -> None:
        """All fields and probe traces back to zero, step counter to 0."""
        for st in self.state:
            for name in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy",
                         "dz2", "iz2", "hx2", "hy2", "ihx2", "ihy2"):
                st[name].zero_()
            if st["probe"] is not None:
                st["probe"].zero_()
        self.t = 0

    @property
    def cells(self) -> int:
        return sum(s.nx * s.ny for s in self.specs)

    def run(self, nsteps: int) -> None:
        call("fd2d_batch_run", C.byref(self.desc), self.t, int(nsteps), self.wave_len,
             self.probe_len, stream())
        self.t += int(nsteps)

    def graph(self, nsteps: int) -> int:
        """Executable graph advancing the batch from the current step by nsteps."""
        key = (self.t, int(nsteps))
        if key not in self._graphs:
            h = C.c_void_p()
            call("fd2d_batch_graph", C.byref(self.desc), self.t, int(nsteps), self.wave_len,
                 self.probe_len, stream(), C.byref(h))
            self._graphs[key] = h.value
        return self._graphs[key]

    def run_graph(self, nsteps: int) -> None:
        """Advance by nsteps with one graph launch (captured on first use)."""
        h = self.graph(nsteps)
        call("fdtd_graph_launch", h, stream())
        self.t += int(nsteps)

    def grid_state(self, k: int) -> dict:
        torch.cuda.synchronize()
        return {name: self.state[k][name].cpu().numpy()
                for name in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy")}

    def probe_trace(self, k: int):
        p = self.state[k]["probe"]
        return None if p is None else p[: self.t].cpu().numpy()

    def close(self) -> None:
        for h in self._graphs.values():
            _lib.lib().fdtd_graph_destroy(h)
        self._graphs.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def single_solver(spec: GridSpec, nsteps_max: int, device=None, **kw) -> fd2d.TMSolver:
    """The same grid as a stand-alone TMSolver (for checking a batch)."""
    cfg = fd2d.TMConfig(nx=spec.nx, ny=spec.ny, ns=nsteps_max, npml=spec.npml,
                        cylinders=spec.cylinders,
                        sources=[spec.source] if spec.source else [],
                        probes=[spec.probe] if spec.probe else [], **kw)
    return fd2d.build_solver(cfg, device)
