"""2-D TM FDTD on MI355X -- host mirror of the PyCUDA 2-D programs.

Names, argument order and data layout follow fd2d/pycuda/fd2d_3_x.py:
``pmlayer`` / ``medium`` namedtuples of device arrays, ``pmlparam`` and
``dielectric`` for the coefficients, ``dfield``/``efield``/``hfield`` for one
step, and row-major ``nx*ny`` fields with ``n = i*ny + j``.  Every function
calls a HIP kernel of libfdtd_hip.so through ctypes; PyTorch only owns the
device buffers.

``TMSolver`` is the fused path: it advances the state by many steps per call
through ``fd2d_tm_run`` (temporally blocked kernels, several leapfrog steps
per launch), with hard point sources, probe traces and snapshots, and gives
the same values as calling dfield -> source -> efield -> hfield per step.
"""
from __future__ import annotations

import ctypes as C
from collections import namedtuple
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._dev import fptr, ptr, require_gpu, stream
from ._lib import call

# state tensors a TM descriptor points to
_DESC_TENSORS = ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy", "dz2", "iz2", "hx2", "hy2", "ihx2", "ihy2")

pmlayer = namedtuple("pmlayer", ("fx1", "fx2", "fx3", "fy1", "fy2", "fy3",
                                 "gx2", "gx3", "gy2", "gy3"))
medium = namedtuple("medium", ("naz", "nbz"))

__all__ = ["pmlayer", "medium", "new_pmlayer", "pmlparam", "dielectric", "dfield", "efield",
           "hfield", "energy", "TMSolver", "TMConfig", "simulate", "c2_config", "time_step",
           "ftrans", "ezinct", "dfield_inc", "inctdz", "hxinct", "incthx", "incthy", "fourier",
           "new_ftrans", "amplitude_phase", "PlaneWaveSolver"]


def time_step(ds: float = 0.01) -> float:
    """dt = np.float32(ds/6e8) with ds = np.float32(0.01) (fd2d/pycuda/fd2d_3_4.py:313-314)."""
    return float(np.float32(np.float32(ds) / np.float32(6e8)))


def _pml_struct(pml: pmlayer | None):
    if pml is None:
        return None
    return C.byref(_lib.FdtdPml(*[fptr(getattr(pml, k)) for k in pmlayer._fields]))


def new_pmlayer(nx: int, ny: int, device=None) -> pmlayer:
    """The default fills of fd2d/pycuda/fd2d_3_2.py:293-304 (f*1 = 0, the rest 1)."""
    dev = device or require_gpu()
    mk = {"x": nx, "y": ny}
    return pmlayer(**{k: (torch.zeros if k[2] == "1" else torch.ones)(
        mk[k[1]], dtype=torch.float32, device=dev) for k in pmlayer._fields})


def pmlparam(nx: int, ny: int, npml: int, pml: pmlayer) -> None:
    """Sullivan PML profiles (fd2d/pycuda/fd2d_3_2.py:264-273) on the device in
    FP64, stored FP32; also writes the default values outside the layer."""
    call("fd2d_pmlparam", int(nx), int(ny), int(npml), _pml_struct(pml), stream())


def dielectric(nx: int, ny: int, npml: int, rgrid: float, dt: float, epsr: float,
               sigma: float, device=None, cylinders=None, int_sub: bool = False) -> medium:
    """naz, nbz of a lossy cylinder (fd2d/pycuda/fd2d_3_4.py:233-254), on the
    device in FP64, stored FP32.  Default: one cylinder of radius rgrid at
    (nx/2-1, ny/2-1); ``cylinders`` = [(cx, cy, radius, epsr, sigma), ...]
    overrides it."""
    dev = device or require_gpu()
    if cylinders is None:
        cylinders = [(nx / 2 - 1, ny / 2 - 1, rgrid, epsr, sigma)]
    cyl = np.ascontiguousarray(np.asarray(cylinders, dtype=np.float64).reshape(-1, 5))
    if len(cyl) > _lib.MAX_CYLINDERS:
        raise ValueError(f"at most {_lib.MAX_CYLINDERS} cylinders")
    naz = torch.empty((nx, ny), dtype=torch.float32, device=dev)
    nbz = torch.empty((nx, ny), dtype=torch.float32, device=dev)
    call("fd2d_dielectric", int(nx), int(ny), int(npml), float(dt), len(cyl),
         cyl.ctypes.data_as(_lib._DP), 1 if int_sub else 0, ptr(naz), ptr(nbz), stream())
    return medium(naz, nbz)


def dfield(t: int, nx: int, ny: int, pml: pmlayer | None, dz, hx, hy, sources=()) -> None:
    """D update + hard sources [(i, j, value)] (fd2d/pycuda/fd2d_3_2.py:219-233;
    pml None = free space, fd2d/pycuda/fd2d_3_1.py:58-70)."""
    n = len(sources)
    si = (C.c_int * max(n, 1))(*[int(s[0]) for s in sources])
    sj = (C.c_int * max(n, 1))(*[int(s[1]) for s in sources])
    sv = (C.c_float * max(n, 1))(*[float(s[2]) for s in sources])
    call("fd2d_dfield", int(t), int(nx), int(ny), _pml_struct(pml), ptr(dz), ptr(hx), ptr(hy),
         n, si, sj, sv, stream())


def efield(nx: int, ny: int, md: medium, dz, iz, ez) -> None:
    """E from D (fd2d/pycuda/fd2d_3_4.py:175-185; md.nbz None = fd2d_3_1.py:73-82)."""
    m = _lib.FdtdMedium(fptr(md.naz), fptr(md.nbz))
    call("fd2d_efield", int(nx), int(ny), C.byref(m), ptr(dz), ptr(iz), ptr(ez), stream())


def hfield(nx: int, ny: int, pml: pmlayer | None, ez, ihx, ihy, hx, hy) -> None:
    """H update (fd2d/pycuda/fd2d_3_2.py:248-260; pml None = fd2d_3_1.py:85-95)."""
    call("fd2d_hfield", int(nx), int(ny), _pml_struct(pml), ptr(ez), ptr(ihx), ptr(ihy),
         ptr(hx), ptr(hy), stream())


def energy(ez, hx, hy, window=None, work: torch.Tensor | None = None) -> float:
    """FP64 sum of ez^2 + hx^2 + hy^2 over window (i0, i1, j0, j1) (default all)."""
    nx, ny = ez.shape
    i0, i1, j0, j1 = window if window is not None else (0, nx, 0, ny)
    if work is None:
        work = torch.empty(_lib.ENERGY_WORK, dtype=torch.float64, device=ez.device)
    out = C.c_double(0.0)
    call("fd2d_energy", nx, ny, int(i0), int(i1), int(j0), int(j1), ptr(ez), ptr(hx), ptr(hy),
         ptr(work, torch.float64), C.byref(out), stream())
    return out.value


STATE = ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy")


# ------------------------------------------------ plane wave (TF/SF) and DFT

ftrans = namedtuple("ftrans", ("r_pt", "i_pt", "r_in", "i_in"))


def ezinct(ny: int, ezi, hxi, bc) -> None:
    """Incident Ez and its two-step boundary (fd2d/pycuda/fd2d_3_4.py:140-148)."""
    call("fd2d_ezinct", int(ny), ptr(ezi), ptr(hxi), ptr(bc), stream())


def dfield_inc(t: int, nx: int, ny: int, pml: pmlayer | None, ezi, dz, hx, hy,
               inc_pos: int = 3, inc_val: float | None = None) -> None:
    """D update plus the incident source ezi[inc_pos] = inc_val, default
    gaussian(t, 20, 8) at 3 (fd2d/pycuda/fd2d_3_4.py:151-161)."""
    from .fd1d import gaussian
    v = gaussian(t, 20, 8.0) if inc_val is None else inc_val
    call("fd2d_dfield_inc", int(t), int(nx), int(ny), _pml_struct(pml), ptr(ezi), ptr(dz),
         ptr(hx), ptr(hy), int(inc_pos), float(v), stream())


def inctdz(nx: int, ny: int, npml: int, hxi, dz) -> None:
    """TF/SF correction of Dz (fd2d/pycuda/fd2d_3_4.py:164-170)."""
    call("fd2d_inctdz", int(nx), int(ny), int(npml), ptr(hxi), ptr(dz), stream())


def hxinct(ny: int, ezi, hxi) -> None:
    """Incident Hx (fd2d/pycuda/fd2d_3_4.py:186-191)."""
    call("fd2d_hxinct", int(ny), ptr(ezi), ptr(hxi), stream())


def incthx(nx: int, ny: int, npml: int, ezi, hx) -> None:
    """TF/SF correction of Hx (fd2d/pycuda/fd2d_3_4.py:211-217)."""
    call("fd2d_incthx", int(nx), int(ny), int(npml), ptr(ezi), ptr(hx), stream())


def incthy(nx: int, ny: int, npml: int, ezi, hy) -> None:
    """TF/SF correction of Hy (fd2d/pycuda/fd2d_3_4.py:220-226)."""
    call("fd2d_incthy", int(nx), int(ny), int(npml), ptr(ezi), ptr(hy), stream())


def _ft_struct(ft: ftrans) -> _lib.FdtdFtrans:
    return _lib.FdtdFtrans(fptr(ft.r_pt), fptr(ft.i_pt), fptr(ft.r_in), fptr(ft.i_in))


def fourier(t: int, nf: int, nx: int, ny: int, dt: float, freq, ezi, ez, ft: ftrans) -> None:
    """Running DFT of ez and of the incident probe ezi[6]
    (fd2d/pycuda/fd2d_3_4.py:121-136); freq: host frequencies."""
    fr = (C.c_float * nf)(*[float(f) for f in freq])
    call("fd2d_fourier", int(t), int(nf), int(nx), int(ny), float(dt), fr, ptr(ezi), ptr(ez),
         C.byref(_ft_struct(ft)), stream())


def new_ftrans(nf: int, nx: int, ny: int, device=None) -> ftrans:
    dev = device or require_gpu()
    z = lambda *s: torch.zeros(s, dtype=torch.float32, device=dev)  # noqa: E731
    return ftrans(z(nf, nx, ny), z(nf, nx, ny), z(nf), z(nf))


def amplitude_phase(ft: ftrans, nx: int, ny: int, npml: int):
    """Amplitude and phase at each frequency along row nx/2-1, as the
    reference's post-processing (fd2d/pycuda/fd2d_3_4.py:378-385): host
    NumPy on the FP32 transforms.  Returns (amplt, phase), (nf, ny) arrays
    set on columns npml-1 .. ny-npml."""
    r_pt, i_pt = ft.r_pt.cpu().numpy(), ft.i_pt.cpu().numpy()
    r_in, i_in = ft.r_in.cpu().numpy(), ft.i_in.cpu().numpy()
    nf = r_in.shape[0]
    amplt = np.zeros((nf, ny), np.float32)
    phase = np.zeros((nf, ny), np.float32)
    j = slice(npml - 1, ny - npml + 1)
    for n in range(nf):
        rp, ip = r_pt[n, nx // 2 - 1, j], i_pt[n, nx // 2 - 1, j]
        amplt[n, j] = 1 / np.hypot(r_in[n], i_in[n]) * np.hypot(rp, ip)
        phase[n, j] = np.arctan2(ip, rp) - np.arctan2(i_in[n], r_in[n])
    return amplt, phase


class PlaneWaveSolver:
    """A plane wave striking the grid through the TF/SF box, with the running
    DFT (fd2d/pycuda/fd2d_3_4.py:261-375): fd2d_tfsf_run over a TMSolver's
    state set.  inc_wave: host values of the incident source ezi[inc_pos]
    for steps 1.. (default gaussian(t, 20, 8) at 3, as the reference); freq:
    DFT frequencies (none: no DFT)."""

    def __init__(self, nx, ny, npml, pml, md, nsteps_max, freq=(), dt=None, inc_pos=3,
                 inc_wave=None, probes=(), device=None, fused=True, depth=0, rows=0, cols=0):
        from .fd1d import gaussian
        self.tm = TMSolver(nx, ny, pml, md, (), probes, nsteps_max, depth, rows, cols,
                           device=device)
        dev = self.tm.device
        self.npml = int(npml)
        self.ezi = torch.zeros(ny, dtype=torch.float32, device=dev)
        self.hxi = torch.zeros(ny, dtype=torch.float32, device=dev)
        self.bc = torch.zeros(4, dtype=torch.float32, device=dev)
        if inc_wave is None:
            inc_wave = [gaussian(t, 20, 8.0) for t in range(1, nsteps_max + 1)]
        self.inc_wave = np.ascontiguousarray(np.asarray(inc_wave, np.float32))
        self.inc_pos = int(inc_pos)
        self.freq = np.ascontiguousarray(np.asarray(freq, np.float32))
        self.dt = float(time_step() if dt is None else dt)
        self.ft = new_ftrans(len(self.freq), nx, ny, dev) if len(self.freq) else None
        # fused: the temporally blocked launches with the incident table in
        # work (the PML form, lossy or not, up to 3 DFT frequencies)
        nf = len(self.freq)
        fused = fused and nf <= 3 and pml is not None
        n = _lib.lib().fd2d_tfsf_work_elems(int(ny), int(npml)) if fused else 0
        self.work = torch.empty(n, dtype=torch.float32, device=dev) if n else None

    @property
    def t(self) -> int:
        return self.tm.t

    def desc(self) -> _lib.Fd2dTfsfDesc:
        b = _lib.Fd2dTfsfDesc()
        b.npml = self.npml
        b.ezi, b.hxi, b.bc = fptr(self.ezi), fptr(self.hxi), fptr(self.bc)
        b.inc_pos = self.inc_pos
        b.inc_wave = self.inc_wave.ctypes.data_as(C.POINTER(C.c_float))
        b.inc_len = len(self.inc_wave)
        b.nf = len(self.freq)
        b.freq = self.freq.ctypes.data_as(C.POINTER(C.c_float)) if b.nf else None
        b.dt = self.dt
        if self.ft is not None:
            b.ft = _ft_struct(self.ft)
        b.dft_probe = 6
        b.work = fptr(self.work)
        b.work_elems = 0 if self.work is None else self.work.numel()
        return b

    def run(self, nsteps: int) -> None:
        d, b = self.tm.desc(), self.desc()
        call("fd2d_tfsf_run", C.byref(d), C.byref(b), int(self.tm.t), int(nsteps), stream())
        self.tm.t += int(nsteps)

    def state(self) -> dict:
        return self.tm.state()


class TMSolver:
    """Fused 2-D TM time stepping on one grid (fd2d_tm_run).

    pml: pmlayer (or None for free space); md: medium (md.nbz None = lossless);
    sources: [(i, j, wave)] with wave a host array of the hard-source values of
    steps 1..wave_len; probes: [(i, j)] recorded every step up to probe_len.
    The state arrays follow fd2d/pycuda/fd2d_3_4.py:279-286; a partner set of
    the same shapes is allocated for the ping-pong launches.  form:
    _lib.TM_FORM_AUTO (the library's choice), _lib.TM_FORM_STRIP (single-wave
    strips) or _lib.TM_FORM_MULTIWAVE (interior blocks swept by four waves
    that exchange their edge columns); all give the same bits.
    """

    def __init__(self, nx, ny, pml, md, sources=(), probes=(), nsteps_max=None,
                 depth=0, rows=0, cols=0, device=None, form=_lib.TM_FORM_AUTO):
        dev = device or require_gpu()
        _lib.preload()
        self.nx, self.ny, self.device = int(nx), int(ny), dev
        self.pml, self.md = pml, md
        self.lossy = md.nbz is not None
        z = lambda: torch.zeros((nx, ny), dtype=torch.float32, device=dev)  # noqa: E731
        self.dz, self.ez, self.hx, self.hy = z(), z(), z(), z()
        self.iz = z() if self.lossy else None
        self.ihx = z() if pml is not None else None
        self.ihy = z() if pml is not None else None
        self.dz2, self.hx2, self.hy2 = z(), z(), z()
        self.iz2 = z() if self.lossy else None
        self.ihx2 = z() if pml is not None else None
        self.ihy2 = z() if pml is not None else None
        if len(sources) > _lib.MAX_SOURCES or len(probes) > _lib.MAX_PROBES:
            raise ValueError("too many sources or probes")
        self.sources = [(int(i), int(j)) for i, j, _ in sources]
        if sources:
            wl = min(len(w) for _, _, w in sources)
            if nsteps_max is not None:
                wl = min(wl, nsteps_max)
            waves = np.stack([np.asarray(w[:wl], dtype=np.float32) for _, _, w in sources])
            self.src_wave = torch.from_numpy(np.ascontiguousarray(waves)).to(dev)
            self.wave_len = wl
        else:
            self.src_wave, self.wave_len = None, 0
        self.probes = [(int(i), int(j)) for i, j in probes]
        self.probe_len = int(nsteps_max or self.wave_len or 0)
        self.probe_out = (torch.zeros((len(probes), self.probe_len), dtype=torch.float32, device=dev)
                          if probes else None)
        self.depth, self.rows, self.cols = depth, rows, cols
        self.form = int(form)
        # free-space tiles skip reading naz/nbz after a run call's first launch;
        # static_medium: the caller promises naz/nbz do not change between run
        # calls and iz is written only by the runs (forget_tiles otherwise),
        # so the flags of an earlier call serve from the first launch.
        # One workspace per tile geometry (depth, rows, cols) and per depth of
        # the call's first launch (the launch that records reads the rows of
        # that depth): calls of other lengths (a short warm-up, a tail) keep
        # their own flags and leave the others' alone.  A workspace's first
        # call always records, so flags of freed arrays whose memory a new
        # medium reuses never match.
        self.tile_flags = True
        self.static_medium = False
        self._tiles = {}
        self._dkey, self._dcache, self._tcache = None, None, {}
        self.band_rows, self.band_cols = self.default_band()
        self.t = 0
        self._work = torch.empty(_lib.ENERGY_WORK, dtype=torch.float64, device=dev)

    def default_band(self):
        """Rows and columns around the grid centre whose profile values are the
        defaults (f*1 = 0, the rest 1): the band promised to fd2d_tm_run.
        Call again after changing the pmlayer arrays."""
        if self.pml is None:
            return (0, 0), (0, 0)

        def run(mask):
            n = len(mask)
            c = n // 2
            if not mask[c]:
                return (0, 0)
            lo = c
            while lo > 0 and mask[lo - 1]:
                lo -= 1
            hi = c + 1
            while hi < n and mask[hi]:
                hi += 1
            return (lo, hi)
        pm = self.pml
        rows = ((pm.fx1 == 0) & (pm.fx2 == 1) & (pm.fx3 == 1) & (pm.gx2 == 1) & (pm.gx3 == 1))
        cols = ((pm.fy1 == 0) & (pm.fy2 == 1) & (pm.fy3 == 1) & (pm.gy2 == 1) & (pm.gy3 == 1))
        return run(rows.cpu().numpy()), run(cols.cpu().numpy())

    def desc(self) -> _lib.Fd2dTmDesc:
        d = _lib.Fd2dTmDesc()
        d.nx, d.ny = self.nx, self.ny
        for k in STATE:
            setattr(d, k, fptr(getattr(self, k)))
        for k in ("dz2", "iz2", "hx2", "hy2", "ihx2", "ihy2"):
            setattr(d, k, fptr(getattr(self, k)))
        d.md = _lib.FdtdMedium(fptr(self.md.naz), fptr(self.md.nbz))
        if self.pml is not None:
            d.pml = _lib.FdtdPml(*[fptr(getattr(self.pml, k)) for k in pmlayer._fields])
        d.nsrc = len(self.sources)
        for k, (i, j) in enumerate(self.sources):
            d.src_i[k], d.src_j[k] = i, j
        d.src_wave = fptr(self.src_wave)
        d.wave_len = self.wave_len
        d.nprobe = len(self.probes)
        for k, (i, j) in enumerate(self.probes):
            d.probe_i[k], d.probe_j[k] = i, j
        d.probe_out = fptr(self.probe_out)
        d.probe_len = self.probe_len
        d.depth, d.rows, d.cols = self.depth, self.rows, self.cols
        d.band_rows[0], d.band_rows[1] = self.band_rows
        d.band_cols[0], d.band_cols[1] = self.band_cols
        d.form = self.form
        return d

    def plan(self, nsteps: int) -> dict:
        d = self.desc()
        out = [C.c_int() for _ in range(4)]
        call("fd2d_tm_plan", C.byref(d), int(nsteps), *[C.byref(o) for o in out])
        return dict(launches=out[0].value, depth=out[1].value, rows=out[2].value,
                    cols=out[3].value)

    def _desc_key(self) -> tuple:
        """Everything desc() reads (tensors by object and address): a run
        call reuses the last descriptor while this is unchanged."""
        def t(x):
            return None if x is None else (id(x), x.data_ptr())
        pm = self.pml
        return (self.nx, self.ny,
                tuple(t(getattr(self, k)) for k in _DESC_TENSORS),
                t(self.md.naz), t(self.md.nbz), t(self.src_wave), t(self.probe_out),
                None if pm is None else tuple(t(getattr(pm, k)) for k in pmlayer._fields),
                tuple(self.sources), self.wave_len, tuple(self.probes), self.probe_len,
                self.depth, self.rows, self.cols, tuple(self.band_rows), tuple(self.band_cols),
                self.form)

    def forget_tiles(self) -> None:
        """Drop the free-space flags: the next run records afresh.  Call it
        after writing iz (or the medium) other than through runs; with
        static_medium the flags otherwise assume iz is +0 wherever they mark
        free space."""
        self._tiles = {}

    def run(self, nsteps: int) -> None:
        """Advance by nsteps (asynchronous on the current stream).  The
        library records the free-space tiles of the plan in a workspace the
        solver keeps (fd2d_tm_desc.tile_work; one int per tile)."""
        # the descriptor and the tile geometry of a call are rebuilt only
        # when something they are made of changed (tens of microseconds of
        # host time per call otherwise)
        dkey = self._desc_key()
        if dkey != self._dkey:
            self._dkey, self._dcache, self._tcache = dkey, self.desc(), {}
        d = self._dcache
        d.tile_work, d.tile_elems, d.tile_keep = None, 0, 0
        if self.tile_flags and nsteps > 0:
            geo = self._tcache.get(int(nsteps))
            if geo is None:
                n = _lib.lib().fd2d_tm_tile_elems(C.byref(d), int(nsteps))
                key = None
                if n:
                    out = [C.c_int() for _ in range(4)]
                    call("fd2d_tm_plan", C.byref(d), int(nsteps), *[C.byref(o) for o in out])
                    L = max(out[0].value, 1)
                    key = (out[1].value, out[2].value, out[3].value, -(-int(nsteps) // L))
                geo = self._tcache[int(nsteps)] = (n, key)
            n, key = geo
            if n:
                ws = self._tiles.get(key)
                first = ws is None or ws.numel() < n
                if first:
                    ws = self._tiles[key] = torch.zeros(n, dtype=torch.int32, device=self.device)
                d.tile_work, d.tile_elems = ws.data_ptr(), ws.numel()
                d.tile_keep = 1 if self.static_medium and not first else 0
        call("fd2d_tm_run", C.byref(d), int(self.t), int(nsteps), stream())
        self.t += int(nsteps)

    def step_unfused(self) -> None:
        """One step through the per-step kernels (reference launch order)."""
        t = self.t + 1
        src = [(i, j, 0.0) for i, j in self.sources]
        if self.sources:
            vals = self.src_wave[:, t - 1].cpu().numpy()
            src = [(i, j, float(v)) for (i, j), v in zip(self.sources, vals)]
        dfield(t, self.nx, self.ny, self.pml, self.dz, self.hx, self.hy, src)
        efield(self.nx, self.ny, self.md, self.dz, self.iz, self.ez)
        hfield(self.nx, self.ny, self.pml, self.ez, self.ihx, self.ihy, self.hx, self.hy)
        for p, (i, j) in enumerate(self.probes):
            self.probe_out[p, t - 1] = self.ez[i, j]
        self.t = t

    def energy(self, window=None) -> float:
        return energy(self.ez, self.hx, self.hy, window, self._work)

    def state(self) -> dict:
        torch.cuda.synchronize()
        return {k: (None if getattr(self, k) is None else getattr(self, k).cpu().numpy())
                for k in STATE}

    def probe_traces(self) -> np.ndarray | None:
        if self.probe_out is None:
            return None
        torch.cuda.synchronize()
        return self.probe_out[:, : self.t].cpu().numpy()


@dataclass
class TMConfig:
    """A 2-D TM run.  cylinders: [(cx, cy, radius, epsr, sigma)];
    sources: [(i, j, t0, spread)] hard Gaussian point sources."""
    nx: int
    ny: int
    ns: int
    npml: int = 8
    cylinders: list = field(default_factory=list)
    sources: list = field(default_factory=list)
    probes: list = field(default_factory=list)
    snap_every: int = 0
    lossy: bool = True
    depth: int = 0
    rows: int = 0
    cols: int = 0
    form: int = 0


def c2_config(n: int = 4096, ns: int = 2000) -> TMConfig:
    """The n x n TM configuration: npml=8 Sullivan PML, one cylinder of radius
    400 cells (eps_r=4, sigma=0.02) at (n/2-1, n/2-1) as
    fd2d/pycuda/fd2d_3_4.py:319, hard Gaussian (t0=20, spread 6) at (n/2, n/2)
    as fd2d/pycuda/fd2d_3_1.py:69, snapshots every 500 steps, four probes."""
    c = n // 2
    r = 400 if n >= 1024 else max(1, n // 10)
    return TMConfig(nx=n, ny=n, ns=ns, npml=8,
                    cylinders=[(n / 2 - 1, n / 2 - 1, r, 4.0, 0.02)],
                    sources=[(c, c, 20, 6.0)],
                    probes=[(c, c + n // 20), (c, c + n // 7), (c + n // 7, c), (12, c)],
                    snap_every=500 if ns >= 500 else 0)


def c5_config(n: int = 16384, ns: int = 5000, seed: int = 5) -> TMConfig:
    """Large single grid: n x n, npml=8, 16 seeded lossy cylinders (radius
    2-6% of the side, eps_r U[1,8], sigma U[0,0.05]) and 4 seeded hard
    Gaussian point sources (spread U[4,12], t0 = round(20 spread / 6)), one
    probe beside each source."""
    rng = np.random.default_rng(seed)
    lo, hi = n // 8, n - n // 8
    cyl = []
    for _ in range(16):
        r = float(rng.uniform(0.02, 0.06) * n)
        cx, cy = (float(v) for v in rng.uniform(lo, hi, 2))
        cyl.append((round(cx * 2) / 2, round(cy * 2) / 2, round(r * 2) / 2,
                    float(rng.uniform(1.0, 8.0)), float(rng.uniform(0.0, 0.05))))
    src, probes = [], []
    for _ in range(4):
        i, j = (int(v) for v in rng.integers(lo, hi, 2))
        spread = float(rng.uniform(4.0, 12.0))
        src.append((i, j, int(round(20 * spread / 6)), spread))
        probes.append((i + 10, j))
    return TMConfig(nx=n, ny=n, ns=ns, npml=8, cylinders=cyl, sources=src, probes=probes,
                    snap_every=0)


def build_solver(cfg: TMConfig, device=None) -> TMSolver:
    dev = device or require_gpu()
    pml = None
    if cfg.npml >= 0:
        pml = new_pmlayer(cfg.nx, cfg.ny, dev)
        pmlparam(cfg.nx, cfg.ny, cfg.npml, pml)
    dt = time_step()
    md = dielectric(cfg.nx, cfg.ny, max(cfg.npml, 0), 0, dt, 1.0, 0.0, dev,
                    cylinders=cfg.cylinders or np.zeros((0, 5)))
    if not cfg.lossy:
        md = medium(md.naz, None)
    from .fd1d import waveform
    srcs = [(i, j, waveform("gaussian", cfg.ns, 0, np.float32, center=t0, spread=sp))
            for i, j, t0, sp in cfg.sources]
    return TMSolver(cfg.nx, cfg.ny, pml, md, srcs, cfg.probes, nsteps_max=cfg.ns,
                    depth=cfg.depth, rows=cfg.rows, cols=cfg.cols, device=dev, form=cfg.form)


def simulate(cfg: TMConfig, device=None, snapshot_fields=("ez", "hy"), out_dir=None) -> dict:
    """Run cfg; returns final state, probe traces and the snapshots.  With
    out_dir, each snapshot is also written as <field>_t<step>.npy and the
    probe traces as probes.csv (simulation_amd.store)."""
    from . import store
    s = build_solver(cfg, device)
    s.static_medium = True          # the medium stays as built for the whole run
    snaps = {k: [] for k in snapshot_fields}
    snap_t = []
    seg = cfg.snap_every or cfg.ns
    while s.t < cfg.ns:
        s.run(min(seg, cfg.ns - s.t))
        if cfg.snap_every and s.t % cfg.snap_every == 0:
            for k in snapshot_fields:
                snaps[k].append(getattr(s, k).cpu().numpy())
                if out_dir:
                    store.write_snapshot(out_dir, k, s.t, snaps[k][-1])
            snap_t.append(s.t)
    out = dict(state=s.state(), probe=s.probe_traces(), snap_t=snap_t,
               snapshots={k: np.stack(v) for k, v in snaps.items() if v})
    if out_dir and out["probe"] is not None:
        import os
        store.write_probes_csv(os.path.join(out_dir, "probes.csv"), out["probe"],
                               [f"ez[{i};{j}]" for i, j in s.probes])
    return out
