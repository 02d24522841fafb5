"""CPU oracle for the FDTD hot path -- TEST INFRASTRUCTURE ONLY.

ctypes front end of ``oracle/liboracle.so`` (built from ``fdtd_oracle.c`` by
``oracle/Makefile``).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module, and only as the
checker or the timed CPU baseline; the product package ``simulation_amd``
never imports it.

Every function takes and returns NumPy arrays in the reference's layout:
1-D fields of length ``nx``; 2-D fields row-major ``(nx, ny)`` with
``n = i*ny + j`` (fd2d/pycuda/fd2d_3_1.py:63).  ``dtype`` selects the FP32
instance (the reference's ``test_*.c`` programs) or the FP64 instance (its
``fd*_*.c`` programs).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build() -> str:
    """Compile liboracle.so with oracle/Makefile; returns the library path."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _lib.orc_gaussian_f32.restype = C.c_float
        _lib.orc_gaussian_f32.argtypes = [C.c_int, C.c_int, C.c_float]
        _lib.orc_gaussian_f64.restype = C.c_double
        _lib.orc_gaussian_f64.argtypes = [C.c_int, C.c_int, C.c_double]
        _lib.orc_sinusoidal_f32.restype = C.c_float
        _lib.orc_sinusoidal_f32.argtypes = [C.c_int, C.c_float, C.c_float]
        _lib.orc_sinusoidal_f64.restype = C.c_double
        _lib.orc_sinusoidal_f64.argtypes = [C.c_int, C.c_double, C.c_double]
        _lib.orc_fd2d_energy_f32.restype = C.c_double
        _lib.orc_fd2d_energy_f64.restype = C.c_double
    return _lib


def _ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    if a.dtype == np.float32:
        return a.ctypes.data_as(_fp)
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    raise TypeError(a.dtype)


def _suf(dtype) -> str:
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


# ----------------------------------------------------------------- waveforms

def gaussian(t: int, t0: int, sigma: float, dtype=np.float32):
    """fd1d/clang/test_1_2.c:13-15 (FP32) / fd1d/clang/fd1d_1_2.c:146-148 (FP64)."""
    if np.dtype(dtype) == np.float32:
        return np.float32(lib().orc_gaussian_f32(int(t), int(t0), float(sigma)))
    return np.float64(lib().orc_gaussian_f64(int(t), int(t0), float(sigma)))


def sinusoidal(t: int, ds: float, freq: float, dtype=np.float32):
    """fd2d/clang/test_3_2.c:22-25 (FP32) / fd2d/clang/fd2d_3_2.c:22-25 (FP64)."""
    if np.dtype(dtype) == np.float32:
        return np.float32(lib().orc_sinusoidal_f32(int(t), float(ds), float(freq)))
    return np.float64(lib().orc_sinusoidal_f64(int(t), float(ds), float(freq)))


def waveform(kind: str, t0: int, nsteps: int, dtype=np.float32, **kw) -> np.ndarray:
    """Source values of steps t0+1 .. t0+nsteps."""
    out = np.empty(nsteps, dtype=dtype)
    for s in range(nsteps):
        t = t0 + 1 + s
        if kind == "gaussian":
            out[s] = gaussian(t, kw["center"], kw["spread"], dtype)
        elif kind == "sinusoidal":
            out[s] = sinusoidal(t, kw.get("ds", 0.01), kw["freq"], dtype)
        else:
            raise ValueError(kind)
    return out


# ----------------------------------------------------------------------- 1-D

def fd1d_dielectric(nx: int, dt: float, epsr: float, sigma: float):
    """ca, cb in FP64 (fd1d/clang/fd1d_1_5.c:44-55)."""
    ca = np.empty(nx, np.float64)
    cb = np.empty(nx, np.float64)
    lib().orc_fd1d_dielectric_f64(C.c_int(nx), C.c_double(dt), C.c_double(epsr),
                                  C.c_double(sigma), _ptr(ca), _ptr(cb))
    return ca, cb


def fd1d_exfield(t, nx, ca, cb, ex, hy, src_mode=0, src_pos=0, src_val=0.0):
    dt = ex.dtype
    real = C.c_float if dt == np.float32 else C.c_double
    getattr(lib(), "orc_fd1d_exfield_" + _suf(dt))(
        C.c_int(t), C.c_int(nx), _ptr(ca), _ptr(cb), _ptr(ex), _ptr(hy),
        C.c_int(src_mode), C.c_int(src_pos), real(src_val))


def fd1d_hyfield(nx, ex, hy, bc):
    getattr(lib(), "orc_fd1d_hyfield_" + _suf(ex.dtype))(
        C.c_int(nx), _ptr(ex), _ptr(hy), _ptr(bc))


def fd1d_run(nx, t0, nsteps, ex, hy, bc=None, ca=None, cb=None,
             src_mode=0, src_pos=0, src_wave=None, probes=()):
    """Advance (ex, hy, bc) in place; returns probe traces (nprobe, nsteps)."""
    dt = ex.dtype
    pk = np.ascontiguousarray(np.asarray(probes, dtype=np.int32).reshape(-1))
    pout = np.zeros((len(pk), nsteps), dtype=dt)
    getattr(lib(), "orc_fd1d_run_" + _suf(dt))(
        C.c_int(nx), C.c_int(t0), C.c_int(nsteps), _ptr(ca), _ptr(cb),
        _ptr(ex), _ptr(hy), _ptr(bc), C.c_int(src_mode), C.c_int(src_pos),
        _ptr(src_wave), C.c_int(len(pk)), _ptr(pk) if len(pk) else None,
        _ptr(pout) if len(pk) else None)
    return pout


# ----------------------------------------------------------------------- 2-D

PML_FIELDS = ("fx1", "fx2", "fx3", "fy1", "fy2", "fy3", "gx2", "gx3", "gy2", "gy3")


def fd2d_pmlparam(nx: int, ny: int, npml: int) -> dict:
    """Ten FP64 profiles, PyCUDA host arithmetic (fd2d/pycuda/fd2d_3_2.py:264-273)."""
    arrs = {k: np.empty(nx if k[1] == "x" else ny, np.float64) for k in PML_FIELDS}
    lib().orc_fd2d_pmlparam_f64(C.c_int(nx), C.c_int(ny), C.c_int(npml),
                                *[_ptr(arrs[k]) for k in PML_FIELDS])
    return arrs


def fd2d_dielectric(nx, ny, npml, dt, cylinders, int_sub=False):
    """naz, nbz (FP64) for cylinders [(cx, cy, radius, epsr, sigma), ...].

    Single cylinder at (nx/2-1, ny/2-1): the Numba dielectric of
    fd2d/pycuda/fd2d_3_4.py:233-254; ``int_sub`` gives the C programs'
    integer sub-sampling (fd2d/clang/fd2d_3_4.c:149-176)."""
    cyl = np.asarray(cylinders, dtype=np.float64).reshape(-1, 5)
    cols = [np.ascontiguousarray(cyl[:, k]) for k in range(5)]
    naz = np.empty((nx, ny), np.float64)
    nbz = np.empty((nx, ny), np.float64)
    lib().orc_fd2d_dielectric_f64(C.c_int(nx), C.c_int(ny), C.c_int(npml), C.c_double(dt),
                                  C.c_int(len(cyl)), *[_ptr(c) for c in cols],
                                  C.c_int(1 if int_sub else 0), _ptr(naz), _ptr(nbz))
    return naz, nbz


@dataclass
class TMState:
    """The 2-D TM state set of fd2d/pycuda/fd2d_3_4.py:279-286 (ezi/hxi excluded)."""
    nx: int
    ny: int
    dtype: type = np.float32
    dz: np.ndarray = field(init=False)
    iz: np.ndarray = field(init=False)
    ez: np.ndarray = field(init=False)
    hx: np.ndarray = field(init=False)
    hy: np.ndarray = field(init=False)
    ihx: np.ndarray = field(init=False)
    ihy: np.ndarray = field(init=False)

    def __post_init__(self):
        for k in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy"):
            setattr(self, k, np.zeros((self.nx, self.ny), dtype=self.dtype))

    def copy(self):
        s = TMState(self.nx, self.ny, self.dtype)
        for k in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy"):
            setattr(s, k, getattr(self, k).copy())
        return s


def fd2d_tm_run(state: TMState, t0: int, nsteps: int, pml: dict | None,
                naz: np.ndarray, nbz: np.ndarray | None,
                sources=(), src_wave=None, probes=()):
    """Advance ``state`` in place (serial order dfield, sources, efield, hfield).

    ``pml``: dict of the ten 1-D profiles in the state dtype, or None for the
    free-space form; ``nbz`` None selects ez = naz*dz.  ``sources``: [(i, j)],
    ``src_wave``: (nsrc, nsteps).  Returns probe traces (nprobe, nsteps)."""
    dt = state.dtype
    suf = _suf(dt)
    ptrt = _fp if np.dtype(dt) == np.float32 else _dp
    if pml is not None:
        keep = [np.ascontiguousarray(pml[k], dtype=dt) for k in PML_FIELDS]
        tab = (ptrt * 10)(*[_ptr(a) for a in keep])
    else:
        keep, tab = [], None
    naz = np.ascontiguousarray(naz, dtype=dt)
    nbz = None if nbz is None else np.ascontiguousarray(nbz, dtype=dt)
    src = np.asarray(sources, dtype=np.int32).reshape(-1, 2)
    si = np.ascontiguousarray(src[:, 0])
    sj = np.ascontiguousarray(src[:, 1])
    sw = None
    if len(src):
        sw = np.ascontiguousarray(np.asarray(src_wave, dtype=dt).reshape(len(src), nsteps))
    prb = np.asarray(probes, dtype=np.int32).reshape(-1, 2)
    pi_ = np.ascontiguousarray(prb[:, 0])
    pj_ = np.ascontiguousarray(prb[:, 1])
    pout = np.zeros((len(prb), nsteps), dtype=dt)
    getattr(lib(), "orc_fd2d_tm_run_" + suf)(
        C.c_int(state.nx), C.c_int(state.ny), C.c_int(t0), C.c_int(nsteps),
        tab, _ptr(naz), _ptr(nbz),
        _ptr(state.dz), _ptr(state.iz), _ptr(state.ez), _ptr(state.hx), _ptr(state.hy),
        _ptr(state.ihx), _ptr(state.ihy),
        C.c_int(len(src)), _ptr(si) if len(src) else None, _ptr(sj) if len(src) else None,
        _ptr(sw),
        C.c_int(len(prb)), _ptr(pi_) if len(prb) else None, _ptr(pj_) if len(prb) else None,
        _ptr(pout) if len(prb) else None)
    return pout


def fd2d_energy(ez, hx, hy) -> float:
    nx, ny = ez.shape
    return float(getattr(lib(), "orc_fd2d_energy_" + _suf(ez.dtype))(
        C.c_int(nx), C.c_int(ny), _ptr(ez), _ptr(hx), _ptr(hy)))
