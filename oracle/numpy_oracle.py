"""NumPy restatement of the 2-D TM step -- TEST INFRASTRUCTURE ONLY.

Whole-array slicing in the shape of the reference's NumPy program
fd2d/program/fd2d_3_3.py:68-98 (dfield, efield, hfield), with the lossy efield
of fd2d/pycuda/fd2d_3_4.py:175-185 and the hard source of
fd2d/pycuda/fd2d_3_1.py:69.  It is an independent second restatement used to
cross-check oracle/fdtd_oracle.c (same values in FP64 and FP32) and to run
tiny cases without a compiler.  Only tests/ import it.
"""
from __future__ import annotations

import numpy as np


def dfield(nx, ny, pml, dz, hx, hy):
    """fd2d/program/fd2d_3_3.py:68-72 (without its plane-wave source)."""
    h = dz.dtype.type(0.5)
    if pml is None:
        dz[1:nx, 1:ny] += h * (hy[1:nx, 1:ny] - hy[0:nx-1, 1:ny] - hx[1:nx, 1:ny] + hx[1:nx, 0:ny-1])
        return
    dz[1:nx, 1:ny] = (pml["gx3"][1:nx, None] * pml["gy3"][1:ny] * dz[1:nx, 1:ny]
                      + pml["gx2"][1:nx, None] * pml["gy2"][1:ny] * h
                      * (hy[1:nx, 1:ny] - hy[0:nx-1, 1:ny] - hx[1:nx, 1:ny] + hx[1:nx, 0:ny-1]))


def efield(naz, nbz, dz, iz, ez):
    """fd2d/pycuda/fd2d_3_4.py:181-182 (nbz None: fd2d/program/fd2d_3_3.py:82-84)."""
    if nbz is None:
        ez[:] = naz * dz
        return
    ez[:] = naz * (dz - iz)
    iz += nbz * ez


def hfield(nx, ny, pml, ez, ihx, ihy, hx, hy):
    """fd2d/program/fd2d_3_3.py:91-98: the curl difference is halved as a
    whole (0.5*curl) where the C/CUDA kernels halve each term; the two agree
    exactly (a power-of-two scaling commutes with rounding)."""
    h = ez.dtype.type(0.5)
    curl_em = ez[0:nx-1, 0:ny-1] - ez[0:nx-1, 1:ny]
    curl_en = ez[0:nx-1, 0:ny-1] - ez[1:nx, 0:ny-1]
    if pml is None:
        hx[0:nx-1, 0:ny-1] += h * curl_em
        hy[0:nx-1, 0:ny-1] -= h * curl_en
        return
    ihx[0:nx-1, 0:ny-1] += curl_em
    ihy[0:nx-1, 0:ny-1] += curl_en
    hx[0:nx-1, 0:ny-1] = (pml["fy3"][0:ny-1] * hx[0:nx-1, 0:ny-1]
                          + pml["fy2"][0:ny-1] * (h * curl_em + pml["fx1"][0:nx-1, None] * ihx[0:nx-1, 0:ny-1]))
    hy[0:nx-1, 0:ny-1] = (pml["fx3"][0:nx-1, None] * hy[0:nx-1, 0:ny-1]
                          - pml["fx2"][0:nx-1, None] * (h * curl_en + pml["fy1"][0:ny-1] * ihy[0:nx-1, 0:ny-1]))


def tm_run(state: dict, nsteps: int, pml, naz, nbz, sources=(), src_wave=None, probes=()):
    """Advance state {dz, iz, ez, hx, hy, ihx, ihy} (arrays (nx, ny)) in place."""
    nx, ny = state["dz"].shape
    out = np.zeros((len(probes), nsteps), state["dz"].dtype)
    for s in range(nsteps):
        dfield(nx, ny, pml, state["dz"], state["hx"], state["hy"])
        for k, (i, j) in enumerate(sources):
            state["dz"][i, j] = src_wave[k][s]
        efield(naz, nbz, state["dz"], state["iz"], state["ez"])
        hfield(nx, ny, pml, state["ez"], state["ihx"], state["ihy"], state["hx"], state["hy"])
        for p, (i, j) in enumerate(probes):
            out[p, s] = state["ez"][i, j]
    return out


# ------------------------------------------------ plane wave (TF/SF) and DFT
# fd2d/clang/test_3_4.c:38-147 (FP32) and fd2d/clang/fd2d_3_4.c:38-147
# (FP64): serial semantics of the reference's plane-wave program.

def ezinct(ezi, hxi, bc):
    """fd2d/clang/test_3_4.c:56-65: update, then the two-step boundary at
    both ends."""
    h = ezi.dtype.type(0.5)
    ny = len(ezi)
    ezi[1:ny] += h * (hxi[0:ny-1] - hxi[1:ny])
    ezi[0] = bc[0]; bc[0] = bc[1]; bc[1] = ezi[1]
    ezi[ny-1] = bc[3]; bc[3] = bc[2]; bc[2] = ezi[ny-2]


def hxinct(ezi, hxi):
    """fd2d/clang/test_3_4.c:105-111."""
    h = ezi.dtype.type(0.5)
    ny = len(ezi)
    hxi[0:ny-1] += h * (ezi[0:ny-1] - ezi[1:ny])


def inctdz(nx, ny, npml, hxi, dz):
    """fd2d/clang/test_3_4.c:82-89: column npml-1 gains, column ny-npml
    loses half the incident Hx beside the box (in this order)."""
    h = dz.dtype.type(0.5)
    r = slice(npml - 1, nx - npml + 1)
    dz[r, npml - 1] += h * hxi[npml - 2]
    dz[r, ny - npml] -= h * hxi[ny - npml]


def incthx(nx, ny, npml, ezi, hx):
    """fd2d/clang/test_3_4.c:129-136."""
    h = hx.dtype.type(0.5)
    r = slice(npml - 1, nx - npml + 1)
    hx[r, npml - 2] += h * ezi[npml - 1]
    hx[r, ny - npml] -= h * ezi[ny - npml]


def incthy(nx, ny, npml, ezi, hy):
    """fd2d/clang/test_3_4.c:139-146."""
    h = hy.dtype.type(0.5)
    c = slice(npml - 1, ny - npml + 1)
    hy[npml - 2, c] -= h * ezi[c]
    hy[nx - npml, c] += h * ezi[c]


_libm = None


def twiddle(t, freq, dt, dtype):
    """cos, sin of 2*pi*freq*dt*t as the C programs evaluate them: the
    product in double, then cos/sin in double (FP64) or narrowed to float and
    cosf/sinf of the host libm (FP32; fd2d/clang/test_3_4.c:42-43)."""
    import ctypes
    import math
    global _libm
    a = 2 * math.pi * float(freq) * float(dt) * t
    if np.dtype(dtype) == np.float64:
        return np.float64(math.cos(a)), np.float64(math.sin(a))
    if _libm is None:
        _libm = ctypes.CDLL("libm.so.6")
        for f in (_libm.cosf, _libm.sinf):
            f.restype, f.argtypes = ctypes.c_float, [ctypes.c_float]
    af = float(np.float32(a))
    return np.float32(_libm.cosf(af)), np.float32(_libm.sinf(af))


def fourier(t, freq, dt, ezi, ez, ft, probe=6):
    """fd2d/clang/test_3_4.c:38-53: running DFT of ez (planes ft["r_pt"][n],
    ft["i_pt"][n]) and of the incident probe ezi[probe]."""
    for n, f in enumerate(freq):
        c, s = twiddle(t, f, dt, ez.dtype)
        ft["r_in"][n] += c * ezi[probe]
        ft["i_in"][n] -= s * ezi[probe]
        ft["r_pt"][n] += c * ez
        ft["i_pt"][n] -= s * ez


def tfsf_run(state, inc, nsteps, npml, pml, naz, nbz, inc_wave, freq=(), dt=0.0, ft=None,
             t0=0, inc_pos=3):
    """The main loop of fd2d/clang/test_3_4.c:223-233 for nsteps from t0:
    state {dz, iz, ez, hx, hy, ihx, ihy}, inc {ezi, hxi, bc}; inc_wave[t-1]
    is the incident source of step t (the program's gaussian(t, 20, 8))."""
    nx, ny = state["dz"].shape
    for s in range(nsteps):
        t = t0 + s + 1
        ezinct(inc["ezi"], inc["hxi"], inc["bc"])
        dfield(nx, ny, pml, state["dz"], state["hx"], state["hy"])
        inc["ezi"][inc_pos] = inc_wave[t - 1]
        inctdz(nx, ny, npml, inc["hxi"], state["dz"])
        efield(naz, nbz, state["dz"], state["iz"], state["ez"])
        if len(freq):
            fourier(t, freq, dt, inc["ezi"], state["ez"], ft)
        hxinct(inc["ezi"], inc["hxi"])
        hfield(nx, ny, pml, state["ez"], state["ihx"], state["ihy"], state["hx"], state["hy"])
        incthx(nx, ny, npml, inc["ezi"], state["hx"])
        incthy(nx, ny, npml, inc["ezi"], state["hy"])


# ------------------------------------------------ 1-D D/E form (lossy, Debye)
# fd1d/clang/test_2_3.c:26-90 (FP32), fd1d/clang/fd1d_2_3.c (FP64),
# fd1d/clang/test_2_2.c (lossy medium, no relaxation term).

def dielectric_dform(nx, dt, chi, tau, epsr, sigma, debye, dtype):
    """test_2_3.c:75-90 / test_2_2.c:73-86 in the program's precision: free
    space below nx/2; ncx = exp(-dt/tau) in double, stored."""
    import math
    f = np.dtype(dtype).type
    dt, chi, tau, epsr, sigma = f(dt), f(chi), f(tau), f(epsr), f(sigma)
    eps0 = f(8.854e-12)
    md = {"nax": np.ones(nx, dtype), "nbx": np.zeros(nx, dtype)}
    lo = nx // 2
    if debye:
        md["ncx"] = np.zeros(nx, dtype)
        md["ndx"] = np.zeros(nx, dtype)
        md["nax"][lo:] = f(1) / (epsr + sigma * dt / eps0 + chi * dt / tau)
        md["ncx"][lo:] = f(math.exp(float(-dt / tau)))
        md["ndx"][lo:] = chi * dt / tau
    else:
        md["nax"][lo:] = f(1) / (epsr + sigma * dt / eps0)
    md["nbx"][lo:] = sigma * dt / eps0
    return md


def dxfield(dx, hy, src_pos, src_val):
    """test_2_3.c:46-52: update for i in [1, nx), then the additive source."""
    h = dx.dtype.type(0.5)
    dx[1:] += h * (hy[:-1] - hy[1:])
    if src_pos >= 1:
        dx[src_pos] += src_val


def exfield_dform(md, dx, ix, sx, ex):
    """test_2_3.c:55-62 (Debye) / test_2_2.c:54-60 (lossy)."""
    if "ncx" in md:
        ex[1:] = md["nax"][1:] * (dx[1:] - ix[1:] - md["ncx"][1:] * sx[1:])
        ix[1:] += md["nbx"][1:] * ex[1:]
        sx[1:] = md["ncx"][1:] * sx[1:] + md["ndx"][1:] * ex[1:]
    else:
        ex[1:] = md["nax"][1:] * (dx[1:] - ix[1:])
        ix[1:] += md["nbx"][1:] * ex[1:]


def hyfield1d(ex, hy, bc):
    """test_2_3.c:65-72: the two-step boundary at both ends, then H."""
    nx = len(ex)
    h = ex.dtype.type(0.5)
    ex[0] = bc[0]; bc[0] = bc[1]; bc[1] = ex[1]
    ex[nx-1] = bc[3]; bc[3] = bc[2]; bc[2] = ex[nx-2]
    hy[:-1] += h * (ex[:-1] - ex[1:])


def dform_run(st, md, nsteps, src_wave, freq=(), dt=0.0, ft=None, t0=0, src_pos=1, probes=()):
    """The main loop of test_2_3.c (dxfield, exfield, fourier, hyfield) for
    nsteps from t0; st {dx, ix, sx, ex, hy, bc}; returns the probe traces."""
    out = np.zeros((len(probes), nsteps), st["ex"].dtype)
    for s in range(nsteps):
        t = t0 + s + 1
        dxfield(st["dx"], st["hy"], src_pos, src_wave[t - 1])
        exfield_dform(md, st["dx"], st["ix"], st["sx"], st["ex"])
        if len(freq):
            fourier(t, freq, dt, st["ex"], st["ex"], ft, probe=10)
        hyfield1d(st["ex"], st["hy"], st["bc"])
        for p, k in enumerate(probes):
            out[p, s] = st["ex"][k]
    return out
