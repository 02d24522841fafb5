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
