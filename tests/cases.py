"""Run parameters of the golden fixtures (tests/golden/*.npz), shared by the
oracle tests (CPU) and the HIP parity tests (GPU)."""
from __future__ import annotations

import numpy as np

from conftest import golden

FD1D_CASES = [f"{c}_{p}" for p in ("f32", "f64")
              for c in ("fd1d_1_1", "fd1d_c1", "fd1d_1_3", "fd1d_1_4", "fd1d_1_5")]
FD2D_CASES = [f"{c}_{p}" for p in ("f32", "f64") for c in ("fd2d_3_1", "fd2d_3_2", "fd2d_c2")]
PML_KEYS = ("fx1", "fx2", "fx3", "fy1", "fy2", "fy3", "gx2", "gx3", "gy2", "gy3")


def fd1d_case(name: str, waveform):
    """dict of the run: nx, ns, dtype, ca, cb, abc, src_mode, src_pos, wave, probes,
    expected outputs.  ``waveform(kind, ns, dtype, **kw)`` evaluates the source."""
    meta, a = golden(name)
    dt = np.float32 if meta["dtype"] == "float32" else np.float64
    nx, ns, form = meta["nx"], meta["ns"], meta["form"]
    run = dict(nx=nx, ns=ns, dtype=dt, ca=None, cb=None, abc=form != "free", probes=meta["probes"],
               snap_t=meta.get("snap_t", []), meta=meta, exp=a)
    if form in ("free", "abc"):       # fd1d_1_1 / fd1d_1_2: hard Gaussian in the middle
        run.update(src_mode=1, src_pos=nx // 2,
                   wave=waveform("gaussian", ns, dt, center=40, spread=12.0))
    elif form == "cb" and meta.get("source") == "sinusoidal":
        # fd1d_1_4: cb, additive 700 MHz sinusoid at ex[1]
        run.update(cb=a["in_cb"], src_mode=2, src_pos=1,
                   wave=waveform("sinusoidal", ns, dt, freq=700e6))
    elif form == "cb":                # fd1d_1_3: additive Gaussian at ex[1]
        run.update(cb=a["in_cb"], src_mode=2, src_pos=1,
                   wave=waveform("gaussian", ns, dt, center=40, spread=12.0))
    else:                             # fd1d_1_5: ca/cb, additive 700 MHz sinusoid at ex[1]
        run.update(ca=a["in_ca"], cb=a["in_cb"], src_mode=2, src_pos=1,
                   wave=waveform("sinusoidal", ns, dt, freq=700e6))
    return run


def fd2d_case(name: str, waveform):
    meta, a = golden(name)
    dt = np.float32 if meta["dtype"] == "float32" else np.float64
    nx, ny, ns = meta["nx"], meta["ny"], meta["ns"]
    src = meta["source"]
    if src["kind"] == "gaussian":
        w = waveform("gaussian", ns, dt, center=src["t0"], spread=src["spread"])
    else:
        w = waveform("sinusoidal", ns, dt, freq=src["freq"])
    pml = {k: a["in_" + k] for k in PML_KEYS} if "in_fx1" in a else None
    return dict(nx=nx, ny=ny, ns=ns, dtype=dt, pml=pml, naz=a["in_naz"],
                nbz=a.get("in_nbz"), sources=[(src["i"], src["j"])], wave=w,
                probes=[tuple(p) for p in meta["probes"]], meta=meta, exp=a)


TFSF_CASES = ("fd2d_tfsf_f32", "fd2d_tfsf_f64", "fd2d_3_3_f32", "fd2d_3_3_f64")


def tfsf_case(name: str, gaussian):
    """A plane-wave fixture (fd2d/clang/test_3_4.c / fd2d_3_4.c main loop, or
    the lossless test_3_3.c / fd2d_3_3.c without DFT: nbz None, no
    frequencies): coefficients, DFT frequencies, the incident source values
    gaussian(t, 20, 8) of steps 1..ns, and the expected outputs."""
    meta, a = golden(name)
    dt = np.float32 if meta["dtype"] == "float32" else np.float64
    inc = meta["incident"]
    wave = np.array([gaussian(t, inc["t0"], inc["spread"], dt) for t in range(1, meta["ns"] + 1)],
                    dt)
    pml = {k: a["in_" + k] for k in PML_KEYS}
    return dict(nx=meta["nx"], ny=meta["ny"], ns=meta["ns"], npml=meta["npml"], dtype=dt,
                pml=pml, naz=a["in_naz"], nbz=a.get("in_nbz"), freq=a["in_freq"],
                dt=dt(meta["dt"]), inc_pos=inc["pos"], wave=wave, meta=meta, exp=a)


DFORM_CASES = ("fd1d_debye_f32", "fd1d_debye_f64", "fd1d_lossy_f32", "fd1d_lossy_f64",
               "fd1d_2_1_f32", "fd1d_2_1_f64")


def dform_case(name: str, gaussian, sinusoidal=None):
    """A 1-D D/E-form fixture (fd1d/clang/test_2_3.c, test_2_2.c or, with the
    700 MHz sinusoid and no DFT, test_2_1.c main loop)."""
    meta, a = golden(name)
    dt = np.float32 if meta["dtype"] == "float32" else np.float64
    src = meta["source"]
    ts = range(1, meta["ns"] + 1)
    if src["kind"] == "sinusoidal":
        wave = np.array([sinusoidal(t, 0.01, src["freq"], dt) for t in ts], dt)
    else:
        wave = np.array([gaussian(t, src["t0"], src["spread"], dt) for t in ts], dt)
    return dict(nx=meta["nx"], ns=meta["ns"], dtype=dt, debye=meta["debye"], freq=a["in_freq"],
                dt=dt(meta["dt"]), src_pos=src["pos"], wave=wave, probes=tuple(meta["probes"]),
                meta=meta, exp=a)
