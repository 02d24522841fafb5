"""Write the golden fixtures of tests/golden/ from the reference's own C programs.

Run in a container where the reference is mounted (it reads nothing else):

    make -C oracle ref            # builds oracle/_ref/libref_<program>.so
    python tests/golden/make_golden.py

Each reference program (fd1d/clang/*.c, fd2d/clang/*.c) is compiled unchanged
by oracle/Makefile into its own shared library; this script calls the
program's own functions (exfield/hyfield, dfield/efield/hfield, pmlparam,
dielectric, gaussian, sinusoidal) through ctypes, in the order of the
program's main() loop, at small sizes, and stores inputs and outputs as
compressed .npz files.  The fixtures are data (inputs and expected outputs);
no reference source is copied.

Cases (sizes are chosen so that every field path is exercised, including odd
and non-square grids):
  fd1d_1_1_{f32,f64}   test_1_1.c / fd1d_1_1.c   nx=200, 150 steps, no boundary
  fd1d_c1_{f32,f64}    test_1_2.c / fd1d_1_2.c   nx=200, 500 steps, hard Gaussian at 100,
                                                  probe ex[150], snapshots every 100 steps
  fd1d_1_3_{f32,f64}   test_1_3.c / fd1d_1_3.c   nx=200, 300 steps, cb, additive Gaussian
  fd1d_1_4_{f32,f64}   test_1_4.c / fd1d_1_4.c   nx=200, 300 steps, cb, additive 700 MHz sinusoid
  fd1d_1_5_{f32,f64}   test_1_5.c / fd1d_1_5.c   nx=200, 400 steps, ca/cb, additive sinusoid
  fd1d_debye_{f32,f64} test_2_3.c / fd1d_2_3.c   nx=300, 600 steps, D/E form, Debye half space,
                                                  additive Gaussian at 1, DFT at 3 frequencies
  fd1d_lossy_{f32,f64} test_2_2.c / fd1d_2_2.c   the same with the lossy half space (no Debye)
  fd1d_2_1_{f32,f64}   test_2_1.c / fd1d_2_1.c   nx=300, 500 steps, D/E form, lossy half space
                                                  (eps_r 4, sigma 0.04), additive 700 MHz sinusoid
                                                  at dx[1], no DFT
  fd2d_3_1_{f32,f64}   test_3_1.c / fd2d_3_1.c   60x70, 60 steps, free space
  fd2d_3_2_{f32,f64}   test_3_2.c / fd2d_3_2.c   64x80, npml=8, 100 steps, PML, sinusoid
  fd2d_c2_{f32,f64}    test_3_4.c / fd2d_3_4.c   97x131, npml=8, 150 steps: PML dfield/hfield,
                                                  lossy efield (naz, nbz, iz), hard Gaussian
                                                  (t0=20, spread 6) at (nx/2, ny/2), 4 probes
  fd2d_tfsf_{f32,f64}  test_3_4.c / fd2d_3_4.c   70x90, npml=8, 160 steps: plane wave through the
                                                  TF/SF box (incident grid + two-step boundary),
                                                  PML, lossy cylinder, running DFT at 3 frequencies
  fd2d_3_3_{f32,f64}   test_3_3.c / fd2d_3_3.c   70x90, npml=8, 160 steps: plane wave through the
                                                  TF/SF box, PML, lossless medium (efield ez =
                                                  naz*dz; naz of a lossless cylinder), no DFT
  fd2d_dielectric_c    fd2d_3_4.c dielectric (FP64, integer sub-sampling), 97x131
  fd2d_pmlparam_c      fd2d_3_2.c / test_3_2.c pmlparam (FP64 / FP32) for npml 8, 32, 64
"""
from __future__ import annotations

import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_DIR = os.path.join(REPO, "oracle", "_ref")


def load(name):
    path = os.path.join(REF_DIR, f"libref_{name}.so")
    if not os.path.exists(path):
        sys.exit(f"missing {path}: run `make -C oracle ref` where the reference is mounted")
    return C.CDLL(path)


def P(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_float if a.dtype == np.float32 else C.c_double))


def real(dt):
    return C.c_float if dt == np.float32 else C.c_double


def pml_struct(dt):
    ptr = C.POINTER(real(dt))

    class PmLayer(C.Structure):
        _fields_ = [(k, ptr) for k in ("fx1", "fx2", "fx3", "fy1", "fy2", "fy3",
                                        "gx2", "gx3", "gy2", "gy3")]
    return PmLayer


def medium_struct(dt):
    ptr = C.POINTER(real(dt))

    class Medium(C.Structure):
        _fields_ = [("naz", ptr), ("nbz", ptr)]
    return Medium


PML_KEYS = ("fx1", "fx2", "fx3", "fy1", "fy2", "fy3", "gx2", "gx3", "gy2", "gy3")


def save(name, meta, **arrays):
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, meta=np.array(json.dumps(meta)), **arrays)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


# ------------------------------------------------------------------ 1-D

def case_fd1d(prog, tag, dt, nx, ns, form, probes=(), snap_every=0, extra=None):
    lib = load(prog)
    R = real(dt)
    ex = np.zeros(nx, dt)
    hy = np.zeros(nx, dt)
    bc = np.zeros(4, dt)
    inputs = {}
    ca = cb = None
    if form == "cb":                       # fd1d/clang/test_1_3.c:37-43
        lib.dielectric.restype = C.POINTER(R)
        p = lib.dielectric(C.c_int(nx), R(extra["epsr"]))
        cb = np.ctypeslib.as_array(p, shape=(nx,)).copy()
        inputs["cb"] = cb
    elif form == "cacb":                   # fd1d/clang/test_1_5.c:45-57
        ptr = C.POINTER(R)

        class Tuple(C.Structure):
            _fields_ = [("ca", ptr), ("cb", ptr)]
        lib.dielectric.restype = Tuple
        ds = R(0.01)
        dtv = np.float32(np.float32(0.01) / np.float32(6e8)) if dt == np.float32 else 0.01 / 6e8
        tup = lib.dielectric(C.c_int(nx), R(dtv), R(extra["epsr"]), R(extra["sigma"]))
        ca = np.ctypeslib.as_array(tup.ca, shape=(nx,)).copy()
        cb = np.ctypeslib.as_array(tup.cb, shape=(nx,)).copy()
        inputs["ca"] = ca
        inputs["cb"] = cb
        inputs["dt"] = np.array(dtv, dtype=dt)
        del ds
    probe_out = np.zeros((len(probes), ns), dt)
    snaps_ex, snaps_hy, snap_t = [], [], []
    for t in range(1, ns + 1):
        if form == "free" or form == "abc":
            lib.exfield(C.c_int(t), C.c_int(nx), P(ex), P(hy))
        elif form == "cb":
            lib.exfield(C.c_int(t), C.c_int(nx), P(cb), P(ex), P(hy))
        else:
            lib.exfield(C.c_int(t), C.c_int(nx), P(ca), P(cb), P(ex), P(hy))
        if form == "free":
            lib.hyfield(C.c_int(nx), P(ex), P(hy))
        else:
            lib.hyfield(C.c_int(nx), P(ex), P(hy), P(bc))
        for p, k in enumerate(probes):
            probe_out[p, t - 1] = ex[k]
        if snap_every and t % snap_every == 0:
            snaps_ex.append(ex.copy())
            snaps_hy.append(hy.copy())
            snap_t.append(t)
    meta = dict(program=prog, form=form, nx=nx, ns=ns, dtype=str(np.dtype(dt)),
                probes=list(probes), snap_t=snap_t, **(extra or {}))
    out = dict(ex=ex, hy=hy, bc=bc, probe=probe_out)
    if snap_every:
        out["snap_ex"] = np.array(snaps_ex)
        out["snap_hy"] = np.array(snaps_hy)
    save(tag, meta, **{f"in_{k}": v for k, v in inputs.items()},
         **{f"out_{k}": v for k, v in out.items()})


# ------------------------------------------------------------------ 2-D

def ref_pml(lib, dt, nx, ny, npml):
    """The program's own pmlparam over the default fills of its main()."""
    arrs = {k: np.zeros(nx if k[1] == "x" else ny, dt) for k in PML_KEYS}
    for k in ("fx2", "fx3", "gx2", "gx3", "fy2", "fy3", "gy2", "gy3"):
        arrs[k][:] = 1
    S = pml_struct(dt)
    s = S(*[P(arrs[k]) for k in PML_KEYS])
    lib.pmlparam(C.c_int(nx), C.c_int(ny), C.c_int(npml), C.byref(s))
    return arrs


def case_fd2d_3_1(tag, prog, dt, nx, ny, ns, probes):
    lib = load(prog)
    f = {k: np.zeros((nx, ny), dt) for k in ("dz", "ez", "hx", "hy")}
    naz = np.ones((nx, ny), dt)
    probe_out = np.zeros((len(probes), ns), dt)
    for t in range(1, ns + 1):
        lib.dfield(C.c_int(t), C.c_int(nx), C.c_int(ny), P(f["dz"]), P(f["hx"]), P(f["hy"]))
        lib.efield(C.c_int(nx), C.c_int(ny), P(naz), P(f["dz"]), P(f["ez"]))
        lib.hfield(C.c_int(nx), C.c_int(ny), P(f["ez"]), P(f["hx"]), P(f["hy"]))
        for p, (i, j) in enumerate(probes):
            probe_out[p, t - 1] = f["ez"][i, j]
    meta = dict(program=prog, nx=nx, ny=ny, ns=ns, dtype=str(np.dtype(dt)), probes=probes,
                source=dict(kind="gaussian", i=nx // 2, j=ny // 2, t0=20, spread=6.0, mode="hard"))
    save(tag, meta, in_naz=naz, **{f"out_{k}": v for k, v in f.items()}, out_probe=probe_out)


def case_fd2d_3_2(tag, prog, dt, nx, ny, ns, npml, probes):
    lib = load(prog)
    pml = ref_pml(lib, dt, nx, ny, npml)
    S = pml_struct(dt)
    s = S(*[P(pml[k]) for k in PML_KEYS])
    f = {k: np.zeros((nx, ny), dt) for k in ("dz", "ez", "hx", "hy", "ihx", "ihy")}
    naz = np.ones((nx, ny), dt)
    probe_out = np.zeros((len(probes), ns), dt)
    for t in range(1, ns + 1):
        lib.dfield(C.c_int(t), C.c_int(nx), C.c_int(ny), C.byref(s), P(f["dz"]), P(f["hx"]), P(f["hy"]))
        lib.efield(C.c_int(nx), C.c_int(ny), P(naz), P(f["dz"]), P(f["ez"]))
        lib.hfield(C.c_int(nx), C.c_int(ny), C.byref(s), P(f["ez"]), P(f["ihx"]), P(f["ihy"]),
                   P(f["hx"]), P(f["hy"]))
        for p, (i, j) in enumerate(probes):
            probe_out[p, t - 1] = f["ez"][i, j]
    meta = dict(program=prog, nx=nx, ny=ny, ns=ns, npml=npml, dtype=str(np.dtype(dt)),
                probes=probes,
                source=dict(kind="sinusoidal", i=nx // 2 - 5, j=ny // 2 - 5, ds=0.01,
                            freq=1500e6, mode="hard"))
    save(tag, meta, in_naz=naz, **{f"in_{k}": v for k, v in pml.items()},
         **{f"out_{k}": v for k, v in f.items()}, out_probe=probe_out)


def case_fd2d_c2(tag, prog, prog64, dt, nx, ny, ns, npml, rgrid, epsr, sigma, probes):
    """dfield/efield/hfield of fd2d_3_4 (PML + lossy dielectric), driven with
    the config-2 shape: hard Gaussian (t0=20, spread 6) at (nx/2, ny/2)."""
    lib = load(prog)
    lib64 = load(prog64)
    # coefficients: the FP64 programs' own pmlparam and dielectric, narrowed
    # to the state dtype the way the PyCUDA scripts store them
    pml = {k: v.astype(dt) for k, v in ref_pml(lib64, np.float64, nx, ny, npml).items()}
    M64 = medium_struct(np.float64)
    lib64.dielectric.restype = M64
    dtv = float(np.float32(np.float32(0.01) / np.float32(6e8)))
    md = lib64.dielectric(C.c_int(nx), C.c_int(ny), C.c_int(npml), C.c_int(rgrid),
                          C.c_double(dtv), C.c_double(epsr), C.c_double(sigma))
    naz = np.ctypeslib.as_array(md.naz, shape=(nx, ny)).copy().astype(dt)
    nbz = np.ctypeslib.as_array(md.nbz, shape=(nx, ny)).copy().astype(dt)
    S = pml_struct(dt)
    s = S(*[P(pml[k]) for k in PML_KEYS])
    M = medium_struct(dt)
    m = M(P(naz), P(nbz))
    R = real(dt)
    lib.gaussian.restype = R
    lib.gaussian.argtypes = [C.c_int, C.c_int, R]
    f = {k: np.zeros((nx, ny), dt) for k in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy")}
    ezi = np.zeros(max(ny, 8), dt)       # written by the program's dfield (its own plane-wave source)
    probe_out = np.zeros((len(probes), ns), dt)
    si, sj = nx // 2, ny // 2
    for t in range(1, ns + 1):
        lib.dfield(C.c_int(t), C.c_int(nx), C.c_int(ny), C.byref(s), P(ezi), P(f["dz"]),
                   P(f["hx"]), P(f["hy"]))
        f["dz"][si, sj] = lib.gaussian(t, 20, 6.0)
        lib.efield(C.c_int(nx), C.c_int(ny), C.byref(m), P(f["dz"]), P(f["iz"]), P(f["ez"]))
        lib.hfield(C.c_int(nx), C.c_int(ny), C.byref(s), P(f["ez"]), P(f["ihx"]), P(f["ihy"]),
                   P(f["hx"]), P(f["hy"]))
        for p, (i, j) in enumerate(probes):
            probe_out[p, t - 1] = f["ez"][i, j]
    meta = dict(program=prog, coef_program=prog64, nx=nx, ny=ny, ns=ns, npml=npml,
                rgrid=rgrid, epsr=epsr, sigma=sigma, dt=dtv, dtype=str(np.dtype(dt)),
                probes=probes,
                source=dict(kind="gaussian", i=si, j=sj, t0=20, spread=6.0, mode="hard"))
    save(tag, meta, in_naz=naz, in_nbz=nbz, **{f"in_{k}": v for k, v in pml.items()},
         **{f"out_{k}": v for k, v in f.items()}, out_probe=probe_out)


def case_fd2d_tfsf(tag, prog, prog64, dt, nx, ny, ns, npml, rgrid, epsr, sigma, freqs):
    """The plane-wave program's main loop (ezinct, dfield with the incident
    source ezi[3] = gaussian(t, 20, 8), inctdz, efield, fourier, hxinct,
    hfield, incthx, incthy), PML + lossy cylinder + running DFT, with the
    program's own functions in the order of its main()."""
    lib = load(prog)
    lib64 = load(prog64)
    pml = {k: v.astype(dt) for k, v in ref_pml(lib64, np.float64, nx, ny, npml).items()}
    M64 = medium_struct(np.float64)
    lib64.dielectric.restype = M64
    if dt == np.float32:
        dtv = np.float32(np.float32(0.01) / np.float32(6e8))
    else:
        dtv = np.float64(0.01 / 6e8)
    md = lib64.dielectric(C.c_int(nx), C.c_int(ny), C.c_int(npml), C.c_int(rgrid),
                          C.c_double(float(dtv)), C.c_double(epsr), C.c_double(sigma))
    naz = np.ctypeslib.as_array(md.naz, shape=(nx, ny)).copy().astype(dt)
    nbz = np.ctypeslib.as_array(md.nbz, shape=(nx, ny)).copy().astype(dt)
    S = pml_struct(dt)
    s = S(*[P(pml[k]) for k in PML_KEYS])
    M = medium_struct(dt)
    m = M(P(naz), P(nbz))
    R = real(dt)
    ptr = C.POINTER(R)

    class Ftrans(C.Structure):
        _fields_ = [(k, ptr) for k in ("r_pt", "i_pt", "r_in", "i_in")]
    nf = len(freqs)
    freq = np.asarray(freqs, dt)
    ft = {"r_pt": np.zeros((nf, nx, ny), dt), "i_pt": np.zeros((nf, nx, ny), dt),
          "r_in": np.zeros(nf, dt), "i_in": np.zeros(nf, dt)}
    fts = Ftrans(*[P(ft[k]) for k in ("r_pt", "i_pt", "r_in", "i_in")])
    f = {k: np.zeros((nx, ny), dt) for k in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy")}
    ezi = np.zeros(ny, dt)
    hxi = np.zeros(ny, dt)
    bc = np.zeros(4, dt)
    I = C.c_int
    for t in range(1, ns + 1):
        lib.ezinct(I(ny), P(ezi), P(hxi), P(bc))
        lib.dfield(I(t), I(nx), I(ny), C.byref(s), P(ezi), P(f["dz"]), P(f["hx"]), P(f["hy"]))
        lib.inctdz(I(nx), I(ny), I(npml), P(hxi), P(f["dz"]))
        lib.efield(I(nx), I(ny), C.byref(m), P(f["dz"]), P(f["iz"]), P(f["ez"]))
        lib.fourier(I(t), I(nf), I(nx), I(ny), R(dtv), P(freq), P(ezi), P(f["ez"]), C.byref(fts))
        lib.hxinct(I(ny), P(ezi), P(hxi))
        lib.hfield(I(nx), I(ny), C.byref(s), P(f["ez"]), P(f["ihx"]), P(f["ihy"]), P(f["hx"]),
                   P(f["hy"]))
        lib.incthx(I(nx), I(ny), I(npml), P(ezi), P(f["hx"]))
        lib.incthy(I(nx), I(ny), I(npml), P(ezi), P(f["hy"]))
    meta = dict(program=prog, coef_program=prog64, nx=nx, ny=ny, ns=ns, npml=npml, rgrid=rgrid,
                epsr=epsr, sigma=sigma, dt=float(dtv), dtype=str(np.dtype(dt)), freqs=list(freqs),
                incident=dict(pos=3, t0=20, spread=8.0), dft_probe=6)
    save(tag, meta, in_naz=naz, in_nbz=nbz, in_freq=freq, **{f"in_{k}": v for k, v in pml.items()},
         **{f"out_{k}": v for k, v in f.items()}, out_ezi=ezi, out_hxi=hxi, out_bc=bc,
         **{f"out_ft_{k}": v for k, v in ft.items()})


def case_fd1d_dform(tag, prog, dt, nx, ns, debye, freqs, probes, source=None, epsr=2.0,
                    sigma=0.01):
    """The 1-D D/E-form program's main loop (dxfield with its additive source
    at cell 1 -- gaussian(t, 50, 10), or the 700 MHz sinusoid of test_2_1.c
    -- exfield, fourier when the program has a DFT, hyfield) with its own
    dielectric: lossy (test_2_1.c, test_2_2.c) or Debye (test_2_3.c), running
    DFT at freqs (none for test_2_1.c)."""
    lib = load(prog)
    R = real(dt)
    ptr = C.POINTER(R)
    keys = ("nax", "nbx", "ncx", "ndx") if debye else ("nax", "nbx")

    class Medium(C.Structure):
        _fields_ = [(k, ptr) for k in keys]

    class Ftrans(C.Structure):
        _fields_ = [(k, ptr) for k in ("r_pt", "i_pt", "r_in", "i_in")]
    if dt == np.float32:
        dtv = np.float32(np.float32(0.01) / np.float32(6e8))
    else:
        dtv = np.float64(0.01 / 6e8)
    chi, tau = 2.0, 0.001e-6
    lib.dielectric.restype = Medium
    if debye:
        md = lib.dielectric(C.c_int(nx), R(dtv), R(chi), R(tau), R(epsr), R(sigma))
    else:
        md = lib.dielectric(C.c_int(nx), R(dtv), R(epsr), R(sigma))
    mda = {k: np.ctypeslib.as_array(getattr(md, k), shape=(nx,)).copy() for k in keys}
    nf = len(freqs)
    freq = np.asarray(freqs, dt)
    ft = {"r_pt": np.zeros((nf, nx), dt), "i_pt": np.zeros((nf, nx), dt),
          "r_in": np.zeros(nf, dt), "i_in": np.zeros(nf, dt)}
    fts = Ftrans(*[P(ft[k]) for k in ("r_pt", "i_pt", "r_in", "i_in")])
    f = {k: np.zeros(nx, dt) for k in ("dx", "ix", "sx", "ex", "hy")}
    bc = np.zeros(4, dt)
    probe_out = np.zeros((len(probes), ns), dt)
    I = C.c_int
    for t in range(1, ns + 1):
        lib.dxfield(I(t), I(nx), P(f["dx"]), P(f["hy"]))
        if debye:
            lib.exfield(I(nx), C.byref(md), P(f["dx"]), P(f["ix"]), P(f["sx"]), P(f["ex"]))
        else:
            lib.exfield(I(nx), C.byref(md), P(f["dx"]), P(f["ix"]), P(f["ex"]))
        if nf:
            lib.fourier(I(t), I(nf), I(nx), R(dtv), P(freq), P(f["ex"]), C.byref(fts))
        lib.hyfield(I(nx), P(f["ex"]), P(f["hy"]), P(bc))
        for p, k in enumerate(probes):
            probe_out[p, t - 1] = f["ex"][k]
    meta = dict(program=prog, nx=nx, ns=ns, debye=debye, dt=float(dtv), chi=chi, tau=tau,
                epsr=epsr, sigma=sigma, dtype=str(np.dtype(dt)), freqs=list(freqs),
                source=source or dict(kind="gaussian", pos=1, t0=50, spread=10.0, mode="add"),
                dft_probe=10, probes=list(probes))
    outs = {f"out_{k}": v for k, v in f.items() if debye or k != "sx"}
    save(tag, meta, **{f"out_md_{k}": v for k, v in mda.items()}, in_freq=freq, **outs,
         out_bc=bc, out_probe=probe_out, **{f"out_ft_{k}": v for k, v in ft.items()})


def case_fd2d_3_3(tag, prog, prog_coef, dt, nx, ny, ns, npml, rgrid, epsr):
    """The lossless plane-wave program's main loop (ezinct, dfield with the
    incident source ezi[3] = gaussian(t, 20, 8), inctdz, efield ez = naz*dz,
    hxinct, hfield, incthx, incthy) with its own pmlparam; naz is that of a
    lossless cylinder (the FP64 lossy program's dielectric with sigma = 0,
    stored in the program's precision) so that the E-from-D factor varies."""
    lib = load(prog)
    pml = ref_pml(lib, dt, nx, ny, npml)
    lib64 = load(prog_coef)
    M64 = medium_struct(np.float64)
    lib64.dielectric.restype = M64
    dtv = float(np.float32(np.float32(0.01) / np.float32(6e8)))
    md = lib64.dielectric(C.c_int(nx), C.c_int(ny), C.c_int(npml), C.c_int(rgrid),
                          C.c_double(dtv), C.c_double(epsr), C.c_double(0.0))
    naz = np.ctypeslib.as_array(md.naz, shape=(nx, ny)).copy().astype(dt)
    S = pml_struct(dt)
    s = S(*[P(pml[k]) for k in PML_KEYS])
    f = {k: np.zeros((nx, ny), dt) for k in ("dz", "ez", "hx", "hy", "ihx", "ihy")}
    ezi = np.zeros(ny, dt)
    hxi = np.zeros(ny, dt)
    bc = np.zeros(4, dt)
    I = C.c_int
    for t in range(1, ns + 1):
        lib.ezinct(I(ny), P(ezi), P(hxi), P(bc))
        lib.dfield(I(t), I(nx), I(ny), C.byref(s), P(ezi), P(f["dz"]), P(f["hx"]), P(f["hy"]))
        lib.inctdz(I(nx), I(ny), I(npml), P(hxi), P(f["dz"]))
        lib.efield(I(nx), I(ny), P(naz), P(f["dz"]), P(f["ez"]))
        lib.hxinct(I(ny), P(ezi), P(hxi))
        lib.hfield(I(nx), I(ny), C.byref(s), P(f["ez"]), P(f["ihx"]), P(f["ihy"]), P(f["hx"]),
                   P(f["hy"]))
        lib.incthx(I(nx), I(ny), I(npml), P(ezi), P(f["hx"]))
        lib.incthy(I(nx), I(ny), I(npml), P(ezi), P(f["hy"]))
    meta = dict(program=prog, coef_program=prog_coef, nx=nx, ny=ny, ns=ns, npml=npml,
                rgrid=rgrid, epsr=epsr, sigma=0.0, dt=dtv, dtype=str(np.dtype(dt)), freqs=[],
                incident=dict(pos=3, t0=20, spread=8.0), dft_probe=6)
    save(tag, meta, in_naz=naz, in_freq=np.zeros(0, dt), **{f"in_{k}": v for k, v in pml.items()},
         **{f"out_{k}": v for k, v in f.items()}, out_ezi=ezi, out_hxi=hxi, out_bc=bc)


def case_dielectric_c(tag, nx, ny, npml, rgrid, epsr, sigma):
    lib = load("fd2d_3_4")
    M64 = medium_struct(np.float64)
    lib.dielectric.restype = M64
    dtv = float(np.float32(np.float32(0.01) / np.float32(6e8)))
    md = lib.dielectric(C.c_int(nx), C.c_int(ny), C.c_int(npml), C.c_int(rgrid),
                        C.c_double(dtv), C.c_double(epsr), C.c_double(sigma))
    naz = np.ctypeslib.as_array(md.naz, shape=(nx, ny)).copy()
    nbz = np.ctypeslib.as_array(md.nbz, shape=(nx, ny)).copy()
    save(tag, dict(program="fd2d_3_4", nx=nx, ny=ny, npml=npml, rgrid=rgrid, epsr=epsr,
                   sigma=sigma, dt=dtv), out_naz=naz, out_nbz=nbz)


def case_pmlparam_c(tag, nx, ny, npmls):
    out = {}
    for npml in npmls:
        for prog, dt in (("fd2d_3_2", np.float64), ("test_3_2", np.float32)):
            arrs = ref_pml(load(prog), dt, nx, ny, npml)
            for k, v in arrs.items():
                out[f"out_{prog}_{npml}_{k}"] = v
    save(tag, dict(nx=nx, ny=ny, npmls=list(npmls)), **out)


def main():
    for dt, suf in ((np.float32, "f32"), (np.float64, "f64")):
        p = "test" if dt == np.float32 else "fd1d"
        case_fd1d(f"{p}_1_1", f"fd1d_1_1_{suf}", dt, 200, 150, "free", probes=(150,))
        case_fd1d(f"{p}_1_2", f"fd1d_c1_{suf}", dt, 200, 500, "abc", probes=(150,), snap_every=100)
        case_fd1d(f"{p}_1_3", f"fd1d_1_3_{suf}", dt, 200, 300, "cb", probes=(150,),
                  extra=dict(epsr=4.0))
        case_fd1d(f"{p}_1_4", f"fd1d_1_4_{suf}", dt, 200, 300, "cb", probes=(150,),
                  extra=dict(epsr=4.0, source="sinusoidal"))
        case_fd1d(f"{p}_1_5", f"fd1d_1_5_{suf}", dt, 200, 400, "cacb", probes=(150,),
                  extra=dict(epsr=4.0, sigma=0.04))
        q = "test" if dt == np.float32 else "fd2d"
        case_fd1d_dform(f"fd1d_debye_{suf}", f"{p}_2_3", dt, 300, 600, True,
                        (50e6, 200e6, 500e6), probes=(100, 200))
        case_fd1d_dform(f"fd1d_lossy_{suf}", f"{p}_2_2", dt, 300, 600, False,
                        (50e6, 200e6, 500e6), probes=(100, 200))
        case_fd1d_dform(f"fd1d_2_1_{suf}", f"{p}_2_1", dt, 300, 500, False, (), probes=(100, 200),
                        source=dict(kind="sinusoidal", pos=1, freq=700e6, mode="add"),
                        epsr=4.0, sigma=0.04)
        case_fd2d_3_3(f"fd2d_3_3_{suf}", f"{q}_3_3", "fd2d_3_4", dt, 70, 90, 160, 8, 12, 4.0)
        case_fd2d_3_1(f"fd2d_3_1_{suf}", f"{q}_3_1", dt, 60, 70, 60,
                      probes=[(30, 35), (10, 20), (45, 60)])
        case_fd2d_3_2(f"fd2d_3_2_{suf}", f"{q}_3_2", dt, 64, 80, 100, 8,
                      probes=[(27, 35), (5, 5), (60, 70)])
        case_fd2d_tfsf(f"fd2d_tfsf_{suf}", f"{q}_3_4", "fd2d_3_4", dt, 70, 90, 160, 8, 12,
                       30.0, 0.30, (50e6, 300e6, 700e6))
        case_fd2d_c2(f"fd2d_c2_{suf}", f"{q}_3_4", "fd2d_3_4", dt, 97, 131, 150, 8, 30, 30.0, 0.30,
                     probes=[(48, 75), (48, 100), (70, 65), (10, 65)])
    case_dielectric_c("fd2d_dielectric_c", 97, 131, 8, 30, 30.0, 0.30)
    case_pmlparam_c("fd2d_pmlparam_c", 200, 150, (8, 32, 64))


if __name__ == "__main__":
    main()
