"""Configurations at their full size and step count against the oracle:
the whole grid bit for bit against the FP32 oracle (every state array and
every probe trace), and the FP32 drift against the FP64 oracle.

The 4096^2 configuration (2000 steps) and the 8192^2 bench workload (1000
steps, FP32 only) run in every GPU run: the oracle takes about 10 s and 16 s
(4096^2, FP32 and FP64) and 20 s (8192^2) on 16 host threads.  The 16384^2
configuration (5000 steps) needs about 40 times that, so its tests run only
when FDTD_FULL_SIZE=1 is set.  The oracle runs print a line every 250
steps."""
import os
import time

import numpy as np
import pytest
from conftest import same_bits

pytestmark = pytest.mark.gpu

full_only = pytest.mark.skipif(os.environ.get("FDTD_FULL_SIZE") != "1",
                               reason="16384^2 x 5000-step oracle runs take minutes of host "
                                      "time: set FDTD_FULL_SIZE=1")

PML_KEYS = ("fx1", "fx2", "fx3", "fy1", "fy2", "fy3", "gx2", "gx3", "gy2", "gy3")
STATE = ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy")


def _oracle_run(oracle, st, ns, pml, naz, nbz, srcs, waves, probes, label, chunk=250):
    out = []
    t0 = time.time()
    for a in range(0, ns, chunk):
        k = min(chunk, ns - a)
        out.append(oracle.fd2d_tm_run(st, a, k, pml, naz, nbz, srcs, waves[:, a:a + k], probes))
        print(f"[full-size] {label}: {a + k}/{ns} steps, {time.time() - t0:.0f} s", flush=True)
    return np.concatenate(out, axis=1)


def _gpu_run(cfg, gpu):
    from simulation_amd import fd2d
    print(f"[full-size] {cfg.nx}^2, {cfg.ns} steps on the GPU", flush=True)
    s = fd2d.build_solver(cfg, gpu)
    s.run(cfg.ns)
    st = s.state()
    traces = s.probe_traces()
    pml = {k: getattr(s.pml, k).cpu().numpy() for k in PML_KEYS}
    naz, nbz = s.md.naz.cpu().numpy(), s.md.nbz.cpu().numpy()
    return st, traces, pml, naz, nbz


def _check_fp32(oracle, cfg, st, traces, pml, naz, nbz, label):
    """The device-built medium against the oracle's, then the run against the
    FP32 oracle on the same coefficients, bit for bit."""
    from simulation_amd import fd2d
    n, ns = cfg.nx, cfg.ns
    o_naz, o_nbz = oracle.fd2d_dielectric(n, n, cfg.npml, fd2d.time_step(), cfg.cylinders)
    assert same_bits(naz, o_naz.astype(np.float32)) and same_bits(nbz, o_nbz.astype(np.float32))
    del o_naz, o_nbz
    srcs = [(i, j) for i, j, _, _ in cfg.sources]
    w = np.stack([oracle.waveform("gaussian", 0, ns, np.float32, center=t0, spread=sp)
                  for _, _, t0, sp in cfg.sources])
    o = oracle.TMState(n, n, np.float32)
    pr = _oracle_run(oracle, o, ns, pml, naz, nbz, srcs, w, cfg.probes, f"{label} FP32 oracle")
    for k in STATE:
        assert same_bits(st[k], getattr(o, k)), k
    assert same_bits(traces, pr)
    print(f"[full-size] {label}: every state array and probe trace bit-exact against the FP32 "
          "oracle", flush=True)


def _check_fp64(oracle, cfg, st, traces, label):
    """FP32 drift against the FP64 oracle on FP64 coefficients (the
    reference's host set-up before narrowing)."""
    from simulation_amd import fd2d
    n, ns = cfg.nx, cfg.ns
    n64, b64 = oracle.fd2d_dielectric(n, n, cfg.npml, fd2d.time_step(), cfg.cylinders)
    p64 = oracle.fd2d_pmlparam(n, n, cfg.npml)
    srcs = [(i, j) for i, j, _, _ in cfg.sources]
    w64 = np.stack([oracle.waveform("gaussian", 0, ns, np.float64, center=t0, spread=sp)
                    for _, _, t0, sp in cfg.sources])
    o64 = oracle.TMState(n, n, np.float64)
    pr64 = _oracle_run(oracle, o64, ns, p64, n64, b64, srcs, w64, cfg.probes, f"{label} FP64 oracle")
    rel = {k: float(np.linalg.norm(st[k] - getattr(o64, k)) / np.linalg.norm(getattr(o64, k)))
           for k in ("ez", "hx", "hy")}
    relp = float(np.linalg.norm(traces - pr64) / np.linalg.norm(pr64))
    mx = float(np.abs(st["ez"] - o64.ez).max())
    print(f"[full-size] {label} FP32 vs FP64 rel-L2: {rel}, probes {relp:.3e}, "
          f"max |ez| error {mx:.3e}", flush=True)
    assert max(rel.values()) < 1e-4 and relp < 1e-4


def test_c2_full_grid_full_steps(gpu, oracle):
    """4096^2, npml=8, cylinder r=400, hard Gaussian at the centre, four
    probes, 2000 steps."""
    from simulation_amd import fd2d
    cfg = fd2d.c2_config(4096, 2000)
    st, traces, pml, naz, nbz = _gpu_run(cfg, gpu)
    _check_fp32(oracle, cfg, st, traces, pml, naz, nbz, "C2")
    _check_fp64(oracle, cfg, st, traces, "C2")


def test_headline_8192_full_grid(gpu, oracle):
    """The bench's 8192^2 workload (npml=8, cylinder r=400, hard Gaussian at
    the centre, four probes) for 1000 steps: the whole grid bit for bit
    against the FP32 oracle (about 20 s of oracle time)."""
    from simulation_amd import fd2d
    cfg = fd2d.c2_config(8192, 1000)
    st, traces, pml, naz, nbz = _gpu_run(cfg, gpu)
    _check_fp32(oracle, cfg, st, traces, pml, naz, nbz, "8192^2")


@full_only
def test_c5_full_grid_full_steps_fp32(gpu, oracle):
    """16384^2, npml=8, 16 seeded cylinders, 4 sources, 4 probes, 5000 steps,
    bit for bit against the FP32 oracle."""
    from simulation_amd import fd2d
    cfg = fd2d.c5_config(16384, 5000)
    st, traces, pml, naz, nbz = _gpu_run(cfg, gpu)
    _check_fp32(oracle, cfg, st, traces, pml, naz, nbz, "C5")


@full_only
def test_c5_full_grid_full_steps_fp64(gpu, oracle):
    """The same run against the FP64 oracle."""
    from simulation_amd import fd2d
    cfg = fd2d.c5_config(16384, 5000)
    st, traces, _, _, _ = _gpu_run(cfg, gpu)
    _check_fp64(oracle, cfg, st, traces, "C5")


def _c4_case(n=2048, ns=600):
    """2048^2 learned-boundary grid (no PML), a lossy cylinder and a hard
    Gaussian 160 cells below the top edge: the pulse reaches the learned
    frame after about 320 steps and works against it for the rest."""
    cyl = [(400.0, 1200.5, 100.0, 4.0, 0.02)]
    return n, ns, cyl, (160, 1100)


def _c4_runs(gpu, params, n, ns, cyl, src):
    from simulation_amd import fd1d, fd2d, learned
    dt = fd2d.time_step()
    md = fd2d.dielectric(n, n, 0, 0, dt, 1.0, 0.0, gpu, cylinders=cyl)
    w = fd1d.waveform("gaussian", ns, 0, np.float32, center=20, spread=6.0)
    s = learned.LearnedTMSolver(n, n, learned.Net(params, gpu), md, [(*src, w)], [], nsteps_max=ns,
                                device=gpu)
    print(f"[full-size] C4 {n}^2 learned boundary, {ns} steps on the GPU", flush=True)
    s.run(ns)
    return s.state(), md.naz.cpu().numpy(), md.nbz.cpu().numpy(), w


def _c4_oracle(params, naz, nbz, ns, src, w, dtype, label, chunk=100):
    from oracle import lbc_oracle as L
    n = naz.shape[0]
    st = {k: np.zeros((n, n), dtype) for k in ("dz", "iz", "ez", "hx", "hy")}
    hist, t0 = None, time.time()
    for a in range(0, ns, chunk):
        k = min(chunk, ns - a)
        hist = L.lbc_run(st, naz, nbz, params, k, hist, [src], [w], t0=a)
        print(f"[full-size] {label}: {a + k}/{ns} steps, {time.time() - t0:.0f} s", flush=True)
    return st


@full_only
def test_c4_full_grid_zero_network_bit_exact(gpu):
    """With a zero network every prediction is the two-step boundary in FP32
    on both sides: the whole 2048^2 learned-boundary path (blocked launch,
    frame kernel, history ring) bit for bit against the oracle's step."""
    from oracle import lbc_oracle as L
    n, ns, cyl, src = _c4_case()
    params = np.zeros(L.NPARAM, np.float32)
    st, naz, nbz, w = _c4_runs(gpu, params, n, ns, cyl, src)
    o = _c4_oracle(params, naz, nbz, ns, src, w, np.float32, "C4 zero network FP32 oracle")
    for k in ("dz", "iz", "ez", "hx", "hy"):
        assert same_bits(st[k], o[k]), k
    assert np.abs(o["ez"][:8]).max() > 1e-6          # the pulse reached the top frame
    print("[full-size] C4 zero network: every state array bit-exact against the FP32 oracle",
          flush=True)


@full_only
def test_c4_full_grid_shipped_network(gpu, oracle):
    """The shipped network at 2048^2 against the FP32 oracle step and the
    FP64 evaluation of the same scheme (FP64 coefficients, samples,
    prediction and updates)."""
    from simulation_amd import fd2d, learned
    n, ns, cyl, src = _c4_case()
    params = np.load(learned.WEIGHTS, allow_pickle=False)["params"].astype(np.float32)
    st, naz, nbz, w = _c4_runs(gpu, params, n, ns, cyl, src)
    o = _c4_oracle(params, naz, nbz, ns, src, w, np.float32, "C4 shipped network FP32 oracle")
    r32 = {k: float(np.linalg.norm(st[k] - o[k]) / np.linalg.norm(o[k])) for k in ("ez", "hx", "hy")}
    exact = all(same_bits(st[k], o[k]) for k in ("dz", "iz", "ez", "hx", "hy"))
    print(f"[full-size] C4 shipped network vs FP32 oracle: bit-exact={exact}, rel-L2 {r32}", flush=True)
    del o
    n64, b64 = oracle.fd2d_dielectric(n, n, 0, fd2d.time_step(), cyl)
    w64 = oracle.waveform("gaussian", 0, ns, np.float64, center=20, spread=6.0)
    o64 = _c4_oracle(params, n64, b64, ns, src, w64, np.float64, "C4 shipped network FP64 oracle")
    r64 = {k: float(np.linalg.norm(st[k] - o64[k]) / np.linalg.norm(o64[k])) for k in ("ez", "hx", "hy")}
    print(f"[full-size] C4 shipped network FP32 vs FP64 rel-L2 {r64}", flush=True)
    assert max(r32.values()) < 1e-4
    assert max(r64.values()) < 1e-3
