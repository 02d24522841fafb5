"""Error of the FP32 HIP path against the FP64 oracle, per configuration.

The FP32 path is bit-exact against the reference's FP32 programs (the other
GPU tests); these tests state how far that FP32 arithmetic drifts from the
FP64 evaluation of the same scheme, as relative L2 over the compared window.

Large grids use the dependency cone: one leapfrog step moves information at
most one cell along each axis, so the fields of a window W after S steps
depend only on the initial state (all zero), coefficients and sources inside W
grown by S+2 cells on every side.  The FP64 oracle runs that sub-domain alone
(no PML: it lies clear of the grid's PML band) and its window values are the
FP64 values of the full grid -- no approximation.

Set FDTD_REPORT=<path> to append one JSON record per check to that file.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("ez", "hx", "hy")


def _rel(a, b):
    nb = float(np.linalg.norm(b.astype(np.float64)))
    return float(np.linalg.norm(a.astype(np.float64) - b)) / nb if nb > 0 else 0.0


def _report(config, **rec):
    path = os.environ.get("FDTD_REPORT")
    rec = {"config": config, **rec}
    print(json.dumps(rec), flush=True)
    if path:
        with open(path, "a") as f:
            f.write(json.dumps(rec) + "\n")


def test_fd1d_c1_fp32_vs_fp64(gpu, oracle):
    from simulation_amd import fd1d
    cfg = fd1d.Fd1dConfig()
    got = fd1d.simulate(cfg, gpu)
    ex = np.zeros(cfg.nx)
    hy = np.zeros(cfg.nx)
    bc = np.zeros(4)
    w = oracle.waveform("gaussian", 0, cfg.ns, np.float64, center=cfg.center, spread=cfg.spread)
    pr = oracle.fd1d_run(cfg.nx, 0, cfg.ns, ex, hy, bc, src_mode=1, src_pos=cfg.nx // 2,
                         src_wave=w, probes=cfg.probes)
    # compare at the step where the pulse is inside the grid (residual after
    # 500 steps is ~1e-4 of the peak and dominated by the boundary history)
    r = {"ex": _rel(got["ex"], ex), "hy": _rel(got["hy"], hy), "probe": _rel(got["probe"], pr)}
    _report("C1 fd1d ke=200 500 steps", steps=cfg.ns, rel_l2=r,
            max_abs_probe=float(np.abs(got["probe"] - pr).max()))
    assert r["probe"] < 1e-5
    assert np.abs(got["ex"] - ex).max() < 1e-6 and np.abs(got["hy"] - hy).max() < 1e-6


def _cone_check(oracle, cfg, solver_steps, window_center, label, gpu):
    """Run cfg on the GPU for solver_steps, then compare a 1024^2 window around
    window_center with the FP64 oracle on the window grown by the cone."""
    import torch
    from simulation_amd import fd2d
    S = solver_steps
    wn = min(1024, cfg.nx - 2 * (cfg.npml + S + 8))
    ci, cj = window_center
    wi0, wj0 = ci - wn // 2, cj - wn // 2
    m = S + 4
    oi, oj = wi0 - m, wj0 - m
    sub = wn + 2 * m
    assert oi > cfg.npml + 2 and oj > cfg.npml + 2
    assert oi + sub < cfg.nx - cfg.npml - 2 and oj + sub < cfg.ny - cfg.npml - 2
    s = fd2d.build_solver(cfg, gpu)
    s.run(S)
    torch.cuda.synchronize()
    win = {k: getattr(s, k)[wi0:wi0 + wn, wj0:wj0 + wn].cpu().numpy() for k in FIELDS}
    del s
    torch.cuda.empty_cache()

    dt = fd2d.time_step()
    cyl = [(cx - oi, cy - oj, r, e, sg) for cx, cy, r, e, sg in cfg.cylinders]
    naz, nbz = oracle.fd2d_dielectric(sub, sub, 0, dt, cyl)
    st = oracle.TMState(sub, sub, np.float64)
    srcs, waves = [], []
    for i, j, t0, sp in cfg.sources:
        if oi <= i < oi + sub and oj <= j < oj + sub:
            srcs.append((i - oi, j - oj))
            waves.append(oracle.waveform("gaussian", 0, S, np.float64, center=t0, spread=sp))
    oracle.fd2d_tm_run(st, 0, S, None, naz, nbz, srcs, np.array(waves) if waves else None)
    r = {k: _rel(win[k], getattr(st, k)[m:m + wn, m:m + wn]) for k in FIELDS}
    mx = {k: float(np.abs(win[k] - getattr(st, k)[m:m + wn, m:m + wn]).max()) for k in FIELDS}
    _report(label, steps=S, window=[int(wi0), int(wj0), int(wn)], sub_domain=int(sub),
            sources_in_cone=len(srcs), rel_l2=r, max_abs=mx)
    for k in FIELDS:
        assert r[k] < 1e-4, (k, r[k])


def test_fd2d_c2_window_fp32_vs_fp64(gpu, oracle):
    from simulation_amd import fd2d
    cfg = fd2d.c2_config(4096, 2000)
    c = cfg.nx // 2
    _cone_check(oracle, cfg, 800, (c, c), "C2 fd2d 4096^2 npml=8 cylinder r=400", gpu)


def test_fd2d_c5_window_fp32_vs_fp64(gpu, oracle):
    import torch
    from simulation_amd import fd2d
    free = torch.cuda.mem_get_info()[0]
    cfg = fd2d.c5_config(16384, 5000)
    if free < 20 * 2**30:
        pytest.skip(f"16384^2 needs ~16 GiB of device memory, {free / 2**30:.1f} GiB free")
    i, j = cfg.sources[0][:2]
    _cone_check(oracle, cfg, 500, (i, j), "C5 fd2d 16384^2 16 cylinders 4 sources", gpu)


def test_fd2d_c3_grids_fp32_vs_fp64(gpu, oracle):
    """Two grids of the batched configuration, 3000 steps, full grid in FP64
    (PML included)."""
    from simulation_amd import fd2d
    from simulation_amd.batch import BatchSolver, c3_specs
    specs = c3_specs(0)
    pick = sorted(range(len(specs)), key=lambda k: specs[k].nx)[:2]
    sel = [specs[k] for k in pick]
    ns = 3000
    b = BatchSolver(sel, ns, device=gpu)
    b.run(ns)
    dt = fd2d.time_step()
    for k, sp in enumerate(sel):
        got = b.grid_state(k)
        pml = oracle.fd2d_pmlparam(sp.nx, sp.ny, sp.npml)
        naz, nbz = oracle.fd2d_dielectric(sp.nx, sp.ny, sp.npml, dt, sp.cylinders)
        st = oracle.TMState(sp.nx, sp.ny, np.float64)
        si, sj, t0, spread = sp.source
        w = oracle.waveform("gaussian", 0, ns, np.float64, center=t0, spread=spread)
        pr = oracle.fd2d_tm_run(st, 0, ns, pml, naz, nbz, [(si, sj)], w[None, :], [sp.probe])
        r = {f: _rel(got[f], getattr(st, f)) for f in FIELDS}
        r["probe"] = _rel(b.probe_trace(k), pr[0])
        _report(f"C3 batch grid {pick[k]} ({sp.nx}^2)", steps=ns, rel_l2=r)
        assert r["probe"] < 1e-4


def test_fd2d_c4_learned_fp32_vs_fp64(gpu, oracle):
    """The learned boundary (shipped network) on a 128 x 128 grid with a lossy
    cylinder and a hard Gaussian at the centre, 300 steps (the pulse reaches
    and leaves through the learned frame): the GPU's FP32 run against the FP64
    evaluation of the same scheme (oracle/lbc_oracle.py: lbc_run with FP64
    state, samples and prediction, FP64 coefficients), whole grid."""
    from oracle import lbc_oracle as L
    from simulation_amd import fd1d, fd2d, learned
    n, ns, c = 128, 300, 64
    print(f"[fp64] C4 learned {n}x{n} {ns} steps", flush=True)
    net = learned.load(learned.WEIGHTS, gpu)
    cyl = [(n / 2 + 10.5, n / 2 - 12.0, 14.0, 4.0, 0.02)]
    dt = fd2d.time_step()
    md = fd2d.dielectric(n, n, 0, 0, dt, 1.0, 0.0, gpu, cylinders=cyl)
    w32 = fd1d.waveform("gaussian", ns, 0, np.float32, center=20, spread=6.0)
    s = learned.LearnedTMSolver(n, n, net, md, [(c, c, w32)], [], nsteps_max=ns, device=gpu)
    s.run(ns)
    st = s.state()
    n64, b64 = oracle.fd2d_dielectric(n, n, 0, dt, cyl)
    w64 = oracle.waveform("gaussian", 0, ns, np.float64, center=20, spread=6.0)
    o = {k: np.zeros((n, n), np.float64) for k in ("dz", "iz", "ez", "hx", "hy")}
    L.lbc_run(o, n64, b64, net.numpy(), ns, None, [(c, c)], [w64])
    rel = {f: _rel(st[f], o[f]) for f in FIELDS}
    _report("C4 learned boundary 128^2, shipped network", steps=ns, window="whole grid",
            rel_l2=rel)
    assert np.abs(o["ez"]).max() > 1e-3
    assert all(v < 1e-3 for v in rel.values())
