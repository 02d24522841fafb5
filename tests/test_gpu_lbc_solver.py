"""Learned-boundary time stepping (fd2d_lbc_run) against the oracle's
learned step (oracle/lbc_oracle.py: lbc_run).

With a zero network every prediction is the baseline (the two-step chain
ez_d(t+1) = ez_{d+1}(t-1)) in FP32 on both sides, so the whole path --
blocked launch, frame regions and ownership, history ring and its hand-over
between launches -- must match bit for bit.  (Constant outputs cannot be
used for this: the network is odd, its biases cancel.)  A generic network
adds the MFMA's FP32 summation order (and the odd bf16 rounding flip of a
hidden activation): rel L2 <= 1e-4 per field."""
import numpy as np
import pytest
import torch

from oracle import lbc_oracle as L
from conftest import same_bits

pytestmark = pytest.mark.gpu

FIELDS = ("dz", "iz", "ez", "hx", "hy")


def _medium(nx, ny, dev):
    from simulation_amd import fd2d
    dt = fd2d.time_step()
    return fd2d.dielectric(nx, ny, 0, 0, dt, 1.0, 0.0, dev,
                           cylinders=[(nx / 2 + 6.5, ny / 2 - 9.0, 11.0, 3.5, 0.03)])


def _wave(ns):
    from simulation_amd import fd1d
    return fd1d.waveform("gaussian", ns, 0, np.float32, center=12, spread=4.0)


def _oracle_run(nx, ny, md, params, runs, src, wave):
    st = {k: np.zeros((nx, ny), np.float32) for k in FIELDS}
    naz, nbz = md.naz.cpu().numpy(), md.nbz.cpu().numpy()
    hist, t = None, 0
    for n in runs:
        hist = L.lbc_run(st, naz, nbz, params, n, hist, [src], [wave], t0=t)
        t += n
    return st, hist


def _gpu_run(nx, ny, md, params, runs, src, wave, probe, dev, **kw):
    from simulation_amd import learned
    net = learned.Net(params, dev)
    s = learned.LearnedTMSolver(nx, ny, net, md, [(*src, wave)], [probe], nsteps_max=sum(runs),
                                device=dev, **kw)
    for n in runs:
        s.run(n)
    return s


@pytest.mark.parametrize("depth,rows,runs", [(8, 0, (60, 30)), (3, 16, (67, 8)), (8, 24, (2, 82)),
                                             (10, 0, (60, 30)), (9, 20, (61, 30))])
def test_zero_network_bit_exact(gpu, depth, rows, runs):
    nx, ny = 96, 112
    print(f"[lbc solver] zero net depth={depth} rows={rows} runs={runs}", flush=True)
    p = np.zeros(L.NPARAM, np.float32)
    md = _medium(nx, ny, gpu)
    # at least max(32, 3 depth + 6) cells from every edge
    src, probe = ((34, 40), (40, 40)) if depth <= 8 else ((40, 42), (46, 42))
    wave = _wave(sum(runs))
    g = _gpu_run(nx, ny, md, p, runs, src, wave, probe, gpu, depth=depth, rows=rows)
    st, hist = _oracle_run(nx, ny, md, p, runs, src, wave)
    got = g.state()
    for k in FIELDS:
        assert same_bits(got[k], st[k]), k
    h = g.history()
    for a in range(4):
        assert same_bits(h[:, a, :], hist[a]), a
    assert g.probe_traces()[0, -1] == st["ez"][probe]
    assert np.abs(st["ez"][:3]).max() > 1e-3          # the frame saw the pulse
    assert np.abs(hist[0]).max() > 1e-3


def test_generic_network_matches_oracle(gpu):
    nx, ny = 80, 72
    print("[lbc solver] generic net", flush=True)
    p = L.init_params(11)
    p[L.P_W3:L.P_L] *= 3.0
    p[L.P_L:] = (np.random.default_rng(11).standard_normal(L.NPARAM - L.P_L) * 0.05).astype(np.float32)
    md = _medium(nx, ny, gpu)
    src, probe = (34, 36), (40, 36)
    runs = (56, 24)
    wave = _wave(80)
    g = _gpu_run(nx, ny, md, p, runs, src, wave, probe, gpu)
    st, _ = _oracle_run(nx, ny, md, p, runs, src, wave)
    got = g.state()
    assert np.abs(st["ez"][:3]).max() > 1e-3
    for k in ("ez", "hx", "hy"):
        r = np.linalg.norm(got[k].astype(np.float64) - st[k]) / np.linalg.norm(st[k])
        print(f"[lbc solver] {k}: rel L2 {r:.2e}", flush=True)
        assert r < 1e-4, k


def test_lbc_run_rejects_bad_setups(gpu):
    from simulation_amd import _lib, learned
    nx, ny = 96, 96
    md = _medium(nx, ny, gpu)
    net = learned.Net(None, gpu)
    s = learned.LearnedTMSolver(nx, ny, net, md, [(20, 48, _wave(10))], [], nsteps_max=10, device=gpu)
    with pytest.raises(_lib.FdtdError, match="within"):
        s.run(4)
    s = learned.LearnedTMSolver(nx, ny, net, md, [(48, 48, _wave(10))], [], nsteps_max=10, device=gpu)
    with pytest.raises(_lib.FdtdError, match="at least 2"):
        s.run(1)
    # depth 10: the frame's corner regions reach 36 cells deep
    s = learned.LearnedTMSolver(nx, ny, net, md, [(34, 48, _wave(10))], [], nsteps_max=10, depth=10,
                                device=gpu)
    with pytest.raises(_lib.FdtdError, match="within 36"):
        s.run(4)
    s = learned.LearnedTMSolver(nx, ny, net, md, [(48, 48, _wave(10))], [], nsteps_max=10, depth=11,
                                device=gpu)
    with pytest.raises(_lib.FdtdError, match="above 10"):
        s.run(4)


def test_trained_network_gain(gpu):
    """The shipped network (simulation_amd/data/lbc_weights.npz, trained on the
    GPU by python -m simulation_amd.learned) against the method without it --
    the same solver with a zero network, i.e. the two-step boundary it
    corrects -- and against the npml=32 PML, on the reflection case at 512^2
    (reflected energy / peak energy, FP64 accounting).  Measured when the
    network was trained: learned 3.4e-5, zero network 3.7e-3, pml32 1.5e-4."""
    from simulation_amd import learned
    net = learned.load(device=gpu)
    r = learned.reflection(net, 512, methods=("learned", "abc", "pml32"), device=gpu)["ratio"]
    print(f"[lbc gain] " + " ".join(f"{m}={v:.2e}" for m, v in r.items()), flush=True)
    assert r["learned"] < r["abc"] / 20
    assert r["learned"] <= 1e-3
    assert r["learned"] <= r["pml32"]
