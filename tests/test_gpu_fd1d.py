"""HIP 1-D path against the reference's golden fixtures: bit-exact, FP32 and FP64.

Both the per-step kernels (fd1d_exfield / fd1d_hyfield, the PyCUDA launch
pattern of fd1d/pycuda/fd1d_1_2.py:87-89) and the single-launch fd1d_run are
checked against what the reference's C programs wrote.
"""
import numpy as np
import pytest
import torch

from cases import FD1D_CASES, fd1d_case
from conftest import same_bits

pytestmark = pytest.mark.gpu


def _wave(kind, ns, dt, **kw):
    from simulation_amd import fd1d
    return fd1d.waveform(kind, ns, 0, dt, **kw)


def _dev(a, dev, dt):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)


@pytest.mark.parametrize("name", FD1D_CASES)
def test_fd1d_step_kernels_bit_exact(gpu, name):
    from simulation_amd import fd1d
    c = fd1d_case(name, _wave)
    dt, nx = c["dtype"], c["nx"]
    td = torch.float32 if dt == np.float32 else torch.float64
    print(f"[fd1d step] {name} nx={nx} ns={c['ns']}", flush=True)
    ex = torch.zeros(nx, dtype=td, device=gpu)
    hy = torch.zeros(nx, dtype=td, device=gpu)
    bc = torch.zeros(4, dtype=td, device=gpu) if c["abc"] else None
    ca, cb = _dev(c["ca"], gpu, dt), _dev(c["cb"], gpu, dt)
    probe = np.zeros((len(c["probes"]), c["ns"]), dt)
    for s in range(c["ns"]):
        fd1d.exfield(s + 1, nx, ca, cb, ex, hy, c["src_mode"], c["src_pos"], float(c["wave"][s]))
        fd1d.hyfield(nx, ex, hy, bc)
        if c["probes"]:
            probe[:, s] = ex[list(c["probes"])].cpu().numpy()
    e = c["exp"]
    assert same_bits(ex.cpu().numpy(), e["out_ex"])
    assert same_bits(hy.cpu().numpy(), e["out_hy"])
    assert same_bits(probe, e["out_probe"])
    if c["abc"]:
        assert same_bits(bc.cpu().numpy(), e["out_bc"])


@pytest.mark.parametrize("name", FD1D_CASES)
def test_fd1d_single_launch_bit_exact(gpu, name):
    from simulation_amd import fd1d
    c = fd1d_case(name, _wave)
    dt, nx = c["dtype"], c["nx"]
    td = torch.float32 if dt == np.float32 else torch.float64
    print(f"[fd1d run] {name} nx={nx} ns={c['ns']}", flush=True)
    ex = torch.zeros(nx, dtype=td, device=gpu)
    hy = torch.zeros(nx, dtype=td, device=gpu)
    bc = torch.zeros(4, dtype=td, device=gpu) if c["abc"] else None
    ca, cb = _dev(c["ca"], gpu, dt), _dev(c["cb"], gpu, dt)
    wave = _dev(c["wave"], gpu, dt)
    cuts = c["snap_t"] or [c["ns"]]
    t, probes, sx, sh = 0, [], [], []
    for cut in cuts:
        probes.append(fd1d.run(nx, cut - t, ex, hy, bc, ca, cb, c["src_mode"], c["src_pos"],
                               wave[t:cut].contiguous(), c["probes"], t0=t))
        sx.append(ex.cpu().numpy())
        sh.append(hy.cpu().numpy())
        t = cut
    e = c["exp"]
    assert same_bits(ex.cpu().numpy(), e["out_ex"])
    assert same_bits(hy.cpu().numpy(), e["out_hy"])
    assert same_bits(torch.cat(probes, 1).cpu().numpy(), e["out_probe"])
    if c["abc"]:
        assert same_bits(bc.cpu().numpy(), e["out_bc"])
    if c["snap_t"]:
        assert same_bits(np.array(sx), e["out_snap_ex"])
        assert same_bits(np.array(sh), e["out_snap_hy"])


def test_fd1d_c1_simulate_matches_oracle_fp32_and_fp64(gpu, oracle):
    """The ke=200 configuration end to end (simulate), both precisions, against
    the oracle; plus the FP32 error against FP64."""
    from simulation_amd import fd1d
    res = {}
    for dt in (np.float32, np.float64):
        cfg = fd1d.Fd1dConfig(dtype=dt)
        out = fd1d.simulate(cfg, gpu)
        ex = np.zeros(200, dt); hy = np.zeros(200, dt); bc = np.zeros(4, dt)
        w = oracle.waveform("gaussian", 0, 500, dt, center=40, spread=12.0)
        pr = oracle.fd1d_run(200, 0, 500, ex, hy, bc, None, None, 1, 100, w, (150,))
        assert same_bits(out["ex"], ex) and same_bits(out["hy"], hy)
        assert same_bits(out["probe"], pr)
        res[dt] = out
    p32, p64 = res[np.float32]["probe"][0], res[np.float64]["probe"][0]
    rel = np.linalg.norm(p32 - p64) / np.linalg.norm(p64)
    print(f"[fd1d c1] probe rel-L2 FP32 vs FP64 = {rel:.3e}")
    assert rel < 1e-5


def test_fd1d_dielectric_device_matches_oracle(gpu, oracle):
    from simulation_amd import fd1d
    for dt in (np.float32, np.float64):
        ca, cb = fd1d.dielectric(512, 0.01 / 6e8, 4.0, 0.04, gpu, dt)
        rca, rcb = oracle.fd1d_dielectric(512, 0.01 / 6e8, 4.0, 0.04)
        assert same_bits(ca.cpu().numpy(), rca.astype(dt))
        assert same_bits(cb.cpu().numpy(), rcb.astype(dt))


def test_fd1d_hard_source_mirror_symmetry(gpu):
    """Before the pulse reaches a wall, ex is exactly symmetric about the hard
    source cell and hy antisymmetric about kc - 1/2 (a - b == -(b - a) in IEEE)."""
    from simulation_amd import fd1d
    nx, kc = 401, 200
    ex = torch.zeros(nx, device=gpu); hy = torch.zeros(nx, device=gpu)
    bc = torch.zeros(4, device=gpu)
    w = torch.from_numpy(fd1d.waveform("gaussian", 300, 0, np.float32, center=40, spread=12.0)).to(gpu)
    fd1d.run(nx, 300, ex, hy, bc, None, None, fd1d.SRC_HARD, kc, w)
    e, h = ex.cpu().numpy(), hy.cpu().numpy()
    m = np.arange(1, 190)
    assert same_bits(e[kc - m], e[kc + m])
    assert same_bits(h[kc + m], -h[kc - m - 1])
