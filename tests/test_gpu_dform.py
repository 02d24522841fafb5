"""1-D D/E form with lossy and Debye media and the 1-D running DFT on the GPU
against the reference's own programs (fixtures tests/golden/fd1d_debye_f32,
fd1d_lossy_f32 from fd1d/clang/test_2_3.c, test_2_2.c).  Bar: bit-exact."""
import numpy as np
import pytest
import torch

from cases import dform_case
from conftest import same_bits

pytestmark = pytest.mark.gpu

CASES = ("fd1d_debye_f32", "fd1d_lossy_f32", "fd1d_2_1_f32")


def _gauss(t, t0, spread, dt):
    from simulation_amd import fd1d
    return fd1d.gaussian(t, t0, spread, dt)


def _sin(t, ds, freq, dt):
    from simulation_amd import fd1d
    return fd1d.sinusoidal(t, ds, freq, dt)


def _setup(c, dev):
    from simulation_amd import fd1d
    m = c["meta"]
    md = fd1d.dielectric_dform(c["nx"], float(c["dt"]), m["chi"], m["tau"], m["epsr"], m["sigma"],
                               debye=c["debye"], device=dev)
    s = fd1d.DFormSolver(c["nx"], md, c["ns"], freq=c["freq"], dt=float(c["dt"]),
                         src_pos=c["src_pos"], src_wave=c["wave"], probes=c["probes"], device=dev)
    return s, md


def _check(c, s):
    e = c["exp"]
    torch.cuda.synchronize()
    for k in ("dx", "ix", "ex", "hy", "bc") + (("sx",) if c["debye"] else ()):
        assert same_bits(getattr(s, k).cpu().numpy(), e["out_" + k]), k
    if len(c["freq"]):
        for k in ("r_pt", "i_pt", "r_in", "i_in"):
            assert same_bits(getattr(s.ft, k).cpu().numpy(), e["out_ft_" + k]), k


@pytest.mark.parametrize("name", CASES)
def test_dform_coefficients_bit_exact(gpu, name):
    c = dform_case(name, _gauss, _sin)
    _, md = _setup(c, gpu)
    for k in ("nax", "nbx", "ncx", "ndx"):
        if "out_md_" + k in c["exp"]:
            assert same_bits(getattr(md, k).cpu().numpy(), c["exp"]["out_md_" + k]), k


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("split", [False, True])
def test_dform_run_bit_exact(gpu, name, split):
    c = dform_case(name, _gauss, _sin)
    ns = c["ns"]
    chunks = (1, ns // 2 - 1, ns - ns // 2) if split else (ns,)
    print(f"[dform] {name} chunks={chunks}", flush=True)
    s, _ = _setup(c, gpu)
    for n in chunks:
        s.run(n)
    _check(c, s)
    torch.cuda.synchronize()
    assert same_bits(s.probe_out.cpu().numpy(), c["exp"]["out_probe"])


@pytest.mark.parametrize("name", CASES)
def test_dform_step_kernels_bit_exact(gpu, name):
    from simulation_amd import fd1d
    c = dform_case(name, _gauss, _sin)
    print(f"[dform] per-step {name}", flush=True)
    s, md = _setup(c, gpu)
    nx, nf = c["nx"], len(c["freq"])
    for t in range(1, c["ns"] + 1):
        fd1d.dxfield(t, nx, s.dx, s.hy, c["src_pos"], float(c["wave"][t - 1]))
        fd1d.exfield_dform(nx, md, s.dx, s.ix, s.sx, s.ex)
        if nf:
            fd1d.fourier(t, nf, nx, float(c["dt"]), c["freq"], s.ex, s.ft)
        fd1d.hyfield(nx, s.ex, s.hy, s.bc)
    _check(c, s)
    if nf:
        a, p = s.amplitude_phase()
        assert a.shape == (nf, nx) and np.all(np.isfinite(a))


@pytest.mark.parametrize("name", CASES)
def test_dform_run_per_step_path_bit_exact(gpu, name):
    """The per-step kernels of fd1d_dform_run (no work buffer) against the
    reference fixtures."""
    from simulation_amd import fd1d
    c = dform_case(name, _gauss, _sin)
    print(f"[dform] per-step run {name}", flush=True)
    m = c["meta"]
    md = fd1d.dielectric_dform(c["nx"], float(c["dt"]), m["chi"], m["tau"], m["epsr"], m["sigma"],
                               debye=c["debye"], device=gpu)
    s = fd1d.DFormSolver(c["nx"], md, c["ns"], freq=c["freq"], dt=float(c["dt"]),
                         src_pos=c["src_pos"], src_wave=c["wave"], probes=c["probes"], device=gpu,
                         fused=False)
    assert s.work is None
    s.run(c["ns"])
    _check(c, s)
    torch.cuda.synchronize()
    assert same_bits(s.probe_out.cpu().numpy(), c["exp"]["out_probe"])


@pytest.mark.parametrize("nx,debye,nf,chunks", [
    (1029, True, 3, (1, 70, 99)),      # three segments, the last owning 5 cells
    (1600, False, 3, (64, 33)),
    (1100, True, 0, (2, 97)),
    (2048, True, 2, (128,)),
])
def test_fused_matches_per_step(gpu, nx, debye, nf, chunks):
    """The fused run (segments of 512 cells, up to 32 steps per launch, the
    boundary simulated by every segment whose region reaches a grid edge)
    against the per-step kernels: every field, bc, the DFT planes and the
    incident-probe DFT, probes on both edges, the incident probe cell and
    segment boundaries."""
    from simulation_amd import fd1d
    dt = float(np.float32(np.float32(0.01) / np.float32(6e8)))
    freq = (50e6, 200e6, 500e6)[:nf]
    ns = sum(chunks)
    probes = (0, nx - 1, 10, 511, 512, nx // 2, nx - 3)
    print(f"[dform] fused vs per-step nx={nx} debye={debye} nf={nf} chunks={chunks}", flush=True)

    def make(fused):
        md = fd1d.dielectric_dform(nx, dt, 2.0, 0.001e-6, 2.0, 0.01, debye=debye, device=gpu)
        return fd1d.DFormSolver(nx, md, ns, freq=freq, dt=dt, probes=probes, device=gpu,
                                fused=fused)
    a, b = make(True), make(False)
    assert a.work is not None and b.work is None
    for n in chunks:
        a.run(n)
        b.run(n)
    torch.cuda.synchronize()
    for k in ("dx", "ix", "ex", "hy", "bc") + (("sx",) if debye else ()):
        assert torch.equal(getattr(a, k), getattr(b, k)), k
    if nf:
        for k in ("r_pt", "i_pt", "r_in", "i_in"):
            assert torch.equal(getattr(a.ft, k), getattr(b.ft, k)), k
    assert torch.equal(a.probe_out, b.probe_out)
    assert a.ex.abs().max().item() > 0.1
