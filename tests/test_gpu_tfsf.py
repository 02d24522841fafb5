"""Plane wave (TF/SF) and running DFT on the GPU against the reference's own
plane-wave program (fixture tests/golden/fd2d_tfsf_f32.npz, written from
fd2d/clang/test_3_4.c's functions in its main-loop order).

Bar: bit-exact (np.array_equal) for every state array, the incident grid and
its boundary state, and every DFT plane.  The DFT twiddles are evaluated on
the host exactly as the C program evaluates them, so the device sums see
the same FP32 operands."""
import numpy as np
import pytest
import torch

from cases import PML_KEYS, tfsf_case
from conftest import same_bits

pytestmark = pytest.mark.gpu


def _gauss(t, t0, spread, dt):
    from simulation_amd import fd1d
    return fd1d.gaussian(t, t0, spread, dt)


def _t(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)


def _solver(c, dev, fused=True):
    from simulation_amd import fd2d
    pml = fd2d.pmlayer(**{k: _t(c["pml"][k], dev) for k in PML_KEYS})
    md = fd2d.medium(_t(c["naz"], dev), None if c["nbz"] is None else _t(c["nbz"], dev))
    s = fd2d.PlaneWaveSolver(c["nx"], c["ny"], c["npml"], pml, md, c["ns"], freq=c["freq"],
                             dt=float(c["dt"]), inc_pos=c["inc_pos"], inc_wave=c["wave"],
                             device=dev, fused=fused)
    return s, pml, md


def _check(c, s):
    e = c["exp"]
    st = s.state()
    for k in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy"):
        if c["nbz"] is None and k == "iz":
            continue
        assert same_bits(st[k], e["out_" + k]), k
    for k in ("ezi", "hxi", "bc"):
        assert same_bits(getattr(s, k).cpu().numpy(), e["out_" + k]), k
    if len(c["freq"]):
        for k in ("r_pt", "i_pt", "r_in", "i_in"):
            assert same_bits(getattr(s.ft, k).cpu().numpy(), e["out_ft_" + k]), k


@pytest.mark.parametrize("fused,chunks", [(True, (160,)), (True, (1, 63, 96)), (False, (160,))])
def test_lossless_plane_wave_bit_exact(gpu, fused, chunks):
    """The lossless plane-wave program (fd2d/clang/test_3_3.c: efield ez =
    naz*dz, no iz, no DFT) through fd2d_tfsf_run, fused and per-step, against
    the program's own output bit for bit."""
    c = tfsf_case("fd2d_3_3_f32", _gauss)
    print(f"[tfsf] lossless {c['nx']}x{c['ny']} fused={fused} chunks={chunks}", flush=True)
    s, _, _ = _solver(c, gpu, fused)
    assert (s.work is not None) == fused and s.ft is None and s.tm.iz is None
    for n in chunks:
        s.run(n)
    _check(c, s)
    assert np.abs(c["exp"]["out_ez"]).max() > 0.1


@pytest.mark.parametrize("chunks", [(160,), (1, 63, 96)])
def test_plane_wave_run_bit_exact(gpu, chunks):
    c = tfsf_case("fd2d_tfsf_f32", _gauss)
    print(f"[tfsf] run {c['nx']}x{c['ny']} npml={c['npml']} chunks={chunks}", flush=True)
    s, _, _ = _solver(c, gpu)
    for n in chunks:
        s.run(n)
    _check(c, s)


@pytest.mark.parametrize("nf,depth,cols", [(3, 0, 0), (0, 0, 0), (2, 3, 2), (1, 5, 1)])
def test_fused_matches_per_step(gpu, nf, depth, cols):
    """The fused path (incident chunks + temporally blocked launches carrying
    the TF/SF corrections and the DFT) against the per-step kernels, on a
    grid with many strips and row blocks, a lossy cylinder, probes, and runs
    split into chunks of 1 (per-step inside), 70 (two incident chunks) and
    99 steps."""
    from simulation_amd import fd2d
    nx, ny, npml, ns = 300, 262, 12, 170
    freq = (50e6, 300e6, 700e6)[:nf]
    print(f"[tfsf] fused vs per-step nf={nf} depth={depth} cols={cols}", flush=True)

    def make(fused):
        pml = fd2d.new_pmlayer(nx, ny, gpu)
        fd2d.pmlparam(nx, ny, npml, pml)
        md = fd2d.dielectric(nx, ny, npml, 0, fd2d.time_step(), 1.0, 0.0, gpu,
                             cylinders=[(140.0, 120.0, 30.0, 30.0, 0.3)])
        return fd2d.PlaneWaveSolver(nx, ny, npml, pml, md, ns, freq=freq, probes=[(150, 131), (20, 30)],
                                    device=gpu, fused=fused, depth=depth, cols=cols, rows=40)
    a, b = make(True), make(False)
    assert a.work is not None and b.work is None
    for n in (1, 70, 99):
        a.run(n)
        b.run(n)
    sa, sb = a.state(), b.state()
    for k in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy"):
        assert same_bits(sa[k], sb[k]), k
    for k in ("ezi", "hxi", "bc"):
        assert torch.equal(getattr(a, k), getattr(b, k)), k
    assert torch.equal(a.tm.probe_out, b.tm.probe_out)
    if nf:
        for k in ("r_pt", "i_pt", "r_in", "i_in"):
            assert torch.equal(getattr(a.ft, k), getattr(b.ft, k)), k
        assert a.ft.r_pt.abs().max().item() > 0.1
    assert np.abs(sa["ez"]).max() > 0.1


def test_fused_wide_grid(gpu):
    """ny > 8192: the incident chunk runs from global memory (two ny-float
    arrays no longer fit the workgroup's LDS); still the per-step bits."""
    from simulation_amd import fd2d
    nx, ny, npml, ns = 40, 8250, 8, 90
    print(f"[tfsf] fused wide grid {nx}x{ny}", flush=True)

    def make(fused):
        pml = fd2d.new_pmlayer(nx, ny, gpu)
        fd2d.pmlparam(nx, ny, npml, pml)
        md = fd2d.dielectric(nx, ny, npml, 0, fd2d.time_step(), 1.0, 0.0, gpu,
                             cylinders=[(20.0, 60.0, 6.0, 4.0, 0.1)])
        return fd2d.PlaneWaveSolver(nx, ny, npml, pml, md, ns, freq=(300e6,), device=gpu,
                                    fused=fused)
    a, b = make(True), make(False)
    for n in (70, 20):
        a.run(n)
        b.run(n)
    sa, sb = a.state(), b.state()
    for k in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy"):
        assert same_bits(sa[k], sb[k]), k
    for k in ("ezi", "hxi", "bc"):
        assert torch.equal(getattr(a, k), getattr(b, k)), k
    for k in ("r_pt", "i_pt", "r_in", "i_in"):
        assert torch.equal(getattr(a.ft, k), getattr(b.ft, k)), k
    assert np.abs(sa["ez"]).max() > 0.1


def test_plane_wave_step_kernels_bit_exact(gpu):
    """The per-step drop-ins, called in the reference's order."""
    from simulation_amd import fd2d
    c = tfsf_case("fd2d_tfsf_f32", _gauss)
    print("[tfsf] per-step kernels", flush=True)
    s, pml, md = _solver(c, gpu)
    tm, nx, ny, npml = s.tm, c["nx"], c["ny"], c["npml"]
    for t in range(1, c["ns"] + 1):
        fd2d.ezinct(ny, s.ezi, s.hxi, s.bc)
        fd2d.dfield_inc(t, nx, ny, pml, s.ezi, tm.dz, tm.hx, tm.hy, c["inc_pos"],
                        float(c["wave"][t - 1]))
        fd2d.inctdz(nx, ny, npml, s.hxi, tm.dz)
        fd2d.efield(nx, ny, md, tm.dz, tm.iz, tm.ez)
        fd2d.fourier(t, len(c["freq"]), nx, ny, float(c["dt"]), c["freq"], s.ezi, tm.ez, s.ft)
        fd2d.hxinct(ny, s.ezi, s.hxi)
        fd2d.hfield(nx, ny, pml, tm.ez, tm.ihx, tm.ihy, tm.hx, tm.hy)
        fd2d.incthx(nx, ny, npml, s.ezi, tm.hx)
        fd2d.incthy(nx, ny, npml, s.ezi, tm.hy)
    _check(c, s)


def test_amplitude_phase_postprocessing(gpu):
    """The reference's amplitude/phase along row nx/2-1 (host NumPy, as the
    PyCUDA script): finite on the box columns, zero outside, amplitude of
    the incident transform itself is 1 where the field equals it."""
    from simulation_amd import fd2d
    c = tfsf_case("fd2d_tfsf_f32", _gauss)
    s, _, _ = _solver(c, gpu)
    s.run(c["ns"])
    a, p = fd2d.amplitude_phase(s.ft, c["nx"], c["ny"], c["npml"])
    e = c["exp"]
    j = slice(c["npml"] - 1, c["ny"] - c["npml"] + 1)
    r_in, i_in = e["out_ft_r_in"], e["out_ft_i_in"]
    rp, ip = e["out_ft_r_pt"][:, c["nx"] // 2 - 1, j], e["out_ft_i_pt"][:, c["nx"] // 2 - 1, j]
    want = (1 / np.hypot(r_in, i_in))[:, None] * np.hypot(rp, ip)
    assert np.allclose(a[:, j], want, rtol=1e-6)
    assert np.all(np.isfinite(p[:, j])) and not np.any(a[:, : c["npml"] - 1])


def test_plane_wave_rejects_bad_setups(gpu):
    from simulation_amd import _lib
    c = tfsf_case("fd2d_tfsf_f32", _gauss)
    s, _, _ = _solver(c, gpu)
    with pytest.raises(_lib.FdtdError, match="inc_wave"):
        s.run(c["ns"] + 1)
    s.npml = 1
    with pytest.raises(_lib.FdtdError, match="npml"):
        s.run(2)
