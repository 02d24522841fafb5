"""HIP 2-D TM path against the reference's golden fixtures and the oracle.

Bar: bit-exact (np.array_equal) for every state array and probe trace in FP32:
the kernels evaluate the reference's expressions operand for operand without
contraction, and the fused/temporally blocked path only reorders which cells
are computed when.
"""
import numpy as np
import pytest
import torch

from cases import FD2D_CASES, PML_KEYS, fd2d_case
from conftest import same_bits

pytestmark = pytest.mark.gpu

F32_CASES = [c for c in FD2D_CASES if c.endswith("f32")]


def _wave(kind, ns, dt, **kw):
    from simulation_amd import fd1d
    return fd1d.waveform(kind, ns, 0, dt, **kw)


def _t(a, dev):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)


def _setup(c, dev):
    from simulation_amd import fd2d
    pml = None
    if c["pml"] is not None:
        pml = fd2d.pmlayer(**{k: _t(c["pml"][k], dev) for k in PML_KEYS})
    md = fd2d.medium(_t(c["naz"], dev), _t(c["nbz"], dev))
    return pml, md


def _check(c, state, probe):
    e = c["exp"]
    for k in ("dz", "ez", "hx", "hy", "iz", "ihx", "ihy"):
        if "out_" + k in e and state.get(k) is not None:
            assert same_bits(state[k], e["out_" + k]), k
    assert same_bits(probe, e["out_probe"])


@pytest.mark.parametrize("name", F32_CASES)
def test_fd2d_step_kernels_bit_exact(gpu, name):
    from simulation_amd import fd2d
    c = fd2d_case(name, _wave)
    print(f"[fd2d step] {name} {c['nx']}x{c['ny']} ns={c['ns']}", flush=True)
    pml, md = _setup(c, gpu)
    s = fd2d.TMSolver(c["nx"], c["ny"], pml, md, [(*c["sources"][0], c["wave"])], c["probes"],
                      nsteps_max=c["ns"], device=gpu)
    for _ in range(c["ns"]):
        s.step_unfused()
    _check(c, s.state(), s.probe_traces())


FUSED = [(1, 1, 0), (2, 2, 0), (3, 1, 7), (5, 2, 16), (8, 2, 0), (8, 1, 32), (4, 2, 200),
         (9, 1, 13), (10, 1, 0), (11, 2, 21), (12, 1, 40)]


@pytest.mark.parametrize("name", F32_CASES)
@pytest.mark.parametrize("depth,cols,rows", FUSED)
def test_fd2d_fused_bit_exact(gpu, name, depth, cols, rows):
    from simulation_amd import fd2d
    c = fd2d_case(name, _wave)
    print(f"[fd2d fused] {name} depth={depth} cols={cols} rows={rows}", flush=True)
    pml, md = _setup(c, gpu)
    s = fd2d.TMSolver(c["nx"], c["ny"], pml, md, [(*c["sources"][0], c["wave"])], c["probes"],
                      nsteps_max=c["ns"], depth=depth, rows=rows, cols=cols, device=gpu)
    # uneven chunks exercise odd launch counts and the unfused tail step
    for n in (1, 2, 13, c["ns"] - 16):
        s.run(n)
    assert s.t == c["ns"]
    _check(c, s.state(), s.probe_traces())


@pytest.mark.parametrize("nx,ny,npml", [(200, 150, 8), (97, 131, 8), (300, 300, 32),
                                        (256, 512, 64), (20, 30, 9)])
def test_pmlparam_device_matches_oracle(gpu, oracle, nx, ny, npml):
    from simulation_amd import fd2d
    pml = fd2d.new_pmlayer(nx, ny, gpu)
    fd2d.pmlparam(nx, ny, npml, pml)
    ref = oracle.fd2d_pmlparam(nx, ny, npml)
    for k in PML_KEYS:
        assert same_bits(getattr(pml, k).cpu().numpy(), ref[k].astype(np.float32)), k


@pytest.mark.parametrize("int_sub", [False, True])
def test_dielectric_device_matches_oracle(gpu, oracle, int_sub):
    from simulation_amd import fd2d
    nx, ny, npml = 203, 177, 8
    dt = fd2d.time_step()
    cyl = [(nx / 2 - 1, ny / 2 - 1, 40, 30.0, 0.3), (40.5, 60.0, 12.5, 4.0, 0.02),
           (150.0, 30.0, 20.0, 7.5, 0.05)]
    if int_sub:
        cyl = cyl[:1]
    md = fd2d.dielectric(nx, ny, npml, 0, dt, 1, 0, gpu, cylinders=cyl, int_sub=int_sub)
    naz, nbz = oracle.fd2d_dielectric(nx, ny, npml, dt, cyl, int_sub=int_sub)
    assert same_bits(md.naz.cpu().numpy(), naz.astype(np.float32))
    assert same_bits(md.nbz.cpu().numpy(), nbz.astype(np.float32))


def test_energy_matches_oracle(gpu, oracle):
    from simulation_amd import fd2d
    g = torch.Generator().manual_seed(7)
    ez, hx, hy = (torch.randn(333, 517, generator=g) for _ in range(3))
    e = fd2d.energy(ez.to(gpu), hx.to(gpu), hy.to(gpu))
    ref = oracle.fd2d_energy(ez.numpy(), hx.numpy(), hy.numpy())
    assert abs(e - ref) <= 1e-12 * ref
    w = (10, 300, 5, 400)
    e = fd2d.energy(ez.to(gpu), hx.to(gpu), hy.to(gpu), window=w)
    sub = [a.numpy()[10:300, 5:400].copy() for a in (ez, hx, hy)]
    assert abs(e - oracle.fd2d_energy(*sub)) <= 1e-12 * e


@pytest.mark.parametrize("n,ns,depth", [(1024, 300, 8), (640, 257, 6)])
def test_c2_shape_fused_matches_oracle(gpu, oracle, n, ns, depth):
    """The n x n configuration-2 shape (PML + cylinder + hard Gaussian + 4
    probes) through TMSolver, against the C oracle in FP32, bit-exact; and the
    FP32 error against the FP64 oracle is recorded."""
    from simulation_amd import fd2d
    cfg = fd2d.c2_config(n, ns)
    cfg.depth = depth
    print(f"[fd2d c2-shape] n={n} ns={ns} depth={depth}", flush=True)
    s = fd2d.build_solver(cfg, gpu)
    s.run(ns)
    st = s.state()
    pml = {k: getattr(s.pml, k).cpu().numpy() for k in PML_KEYS}
    naz, nbz = s.md.naz.cpu().numpy(), s.md.nbz.cpu().numpy()
    w = oracle.waveform("gaussian", 0, ns, np.float32, center=20, spread=6.0)
    o = oracle.TMState(n, n, np.float32)
    pr = oracle.fd2d_tm_run(o, 0, ns, pml, naz, nbz, [cfg.sources[0][:2]], w[None, :], cfg.probes)
    for k in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy"):
        assert same_bits(st[k], getattr(o, k)), k
    assert same_bits(s.probe_traces(), pr)
    # FP64 oracle on FP64 coefficients (the PyCUDA set-up before narrowing)
    p64 = oracle.fd2d_pmlparam(n, n, cfg.npml)
    n64, b64 = oracle.fd2d_dielectric(n, n, cfg.npml, fd2d.time_step(), cfg.cylinders)
    o64 = oracle.TMState(n, n, np.float64)
    w64 = oracle.waveform("gaussian", 0, ns, np.float64, center=20, spread=6.0)
    pr64 = oracle.fd2d_tm_run(o64, 0, ns, p64, n64, b64, [cfg.sources[0][:2]], w64[None, :],
                              cfg.probes)
    rel = np.linalg.norm(st["ez"] - o64.ez) / np.linalg.norm(o64.ez)
    relp = np.linalg.norm(s.probe_traces() - pr64) / np.linalg.norm(pr64)
    print(f"[fd2d c2-shape] n={n}: rel-L2 vs FP64 ez={rel:.3e} probes={relp:.3e}")
    assert rel < 1e-4 and relp < 1e-4


def test_fused_is_deterministic_and_matches_unfused(gpu):
    """Two fused runs agree bit for bit, and with the per-step kernels."""
    from simulation_amd import fd2d
    cfg = fd2d.c2_config(512, 120)
    outs = []
    for depth in (8, 8, 3):
        cfg.depth = depth
        s = fd2d.build_solver(cfg, gpu)
        s.run(120)
        outs.append(s.state())
    s = fd2d.build_solver(cfg, gpu)
    for _ in range(120):
        s.step_unfused()
    outs.append(s.state())
    for o in outs[1:]:
        for k in ("dz", "ez", "hx", "hy", "iz", "ihx", "ihy"):
            assert same_bits(o[k], outs[0][k]), k


@pytest.mark.parametrize("depth,rows", [(8, 9), (8, 10), (5, 13), (5, 14)])
def test_interior_row_blocks_odd_and_even(gpu, depth, rows):
    """Interior tiles sweep rows + 2*depth positions, two per loop trip with
    alternating carry sets, plus one trailing position when the count is
    odd: both parities against the per-step kernels, bit for bit."""
    from simulation_amd import fd2d
    n, ns = 320, 80
    cfg = fd2d.TMConfig(nx=n, ny=n + 24, ns=ns, npml=8,
                        cylinders=[(150.0, 170.0, 50.0, 4.0, 0.02)],
                        sources=[(236, 250, 20, 6.0)], probes=[(100, 120)], depth=depth,
                        rows=rows, cols=1)
    print(f"[fd2d interior parity] depth={depth} rows={rows}", flush=True)
    a = fd2d.build_solver(cfg, gpu)
    a.run(ns)
    b = fd2d.build_solver(cfg, gpu)
    for _ in range(ns):
        b.step_unfused()
    sa, sb = a.state(), b.state()
    for k in ("dz", "ez", "hx", "hy", "iz", "ihx", "ihy"):
        assert same_bits(sa[k], sb[k]), k
    assert same_bits(a.probe_traces(), b.probe_traces())
    assert np.abs(sa["ez"]).max() > 0.01


@pytest.mark.parametrize("rows", [4, 32])
def test_many_sources_and_probes_tile_lists(gpu, rows):
    """16 hard sources and 16 probes on strip overlaps (each in two strips):
    at rows 32 the boundary tiles fit the launch's lists (near their
    capacity); at rows 4 the source tiles overflow them and every workgroup
    takes its tile from the whole grid.  Both against the per-step kernels,
    bit for bit, state and probe traces."""
    from simulation_amd import fd2d
    n, ns = 400, 90
    srcs = [(40 + 20 * k, 48 * (1 + k % 7) + 3, 10 + k, 3.0 + 0.25 * k) for k in range(16)]
    probes = [(45 + 20 * k, 48 * (1 + (k + 3) % 7) - 5) for k in range(16)]
    cfg = fd2d.TMConfig(nx=n, ny=n, ns=ns, npml=8, cylinders=[(200.0, 210.0, 40.0, 4.0, 0.02)],
                        sources=srcs, probes=probes, depth=8, rows=rows)
    print(f"[fd2d lists] rows={rows}", flush=True)
    a = fd2d.build_solver(cfg, gpu)
    a.run(ns)
    b = fd2d.build_solver(cfg, gpu)
    for _ in range(ns):
        b.step_unfused()
    sa, sb = a.state(), b.state()
    for k in ("dz", "ez", "hx", "hy", "iz", "ihx", "ihy"):
        assert same_bits(sa[k], sb[k]), k
    assert same_bits(a.probe_traces(), b.probe_traces())
    assert np.abs(sa["ez"]).max() > 0.1


def test_free_space_tiles_same_bits_and_no_stale_flags(gpu):
    """Runs that skip naz/nbz in free-space tiles (the tile workspace) give
    the bits of runs that read them; a medium changed between two run calls
    is seen by the second call (its flags are recorded afresh)."""
    from simulation_amd import fd2d
    n = 384
    cfg = fd2d.c2_config(n, 200)
    cfg.depth, cfg.rows = 6, 16
    print("[fd2d] free-space tiles", flush=True)
    a = fd2d.build_solver(cfg, gpu)
    b = fd2d.build_solver(cfg, gpu)
    b.tile_flags = False
    for s in (a, b):
        s.run(80)
    assert a._tiles and not b._tiles
    sa, sb = a.state(), b.state()
    for k in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy"):
        assert same_bits(sa[k], sb[k]), k
    # a lossy patch appears in what was free space (the cylinder has radius
    # 38 around (191, 191); the pulse reaches the patch within the 80 steps)
    for s in (a, b):
        s.md.naz[180:200, 232:252] = 0.5
        s.md.nbz[180:200, 232:252] = 0.01
        s.run(80)
    sa, sb = a.state(), b.state()
    for k in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy"):
        assert same_bits(sa[k], sb[k]), k
    assert same_bits(a.probe_traces(), b.probe_traces())
    assert np.abs(sa["ez"][180:200, 232:252]).max() > 1e-6


def test_free_space_flags_kept_across_calls(gpu):
    """With static_medium the flags recorded by one call serve the next
    calls (same arrays and tiles, depth no deeper); a deeper call records
    afresh.  Same bits as a run that reads every naz/nbz."""
    from simulation_amd import fd2d
    n = 320
    cfg = fd2d.c2_config(n, 300)
    cfg.depth, cfg.rows = 6, 20
    print("[fd2d] free-space flags across calls", flush=True)
    a = fd2d.build_solver(cfg, gpu)
    a.static_medium = True
    b = fd2d.build_solver(cfg, gpu)
    b.tile_flags = False
    for chunk in (25, 60, 12, 31, 2, 70):     # launch depths 5, 6, 6, 6 or 5, 1, 6
        a.run(chunk)
        b.run(chunk)
    sa, sb = a.state(), b.state()
    for k in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy"):
        assert same_bits(sa[k], sb[k]), k
    assert same_bits(a.probe_traces(), b.probe_traces())


def test_free_space_flags_new_medium_in_freed_memory(gpu):
    """A solver built after another was freed (its arrays, and so perhaps the
    same memory, now hold another medium) records its own flags before it
    keeps any: same bits as a run that reads every naz/nbz.  Calls of other
    lengths in between (a warm-up of 5 steps) keep their own flags."""
    import gc
    import torch
    from simulation_amd import fd2d
    n = 288
    cfg = fd2d.c2_config(n, 200)
    cfg.depth, cfg.rows = 6, 12
    print("[fd2d] free-space flags, new medium in freed memory", flush=True)
    a = fd2d.build_solver(cfg, gpu)
    a.static_medium = True
    a.run(36)
    a.run(36)
    torch.cuda.synchronize()
    del a
    gc.collect()
    cfg2 = fd2d.c2_config(n, 200)
    cfg2.depth, cfg2.rows = 6, 12
    cfg2.cylinders = [(110.0, 180.0, 20, 3.0, 0.03), (180.0, 100.0, 15, 2.0, 0.01)]
    b = fd2d.build_solver(cfg2, gpu)
    b.static_medium = True
    c = fd2d.build_solver(cfg2, gpu)
    c.tile_flags = False
    for chunk in (36, 5, 36, 24):
        b.run(chunk)
        c.run(chunk)
    assert len(b._tiles) == 2
    sb, sc = b.state(), c.state()
    for k in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy"):
        assert same_bits(sb[k], sc[k]), k
    assert same_bits(b.probe_traces(), c.probe_traces())
    assert np.abs(sb["ez"][95:125, 165:195]).max() > 1e-6


def test_free_space_flags_after_foreign_state(gpu, tmp_path):
    """A state loaded from another medium's run (nonzero iz where this medium
    is free space) into a solver that keeps its free-space flags across
    calls: the load drops the flags, so the next run records afresh and
    gives the bits of a solver without flags."""
    from simulation_amd import fd2d, store
    n = 320
    cfg = fd2d.c2_config(n, 400)
    cfg.depth, cfg.rows = 6, 20
    print("[fd2d] free-space flags after a foreign state", flush=True)
    a = fd2d.build_solver(cfg, gpu)
    a.static_medium = True
    a.run(60)                                     # records; flags kept
    other = fd2d.c2_config(n, 400)
    other.depth, other.rows = 6, 20
    other.cylinders = [(100.0, 220.0, 30, 3.0, 0.04)]
    other.sources = [(110, 200, 20, 6.0)]
    c = fd2d.build_solver(other, gpu)
    c.run(80)
    assert np.abs(c.state()["iz"][80:120, 200:240]).max() > 0   # iz where a's medium is free
    path = str(tmp_path / "foreign.npz")
    store.save_checkpoint(path, c)
    b = fd2d.build_solver(cfg, gpu)
    b.tile_flags = False
    for s in (a, b):
        store.load_checkpoint(path, s)
        s.run(60)
    sa, sb = a.state(), b.state()
    for k in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy"):
        assert same_bits(sa[k], sb[k]), k
