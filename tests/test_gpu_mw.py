"""The multi-wave form of the fused TM launch (interior blocks swept by four
waves that exchange their edge columns through LDS, frame in single-wave
strips) against the single-wave strip form and the oracle.

Bar: bit-exact (np.array_equal) for every state array and probe trace.  The
cases cover depths 1-8 (odd launch counts, short tails), row blocks shorter
and longer than the depth, super-strips cut by the grid edge, grids with no
interior at all, the lossless and free-space forms, and sources and probes
inside the interior, on its edges and in the halo of a neighbouring block.
"""
import numpy as np
import pytest
from conftest import same_bits

pytestmark = pytest.mark.gpu


def _cfg(nx, ny, ns, npml=8, lossy=True, sources=None, probes=None, cylinders=None):
    from simulation_amd import fd2d
    c = nx // 2, ny // 2
    return fd2d.TMConfig(
        nx=nx, ny=ny, ns=ns, npml=npml, lossy=lossy,
        cylinders=cylinders if cylinders is not None else [(nx / 2 - 1, ny / 2 - 1, min(nx, ny) / 5,
                                                            4.0, 0.02)],
        sources=sources if sources is not None else [(c[0], c[1], 20, 6.0)],
        probes=probes if probes is not None else [(c[0], c[1] + 7), (c[0] + 3, c[1]), (12, c[1])])


def _pair(cfg, dev, depth=0, rows=0):
    from simulation_amd import _lib, fd2d
    cfg.depth, cfg.rows = depth, rows
    cfg.form = _lib.TM_FORM_MULTIWAVE
    a = fd2d.build_solver(cfg, dev)
    cfg.rows = 0
    cfg.form = _lib.TM_FORM_STRIP
    b = fd2d.build_solver(cfg, dev)
    return a, b


def _same(a, b):
    sa, sb = a.state(), b.state()
    for k in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy"):
        if sa[k] is None:
            assert sb[k] is None
            continue
        assert same_bits(sa[k], sb[k]), k
    if a.probe_out is not None:
        assert same_bits(a.probe_traces(), b.probe_traces())


CASES = [
    # nx, ny, ns, depth, rows, chunks
    (640, 640, 96, 8, 0, (96,)),
    (640, 640, 96, 8, 20, (40, 56)),
    (640, 640, 90, 5, 7, (90,)),
    (520, 700, 70, 7, 64, (1, 33, 36)),
    (300, 517, 64, 8, 300, (64,)),
    (300, 517, 60, 3, 11, (60,)),
    (257, 1031, 48, 6, 33, (48,)),
    (97, 131, 40, 4, 16, (40,)),      # interior narrower than a super-strip
    (60, 70, 40, 8, 0, (40,)),        # no interior: frame only
    (700, 300, 34, 1, 50, (34,)),
    (700, 300, 34, 2, 50, (34,)),
]


@pytest.mark.parametrize("nx,ny,ns,depth,rows,chunks", CASES)
def test_multiwave_matches_strips(gpu, nx, ny, ns, depth, rows, chunks):
    print(f"[mw] {nx}x{ny} ns={ns} depth={depth} rows={rows} chunks={chunks}", flush=True)
    a, b = _pair(_cfg(nx, ny, ns), gpu, depth, rows)
    for n in chunks:
        a.run(n)
        b.run(n)
    _same(a, b)
    assert np.abs(a.state()["ez"]).max() > 1e-3


@pytest.mark.parametrize("lossy,npml", [(False, 8), (True, -1), (False, -1)])
def test_multiwave_forms(gpu, lossy, npml):
    """Lossless (ez = naz*dz), free space (no profiles) and both."""
    print(f"[mw] forms lossy={lossy} npml={npml}", flush=True)
    a, b = _pair(_cfg(400, 530, 64, npml=npml, lossy=lossy), gpu, 8, 40)
    a.run(64)
    b.run(64)
    _same(a, b)


def test_multiwave_sources_and_probes_everywhere(gpu):
    """Sources and probes inside the interior, on a super-strip boundary, in
    a neighbouring block's halo rows and columns, and in the frame."""
    nx, ny, ns = 600, 760, 80
    # the interior starts at column 8 + 8 + 1 = 17; super-strips write 240
    # columns: boundaries at 257, 497; blocks of 50 rows start at 16
    src = [(300, 257, 20, 6.0), (66, 256, 25, 5.0), (115, 497, 18, 4.0), (5, 30, 20, 6.0),
           (300, 10, 22, 7.0), (590, 700, 20, 6.0)]
    probes = [(300, 258), (66, 250), (116, 497), (65, 496), (4, 4), (66, 17), (599, 759),
              (200, 16), (200, 17)]
    cfg = _cfg(nx, ny, ns, sources=src, probes=probes)
    print("[mw] sources and probes", flush=True)
    a, b = _pair(cfg, gpu, 8, 50)
    a.run(37)
    b.run(37)
    a.run(43)
    b.run(43)
    _same(a, b)


def test_multiwave_matches_oracle(gpu, oracle):
    """C2-shaped run at 256^2 against the C oracle (fd2d/clang/test_3_4.c
    loop order), all state arrays and probes."""
    from simulation_amd import fd2d
    n, ns = 256, 120
    cfg = fd2d.c2_config(n, ns)
    cfg.depth, cfg.rows, cfg.form = 8, 24, 2
    print("[mw] oracle 256^2", flush=True)
    s = fd2d.build_solver(cfg, gpu)
    s.run(57)
    s.run(63)
    st = s.state()
    pml = {k: getattr(s.pml, k).cpu().numpy() for k in fd2d.pmlayer._fields}
    o = oracle.TMState(n, n, np.float32)
    w = oracle.waveform("gaussian", 0, ns, np.float32, center=20, spread=6.0)
    pr = oracle.fd2d_tm_run(o, 0, ns, pml, s.md.naz.cpu().numpy(), s.md.nbz.cpu().numpy(),
                            [(n // 2, n // 2)], w[None, :], cfg.probes)
    for k in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy"):
        assert same_bits(st[k], getattr(o, k)), k
    assert same_bits(s.probe_traces(), pr)
