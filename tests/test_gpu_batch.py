"""Batched uneven grids: every grid of a batch evolves exactly (bit for bit)
as it does alone through fd2d_tm_run, which is itself bit-exact against the
oracle; the captured HIP graph replays the same launches."""
import ctypes as C

import numpy as np
import pytest
import torch
from conftest import same_bits

pytestmark = pytest.mark.gpu

FIELDS = ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy")


def _specs():
    from simulation_amd.batch import GridSpec
    return [
        GridSpec(97, 131, 8, [(47.0, 64.0, 30.0, 30.0, 0.3)], (48, 65, 20, 6.0), (60, 70)),
        GridSpec(200, 200, 8, [(60.5, 80.0, 20.0, 4.0, 0.02), (130.0, 120.0, 35.0, 7.0, 0.05)],
                 (100, 100, 27, 8.0), (110, 100)),
        GridSpec(150, 256, 8, [(75.0, 128.0, 40.0, 2.0, 0.0)], (70, 30, 13, 4.0), None),
        GridSpec(333, 141, 8, [], (20, 70, 40, 12.0), (300, 70)),
        GridSpec(64, 64, 8, [(31.0, 31.0, 10.0, 8.0, 0.04)], None, (32, 32)),
    ]


@pytest.mark.parametrize("depth,rows,cols", [(4, 16, 1), (8, 32, 2), (3, 7, 2), (8, 0, 1)])
def test_batch_matches_single_grid_runs(gpu, depth, rows, cols):
    from simulation_amd.batch import BatchSolver, single_solver
    specs = _specs()
    ns = 61 if depth != 8 else 64
    print(f"[batch] depth={depth} rows={rows} cols={cols} ns={ns}", flush=True)
    b = BatchSolver(specs, ns, depth=depth, rows=rows, cols=cols, device=gpu)
    b.run(ns - 16)
    b.run(16)
    for k, sp in enumerate(specs):
        s = single_solver(sp, ns, gpu)
        s.run(ns)
        ref = s.state()
        got = b.grid_state(k)
        for f in FIELDS:
            assert same_bits(got[f], ref[f]), (k, f)
        if sp.probe:
            assert same_bits(b.probe_trace(k), s.probe_traces()[0]), k


def test_batch_grid_matches_oracle(gpu, oracle):
    from simulation_amd import fd2d
    from simulation_amd.batch import BatchSolver
    specs = _specs()[:2]
    ns = 48
    b = BatchSolver(specs, ns, depth=6, rows=24, cols=1, device=gpu)
    b.run(ns)
    for k, sp in enumerate(specs):
        st = b.state[k]
        pml = {f: getattr(st["pml"], f).cpu().numpy() for f in fd2d.pmlayer._fields}
        o = oracle.TMState(sp.nx, sp.ny, np.float32)
        si, sj, t0, spread = sp.source
        w = oracle.waveform("gaussian", 0, ns, np.float32, center=t0, spread=spread)
        pr = oracle.fd2d_tm_run(o, 0, ns, pml, st["md"].naz.cpu().numpy(), st["md"].nbz.cpu().numpy(),
                                [(si, sj)], w[None, :], [sp.probe])
        got = b.grid_state(k)
        for f in FIELDS:
            assert same_bits(got[f], getattr(o, f)), (k, f)
        assert same_bits(b.probe_trace(k), pr[0])


def test_batch_graph_replays_match_eager(gpu):
    """One captured graph (t0=0, 20 steps) replayed three times, with eager
    device work and synchronisation between replays, against three eager
    runs of the same launches."""
    from simulation_amd import _lib
    from simulation_amd.batch import BatchSolver
    specs = _specs()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        ge = BatchSolver(specs, 20, depth=5, rows=16, cols=1, device=gpu)
        gg = BatchSolver(specs, 20, depth=5, rows=16, cols=1, device=gpu)
        torch.cuda.synchronize()
        h = gg.graph(20)
        for rep in range(3):
            _lib.call("fd2d_batch_run", C.byref(ge.desc), 0, 20, ge.wave_len, ge.probe_len,
                      s.cuda_stream)
            _lib.call("fdtd_graph_launch", h, s.cuda_stream)
            chk = sum(float(gg.state[k]["hx"].abs().sum()) for k in range(len(specs)))  # eager work
            torch.cuda.synchronize()
            assert chk > 0
            for k in range(len(specs)):
                a, b = ge.grid_state(k), gg.grid_state(k)
                for f in FIELDS:
                    assert same_bits(a[f], b[f]), (rep, k, f)
        gg.close()


def test_batch_rejects_bad_grids(gpu):
    from simulation_amd import _lib
    from simulation_amd.batch import BatchSolver, GridSpec
    with pytest.raises(_lib.FdtdError, match="source outside"):
        BatchSolver([GridSpec(64, 64, 8, [], (70, 10, 20, 6.0), None)], 10, device=gpu)
    b = BatchSolver([GridSpec(64, 64, 8, [], (30, 10, 20, 6.0), None)], 10, depth=4, device=gpu)
    with pytest.raises(_lib.FdtdError, match="even number"):
        b.run(1)
    with pytest.raises(_lib.FdtdError, match="exceed"):
        b.run(12)


def test_full_c3_batch_in_one_graph(gpu):
    """The whole seeded batch of 256 grids (sides 128..1024, per-grid row
    blocks) advanced 64 steps by one captured graph; 34 of its grids -- the
    smallest, the largest and every eighth of the list --
    against their stand-alone fd2d_tm_run, bit for bit, probes included."""
    from simulation_amd.batch import BatchSolver, c3_specs, single_solver
    specs = c3_specs(0)
    ns = 64
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        b = BatchSolver(specs, ns, device=gpu)
        print(f"[batch] full C3: {len(specs)} grids, {b.desc.nunits} units, {b.cells} cells",
              flush=True)
        b.run_graph(ns)
        torch.cuda.synchronize()
        sizes = [sp.nx * sp.ny for sp in specs]
        pick = {int(np.argmin(sizes)), int(np.argmax(sizes))}
        pick |= set(range(0, len(specs), len(specs) // 30))
        pick = sorted(pick)
        assert len(pick) >= 32
        for k in pick:
            sp = specs[k]
            one = single_solver(sp, ns, gpu)
            one.run(ns)
            ref = one.state()
            got = b.grid_state(k)
            for f in FIELDS:
                assert same_bits(got[f], ref[f]), (k, sp.nx, f)
            if sp.probe:
                assert same_bits(b.probe_trace(k), one.probe_traces()[0]), k
            del one
        b.close()


def test_batch_free_space_units_same_bits(gpu):
    """Batches that skip naz/nbz in free-space units (fd2d_batch_desc.unit_work)
    give the bits of batches that read them, also after a medium changes
    between two runs (every run records its units afresh)."""
    from simulation_amd.batch import BatchSolver
    specs = _specs()
    print("[batch] free-space units", flush=True)
    a = BatchSolver(specs, 80, depth=6, rows=20, cols=1, device=gpu)
    b = BatchSolver(specs, 80, depth=6, rows=20, cols=1, device=gpu, unit_flags=False)
    assert a._unit_work is not None and b._unit_work is None

    def same():
        for k in range(len(specs)):
            ga, gb = a.grid_state(k), b.grid_state(k)
            for f in FIELDS:
                assert same_bits(ga[f], gb[f]), (k, f)
            if specs[k].probe:
                assert same_bits(a.probe_trace(k), b.probe_trace(k)), k
    for x in (a, b):
        x.run(40)
    same()
    assert int((a._unit_work & 1).sum()) > 0          # some units were free space
    # a lossy patch appears in the free space of grid 3 (333 x 141, no
    # cylinders; source at (20, 70)), inside units recorded as free space
    for x in (a, b):
        x.state[3]["md"].naz[40:60, 60:80] = 0.5
        x.state[3]["md"].nbz[40:60, 60:80] = 0.01
        x.run(40)
    same()
    assert np.abs(a.grid_state(3)["ez"][40:60, 60:80]).max() > 1e-6
