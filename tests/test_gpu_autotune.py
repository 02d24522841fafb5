"""Autotuner: picks one of the timed launch shapes, leaves the solver's state
untouched, and the tuned solver still computes the reference values bit for
bit."""
import numpy as np
import pytest
from conftest import same_bits

pytestmark = pytest.mark.gpu


def test_tune_single_grid_restores_state_and_stays_exact(gpu):
    from simulation_amd import autotune, fd2d
    cfg = fd2d.c2_config(512, 120)
    a = fd2d.build_solver(cfg, gpu)
    b = fd2d.build_solver(cfg, gpu)
    a.run(40)
    b.run(40)
    shapes = [(8, 16, 1), (4, 32, 2), (6, 16, 2)]
    best = autotune.tune(a, steps=24, configs=shapes)
    print("[autotune] best", best, flush=True)
    assert (best["depth"], best["rows"], best["cols"]) in shapes
    assert a.t == 40
    sa, sb = a.state(), b.state()
    for k in fd2d.STATE:
        assert same_bits(sa[k], sb[k]), k
    a.run(60)
    b.run(60)
    sa, sb = a.state(), b.state()
    for k in fd2d.STATE:
        assert same_bits(sa[k], sb[k]), k
    assert same_bits(a.probe_traces(), b.probe_traces())


def test_tune_batch_plans_fastest_and_resets(gpu):
    from simulation_amd import autotune
    from simulation_amd.batch import BatchSolver, c3_specs
    specs = c3_specs(3, count=6, lo=96, hi=300)
    ref = BatchSolver(specs, 64, device=gpu)
    ref.run(64)
    b = BatchSolver(specs, 64, device=gpu)
    shapes = [(8, 32, 1, 0), (4, 16, 2, 0)]
    best = autotune.tune_batch(b, steps=32, configs=shapes)
    print("[autotune batch] best", best, flush=True)
    assert (best["depth"], best["rows"], best["cols"], best["waves"]) in shapes
    assert b.t == 0
    b.run(64)
    for k in range(len(specs)):
        x, y = b.grid_state(k), ref.grid_state(k)
        for f in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy"):
            assert same_bits(x[f], y[f]), (k, f)
        assert same_bits(b.probe_trace(k), ref.probe_trace(k))
