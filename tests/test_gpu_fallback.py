"""The learned boundary with the plain PML kept as fallback
(learned.GuardedLearnedSolver) on the GPU.

* A forced hand-over continues on the npml=32 form from the state the
  learned run held before the failing chunk, and gives the same bits as a
  PML run built by hand (fd2d.new_pmlayer / pmlparam / TMSolver, the window
  copied into the centre of the padded grid) from that same state.
* A guarded run that never hands over gives the bits of the unguarded
  learned run.
* A boundary that goes unstable (its linear path scaled up 50x) is caught
  by the energy check and handed over before the fields blow up; the PML run
  then stays bounded.
"""
import numpy as np
import pytest
import torch
from conftest import same_bits

pytestmark = pytest.mark.gpu

N = 192


def _medium(n, dev):
    from simulation_amd import fd2d
    return fd2d.dielectric(n, n, 0, 0, fd2d.time_step(), 1.0, 0.0, dev,
                           cylinders=[(n / 2 + 20.5, n / 2 - 15.0, 18.0, 4.0, 0.02)])


def _net(dev, scale_linear=1.0):
    from simulation_amd import learned
    import os
    p = (np.load(learned.WEIGHTS, allow_pickle=False)["params"].copy()
         if os.path.exists(learned.WEIGHTS) else np.zeros(learned.NPARAM, np.float32))
    p[learned.P_L:] *= scale_linear
    return learned.Net(p, dev)


def _source(ns):
    from simulation_amd import fd1d
    return fd1d.waveform("gaussian", ns, 0, np.float32, center=20, spread=6.0)


def test_forced_handover_matches_pml_run_from_the_same_state(gpu):
    from simulation_amd import fd2d, learned
    ns = 200
    md = _medium(N, gpu)
    src, probe = (N // 2, N // 2 - 10), (N // 2 + 7, N // 2 + 3)
    pol = learned.FallbackPolicy(every=32, force_at=96)
    print("[fallback] forced hand-over", flush=True)
    g = learned.GuardedLearnedSolver(N, N, _net(gpu), md, [(*src, _source(ns))], [probe],
                                     nsteps_max=ns, policy=pol, device=gpu)
    g.run(ns)
    assert [e["step"] for e in g.events] == [64]
    assert g.t == ns and g.pml is not None
    # by hand: the learned run to step 64, its window in the centre of the
    # padded grid, an npml=32 PML around it
    ref = learned.LearnedTMSolver(N, N, _net(gpu), md, [(*src, _source(ns))], [probe],
                                  nsteps_max=ns, device=gpu)
    ref.run(64)
    p = 32
    M = N + 2 * p
    pml = fd2d.new_pmlayer(M, M, gpu)
    fd2d.pmlparam(M, M, p, pml)
    naz = torch.ones((M, M), device=gpu)
    nbz = torch.zeros((M, M), device=gpu)
    naz[p:p + N, p:p + N] = md.naz
    nbz[p:p + N, p:p + N] = md.nbz
    s = fd2d.TMSolver(M, M, pml, fd2d.medium(naz, nbz), [(src[0] + p, src[1] + p, _source(ns))],
                      [(probe[0] + p, probe[1] + p)], nsteps_max=ns, device=gpu)
    for k in ("dz", "iz", "ez", "hx", "hy"):
        getattr(s, k)[p:p + N, p:p + N] = getattr(ref.tm, k)
    s.probe_out[:, :64] = ref.tm.probe_out[:, :64]
    s.t = 64
    s.run(ns - 64)
    a, b = g.pml.state(), s.state()
    for k in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy"):
        assert same_bits(a[k], b[k]), k
    assert same_bits(g.probe_traces(), s.probe_traces())
    assert np.abs(a["ez"]).max() > 1e-3


def test_guard_without_handover_keeps_the_learned_bits(gpu):
    from simulation_amd import learned
    ns = 300
    md = _medium(N, gpu)
    src = (N // 2, N // 2)
    print("[fallback] no hand-over", flush=True)
    g = learned.GuardedLearnedSolver(N, N, _net(gpu), md, [(*src, _source(ns))], [],
                                     nsteps_max=ns, policy=learned.FallbackPolicy(every=50,
                                                                                  quiet_after=60),
                                     device=gpu)
    g.run(ns)
    r = learned.LearnedTMSolver(N, N, _net(gpu), md, [(*src, _source(ns))], [], nsteps_max=ns,
                                device=gpu)
    r.run(50)
    r.run(250)
    assert g.events == [] and g.pml is None
    a, b = g.learned.state(), r.state()
    for k in ("dz", "iz", "ez", "hx", "hy"):
        assert same_bits(a[k], b[k]), k
    assert len(g.energies) == 6 and all(np.isfinite(e) for _, e in g.energies)


def test_unstable_boundary_is_caught_and_handed_over(gpu):
    """A learned boundary made unstable (linear path x50) blows up; the
    energy check hands over at the first check past the growth factor (or
    at a non-finite energy), from the last good state, and the PML run
    stays bounded by the energy seen before the blow-up."""
    from simulation_amd import learned
    ns = 800
    md = _medium(N, gpu)
    print("[fallback] unstable boundary", flush=True)
    g = learned.GuardedLearnedSolver(N, N, _net(gpu, 50.0), md, [], [], nsteps_max=ns,
                                     policy=learned.FallbackPolicy(every=32), device=gpu)
    g.learned.tm.dz.copy_(learned.pulse(N, N, N / 2 + 5, N / 2 - 3, 3.0, gpu))
    g.run(ns)
    assert len(g.events) == 1, g.energies
    ev = g.events[0]
    print("[fallback] event", ev, flush=True)
    assert "above" in ev["reason"] or "nan" in ev["reason"] or "inf" in ev["reason"]
    good = [e for t, e in g.energies if t <= ev["step"]]
    assert good and all(np.isfinite(good))
    e_end = g.pml.energy()
    assert np.isfinite(e_end) and e_end <= max(good)
