"""The learned boundary's hand-over rule (CPU): when an energy check hands a
run over to the PML, from the energies a guarded run would see."""
import math

from simulation_amd.learned import FallbackPolicy, fallback_reason


def test_non_finite_energy_hands_over():
    p = FallbackPolicy()
    assert "nan" in fallback_reason(p, float("nan"), 1.0, 64)
    assert "inf" in fallback_reason(p, math.inf, 1.0, 64)


def test_growth_past_the_factor_hands_over():
    p = FallbackPolicy(growth=4.0)
    assert fallback_reason(p, 3.9, 1.0, 640) is None
    assert fallback_reason(p, 4.0, 1.0, 640) is None        # at the factor: not past it
    assert "above" in fallback_reason(p, 4.01, 1.0, 640)


def test_growth_ignored_while_sources_inject_or_below_floor():
    p = FallbackPolicy(growth=2.0, quiet_after=100, floor=1e-6)
    assert fallback_reason(p, 50.0, 1.0, 64) is None          # source still on
    assert fallback_reason(p, 50.0, 1.0, 100) is not None
    assert fallback_reason(p, 5e-6, 1e-7, 200) is None        # nothing seen above the floor yet
    # a first check (nothing seen yet) never hands over on growth
    assert fallback_reason(FallbackPolicy(), 1e9, 0.0, 64) is None


def test_forced_handover():
    p = FallbackPolicy(force_at=96)
    assert fallback_reason(p, 1.0, 1.0, 64) is None
    assert "forced" in fallback_reason(p, 1.0, 1.0, 96)
    assert "forced" in fallback_reason(p, 1.0, 1.0, 128)


def test_a_decaying_run_never_hands_over():
    """Energies of a pulse leaving through an absorbing boundary: rise while
    it spreads from its initial shape, then decay."""
    p = FallbackPolicy()
    energies = [0.8, 1.0, 0.97, 0.9, 0.5, 0.1, 1e-3, 1e-5, 3e-5, 2e-5]
    emax = 0.0
    for k, e in enumerate(energies):
        assert fallback_reason(p, e, emax, 64 * (k + 1)) is None
        emax = max(emax, e)
