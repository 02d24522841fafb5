"""The C-ABI library on the CPU: it loads, exports every symbol include/fdtd.h
declares, and rejects bad arguments on the host before any device work."""
import ctypes as C
import os
import subprocess

import pytest

from conftest import ROOT


@pytest.fixture(scope="module")
def L():
    from simulation_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libfdtd_hip.so not built (python -m simulation_amd.build)")
    return _lib


def test_header_symbols_are_exported(L):
    syms = L.header_symbols()
    assert len(syms) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    assert set(syms) == set(L.SIGNATURES), set(syms) ^ set(L.SIGNATURES)


def test_abi_version_and_host_waveforms(L, oracle):
    lib = L.lib()
    assert lib.fdtd_abi_version() == 4
    for t in (1, 7, 20, 40, 333):
        assert lib.fdtd_gaussian_f32(t, 20, 6.0) == oracle.gaussian(t, 20, 6.0)
        assert lib.fdtd_gaussian_f64(t, 40, 12.0) == oracle.gaussian(t, 40, 12.0, "float64")
        assert lib.fdtd_sinusoidal_f32(t, 0.01, 700e6) == oracle.sinusoidal(t, 0.01, 700e6)
        assert lib.fdtd_sinusoidal_f64(t, 0.01, 1500e6) == oracle.sinusoidal(t, 0.01, 1500e6, "float64")


def _err(L, name, *args):
    rc = getattr(L.lib(), name)(*args)
    return rc, L.lib().fdtd_last_error().decode()


def test_argument_validation_is_host_side(L):
    fake = 0x1000  # never dereferenced: every call below is rejected first
    rc, msg = _err(L, "fd1d_exfield", 1, 1, None, None, fake, fake, 0, 0, 0.0, None)
    assert rc == L.FDTD_EINVAL and "nx" in msg
    rc, msg = _err(L, "fd1d_exfield", 1, 10, fake, None, fake, fake, 0, 0, 0.0, None)
    assert rc == L.FDTD_EINVAL and "cb" in msg
    rc, msg = _err(L, "fd1d_exfield", 1, 10, None, None, fake, fake, 1, 10, 0.0, None)
    assert rc == L.FDTD_EINVAL and "src_pos" in msg
    rc, msg = _err(L, "fd1d_hyfield", 2, fake, fake, None, None)
    assert rc == L.FDTD_EINVAL
    big = L.lib().fd1d_run_max_nx(4) + 1
    rc, msg = _err(L, "fd1d_run", big, 0, 1, None, None, fake, fake, None, 0, 0, None, 0, None,
                   None, None)
    assert rc == L.FDTD_EINVAL and "limit" in msg
    rc, msg = _err(L, "fd2d_dfield", 1, 1, 5, None, fake, fake, fake, 0, None, None, None, None)
    assert rc == L.FDTD_EINVAL
    pml = L.FdtdPml()
    pml.fx1 = C.cast(fake, C.POINTER(C.c_float))
    rc, msg = _err(L, "fd2d_hfield", 8, 8, C.byref(pml), fake, fake, fake, fake, fake, None)
    assert rc == L.FDTD_EINVAL and "complete" in msg
    rc, msg = _err(L, "fd2d_pmlparam", 8, 8, 7, None, None)
    assert rc == L.FDTD_EINVAL
    rc, msg = _err(L, "fd2d_tm_run", None, 0, 1, None)
    assert rc == L.FDTD_EINVAL and "NULL" in msg
    d = L.Fd2dTmDesc()
    d.nx, d.ny = 64, 64
    for k in ("dz", "ez", "hx", "hy", "iz"):
        setattr(d, k, C.cast(fake, C.POINTER(C.c_float)))
    d.md.naz = C.cast(fake, C.POINTER(C.c_float))
    d.nsrc = 1
    d.src_i[0], d.src_j[0] = 70, 3
    d.src_wave = C.cast(fake, C.POINTER(C.c_float))
    d.wave_len = 10
    rc, msg = _err(L, "fd2d_tm_run", C.byref(d), 0, 5, None)
    assert rc == L.FDTD_EINVAL and "outside" in msg
    d.src_i[0] = 3
    rc, msg = _err(L, "fd2d_tm_run", C.byref(d), 0, 11, None)
    assert rc == L.FDTD_EINVAL and "wave_len" in msg
    d.depth = 99
    rc, msg = _err(L, "fd2d_tm_run", C.byref(d), 0, 5, None)
    assert rc == L.FDTD_EINVAL and "depth" in msg
    rc, msg = _err(L, "fd2d_energy", 8, 8, 0, 9, 0, 8, fake, fake, fake, fake, None, None)
    assert rc == L.FDTD_EINVAL and "window" in msg


def test_tm_plan_even_launch_counts(L):
    """Launch plans always end in the caller's state set: an even number of
    ping-pong launches, plus one in-place step when that is impossible."""
    fake = C.cast(0x1000, C.POINTER(C.c_float))
    d = L.Fd2dTmDesc()
    d.nx, d.ny = 4096, 4096
    for k in ("dz", "ez", "hx", "hy"):
        setattr(d, k, fake)
    d.md.naz = fake
    out = [C.c_int() for _ in range(4)]
    for depth in (1, 2, 5, 8):
        d.depth = depth
        for ns in (1, 2, 3, 7, 8, 9, 17, 500, 2000):
            rc = L.lib().fd2d_tm_plan(C.byref(d), ns, *[C.byref(o) for o in out])
            assert rc == 0, L.lib().fdtd_last_error()
            launches = out[0].value
            assert out[1].value == depth
            assert launches >= (ns + depth - 1) // depth


def test_product_path_has_no_oracle_or_cpu_fallback():
    """simulation_amd never imports or loads the oracle, and every compute
    entry point goes through libfdtd_hip.so."""
    pkg = os.path.join(ROOT, "simulation_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in txt.lower().replace("the oracle", ""), f


def test_lbc_sizes_match_the_oracle_layout():
    """Buffer sizes the learned-boundary entry points expect (the caller
    allocates from these) agree with the sample layout of the oracle."""
    from oracle import lbc_oracle as LO
    from simulation_amd import _lib
    L = _lib.lib()
    assert L.lbc_param_count() == LO.NPARAM
    for nx, ny in ((64, 64), (96, 112), (2048, 2048)):
        assert L.lbc_hist_elems(nx, ny) == len(LO.positions(nx, ny)) * LO.HIST * LO.NSAMP
    assert L.fd2d_lbc_margin() >= 3 * 8 + 6      # deepest frame region (3K+6 at K=8)


def test_lbc_work_and_operand_buffer_sizes():
    """The frame tables (one set per depth 1..8) need a work buffer that grows
    with the perimeter; grids below 64 x 64 are rejected (size 0).  The bf16
    buffer holds the row-major operands (W1 128-wide, W2, W3 16 rows, W2^T,
    W3^T 32-wide), 64 lanes x 26 per-lane operands of 8 values and the flag
    that says whether the hidden layers' output layer is non-zero."""
    from simulation_amd import _lib
    L = _lib.lib()
    assert L.lbc_work_bytes(32, 32) == 0
    a, b = L.lbc_work_bytes(512, 512), L.lbc_work_bytes(2048, 2048)
    assert 0 < a < b < 8 * a
    # operands, then the hidden-output flag padded to 8 elements
    assert L.lbc_wbuf_elems() == 64 * 128 + 64 * 64 + 16 * 64 + 64 * 64 + 64 * 32 + 64 * 26 * 8 + 8
