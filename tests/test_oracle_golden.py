"""The CPU oracle against the golden fixtures written by the reference's own C
programs (tests/golden/make_golden.py): bit-exact, FP32 and FP64."""
import numpy as np
import pytest

from cases import (DFORM_CASES, FD1D_CASES, FD2D_CASES, TFSF_CASES, dform_case, fd1d_case,
                   fd2d_case, tfsf_case)
from conftest import same_bits


def _wave(O):
    return lambda kind, ns, dt, **kw: O.waveform(kind, 0, ns, dt, **kw)


@pytest.mark.parametrize("name", FD1D_CASES)
def test_fd1d_oracle_matches_reference(oracle, name):
    c = fd1d_case(name, _wave(oracle))
    dt, nx = c["dtype"], c["nx"]
    ex = np.zeros(nx, dt)
    hy = np.zeros(nx, dt)
    bc = np.zeros(4, dt) if c["abc"] else None
    ca = None if c["ca"] is None else np.ascontiguousarray(c["ca"])
    cb = None if c["cb"] is None else np.ascontiguousarray(c["cb"])
    snaps_ex, snaps_hy = [], []
    probes = []
    t = 0
    cuts = c["snap_t"] or [c["ns"]]
    for cut in cuts:
        n = cut - t
        probes.append(oracle.fd1d_run(nx, t, n, ex, hy, bc, ca, cb, c["src_mode"], c["src_pos"],
                                      np.ascontiguousarray(c["wave"][t:cut]), c["probes"]))
        t = cut
        snaps_ex.append(ex.copy())
        snaps_hy.append(hy.copy())
    e = c["exp"]
    assert same_bits(ex, e["out_ex"])
    assert same_bits(hy, e["out_hy"])
    assert same_bits(np.concatenate(probes, axis=1), e["out_probe"])
    if c["abc"]:
        assert same_bits(bc, e["out_bc"])
    if c["snap_t"]:
        assert same_bits(np.array(snaps_ex), e["out_snap_ex"])
        assert same_bits(np.array(snaps_hy), e["out_snap_hy"])
    # the case is not trivially zero: the pulse is on the grid or passed the probe
    assert max(np.abs(e["out_ex"]).max(), np.abs(e["out_probe"]).max()) > 0.1


@pytest.mark.parametrize("name", FD2D_CASES)
def test_fd2d_oracle_matches_reference(oracle, name):
    c = fd2d_case(name, _wave(oracle))
    st = oracle.TMState(c["nx"], c["ny"], c["dtype"])
    pr = oracle.fd2d_tm_run(st, 0, c["ns"], c["pml"], c["naz"], c["nbz"], c["sources"],
                            c["wave"][None, :], c["probes"])
    e = c["exp"]
    for k in ("dz", "ez", "hx", "hy", "iz", "ihx", "ihy"):
        if "out_" + k in e:
            assert same_bits(getattr(st, k), e["out_" + k]), k
    assert same_bits(pr, e["out_probe"])
    assert np.abs(st.ez).max() > 1e-4


def test_dielectric_integer_subsampling_matches_reference(oracle):
    """The accumulation and naz/nbz formulas, pinned bit-exactly by the C
    program's dielectric (fd2d/clang/fd2d_3_4.c:149-176, integer m/3)."""
    from conftest import golden
    m, a = golden("fd2d_dielectric_c")
    cyl = [(m["nx"] / 2 - 1, m["ny"] / 2 - 1, m["rgrid"], m["epsr"], m["sigma"])]
    naz, nbz = oracle.fd2d_dielectric(m["nx"], m["ny"], m["npml"], m["dt"], cyl, int_sub=True)
    assert same_bits(naz, a["out_naz"])
    assert same_bits(nbz, a["out_nbz"])
    assert (a["out_nbz"] > 0).sum() > 100


@pytest.mark.parametrize("npml", [8, 32, 64])
def test_pmlparam_matches_reference_after_fp32_storage(oracle, npml):
    """The PyCUDA profiles (FP64 pow, stored FP32) equal the C program's FP64
    profiles (fd2d/clang/fd2d_3_2.c:70-81, a different association of the
    cube) once both are stored as float32, and differ from them in FP64 by at
    most 2 ulp; the FP32 program's own float arithmetic (test_3_2.c:70-81)
    agrees within 2e-7 relative."""
    from conftest import golden
    m, a = golden("fd2d_pmlparam_c")
    p = oracle.fd2d_pmlparam(m["nx"], m["ny"], npml)
    for k in oracle.PML_FIELDS:
        ref64 = a[f"out_fd2d_3_2_{npml}_{k}"]
        ref32 = a[f"out_test_3_2_{npml}_{k}"]
        assert same_bits(p[k].astype(np.float32), ref64.astype(np.float32)), k
        ulp = np.spacing(np.abs(ref64))
        assert np.all(np.abs(p[k] - ref64) <= 2 * ulp), k
        assert np.allclose(p[k].astype(np.float32), ref32, rtol=2e-7, atol=0), k


@pytest.mark.parametrize("name", TFSF_CASES)
def test_tfsf_oracle_matches_reference(oracle, name):
    """Plane wave through the TF/SF box with the running DFT: the NumPy
    restatement against the reference program's own functions, bit-exact
    (fields, incident grid and its boundary state, all DFT planes)."""
    from oracle import numpy_oracle as NO
    c = tfsf_case(name, oracle.gaussian)
    dt, nx, ny, nf = c["dtype"], c["nx"], c["ny"], len(c["freq"])
    st = {k: np.zeros((nx, ny), dt) for k in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy")}
    inc = {"ezi": np.zeros(ny, dt), "hxi": np.zeros(ny, dt), "bc": np.zeros(4, dt)}
    ft = {"r_pt": np.zeros((nf, nx, ny), dt), "i_pt": np.zeros((nf, nx, ny), dt),
          "r_in": np.zeros(nf, dt), "i_in": np.zeros(nf, dt)}
    NO.tfsf_run(st, inc, c["ns"], c["npml"], c["pml"], c["naz"], c["nbz"], c["wave"],
                c["freq"], c["dt"], ft, inc_pos=c["inc_pos"])
    e = c["exp"]
    for k, v in list(st.items()) + list(inc.items()):
        if c["nbz"] is None and k == "iz":
            continue
        assert same_bits(v, e["out_" + k]), k
    for k, v in ft.items():
        if nf:
            assert same_bits(v, e["out_ft_" + k]), k
    # the plane wave crossed the box and the cylinder scattered it
    assert np.abs(e["out_ez"]).max() > 0.1
    if nf:
        assert np.abs(e["out_ft_r_pt"]).max() > 1.0


@pytest.mark.parametrize("name", DFORM_CASES)
def test_dform_oracle_matches_reference(oracle, name):
    """1-D D/E form with a lossy or Debye half space and the running DFT: the
    NumPy restatement (coefficients included) against the reference
    program's own functions, bit-exact."""
    from oracle import numpy_oracle as NO
    c = dform_case(name, oracle.gaussian, oracle.sinusoidal)
    dt, nx, m = c["dtype"], c["nx"], c["meta"]
    md = NO.dielectric_dform(nx, c["dt"], m["chi"], m["tau"], m["epsr"], m["sigma"], c["debye"], dt)
    e = c["exp"]
    for k, v in md.items():
        assert same_bits(v, e["out_md_" + k]), k
    st = {k: np.zeros(nx, dt) for k in ("dx", "ix", "sx", "ex", "hy")}
    st["bc"] = np.zeros(4, dt)
    nf = len(c["freq"])
    ft = {"r_pt": np.zeros((nf, nx), dt), "i_pt": np.zeros((nf, nx), dt),
          "r_in": np.zeros(nf, dt), "i_in": np.zeros(nf, dt)}
    pr = NO.dform_run(st, md, c["ns"], c["wave"], c["freq"], c["dt"], ft, src_pos=c["src_pos"],
                      probes=c["probes"])
    for k in ("dx", "ix", "ex", "hy", "bc") + (("sx",) if c["debye"] else ()):
        assert same_bits(st[k], e["out_" + k]), k
    for k, v in ft.items():
        assert same_bits(v, e["out_ft_" + k]), k
    assert same_bits(pr, e["out_probe"])
    assert np.abs(e["out_ex"]).max() > 0.1
