"""Pins of the oracle beyond the reference's fixtures: closed forms, an
exact integer-lattice count for the sub-sampled cylinder, symmetries and an
independent NumPy restatement.  Each check is chosen so that a plausible
slip (a dropped term, a wrong index or offset, a transposed operand) fails."""
from fractions import Fraction

import numpy as np
import pytest

from oracle import numpy_oracle as NO


# ------------------------------------------------------------------ PML

@pytest.mark.parametrize("nx,ny,npml", [(100, 100, 8), (97, 131, 8), (300, 200, 32), (200, 300, 64)])
def test_pmlparam_structure_and_closed_forms(oracle, nx, ny, npml):
    p = oracle.fd2d_pmlparam(nx, ny, npml)
    # outer cell: xm = 0.33 exactly -> gx2 = 1/1.33, gx3 = 0.67/1.33
    assert p["gx2"][0] == 1 / (1 + 0.33) and p["gx3"][0] == (1 - 0.33) / (1 + 0.33)
    # inner cell of the f-profile: xn = 0.33*(0.5/npml)^3 to within an ulp
    xn = 0.33 * (0.5 / npml) ** 3
    assert abs(p["fx1"][npml - 1] - xn) <= np.spacing(xn)
    # mirror placement: f at n and N-2-n, g at n and N-1-n, for both axes
    for ax, N in (("x", nx), ("y", ny)):
        n = np.arange(npml)
        for k in ("1", "2", "3"):
            f = p["f" + ax + k]
            assert np.array_equal(f[n], f[N - 2 - n])
        for k in ("2", "3"):
            g = p["g" + ax + k]
            assert np.array_equal(g[n], g[N - 1 - n])
        # defaults in the interior
        core = slice(npml, N - 1 - npml)
        assert np.all(p["f" + ax + "1"][core] == 0) and np.all(p["f" + ax + "2"][core] == 1)
        assert np.all(p["g" + ax + "3"][npml:N - npml] == 1)
        # algebraic identity (1-x)/(1+x) = 2/(1+x) - 1, and monotone grading
        f2, f3 = p["f" + ax + "2"][:npml], p["f" + ax + "3"][:npml]
        assert np.allclose(f3, 2 * f2 - 1, rtol=0, atol=4e-16)
        assert np.all(np.diff(p["f" + ax + "1"][:npml]) < 0)
    # x and y profiles coincide on square grids only through npml, not position
    assert np.array_equal(p["fx1"][:npml], p["fy1"][:npml])


def test_pmlparam_fp32_storage_independent_of_cube_rounding():
    """The device forms x^3 rounded once (fused multiply-add error terms); the
    reference uses libm pow.  Stored as float32, all five profiles agree for
    every npml <= 300 whichever of the two roundings (or exact rationals) is
    used, so the FP32 coefficients do not depend on pow's last ulp."""
    def prof(c):
        x = 0.33 * c
        return (np.float32(x), np.float32(1 / (1 + x)), np.float32((1 - x) / (1 + x)))
    for npml in range(1, 301):
        for n in range(npml):
            for x in ((npml - n) / npml, (npml - n - 0.5) / npml):
                cr = float(Fraction(x) ** 3)
                assert prof(x ** 3) == prof(cr)


# ----------------------------------------------------------- dielectric

def lattice_count(nx, ny, npml, cx, cy, r):
    """Exact number of the 3x3 sub-samples (i - m/3, j - n/3) of each cell
    within radius r of (cx, cy), by integer arithmetic on the 6x-refined
    lattice: (6(cx-i)+2m)^2 + (6(cy-j)+2n)^2 <= 36 r^2 (cx, cy half-integers
    or integers; r a multiple of 1/2)."""
    i = np.arange(nx)[:, None]
    j = np.arange(ny)[None, :]
    X = int(round(6 * cx)) - 6 * i
    Y = int(round(6 * cy)) - 6 * j
    R = int(round(36 * r * r))
    cnt = np.zeros((nx, ny), np.int64)
    for m in (-1, 0, 1):
        for n in (-1, 0, 1):
            cnt += ((X + 2 * m) ** 2 + (Y + 2 * n) ** 2 <= R)
    mask = np.zeros((nx, ny), bool)
    mask[npml:nx - npml, npml:ny - npml] = True
    return np.where(mask, cnt, 0)


@pytest.mark.parametrize("nx,ny,r", [(100, 100, 14), (97, 131, 30), (120, 90, 12.5), (64, 64, 0)])
def test_dielectric_subsample_count_is_exact(oracle, nx, ny, r):
    """The float sub-sampling (m/3, Numba form) counts exactly the lattice
    points: no near-tie exists, since off-centre sub-samples give integer
    lattice norms that differ from 36 r^2 by at least 1."""
    epsr, sigma, npml = 30.0, 0.3, 8
    dt = float(np.float32(np.float32(0.01) / np.float32(6e8)))
    cx, cy = nx / 2 - 1, ny / 2 - 1
    naz, nbz = oracle.fd2d_dielectric(nx, ny, npml, dt, [(cx, cy, r, epsr, sigma)])
    cnt = lattice_count(nx, ny, npml, cx, cy, r)
    eps0 = 8.854e-12
    got_sig = np.rint(nbz * eps0 * 9 / (sigma * dt)).astype(np.int64)
    got_eps = np.rint((1 / naz - nbz - 1) * 9 / (epsr - 1)).astype(np.int64)
    assert np.array_equal(got_sig, cnt)
    assert np.array_equal(got_eps, cnt)
    if r == 0:
        assert cnt.sum() == (1 if float(cx).is_integer() and float(cy).is_integer() else 0)


def test_dielectric_values_symmetry_and_limits(oracle):
    nx, ny, npml, r = 101, 101, 8, 20
    dt = 1e-11
    cx = cy = 49.0
    naz, nbz = oracle.fd2d_dielectric(nx, ny, npml, dt, [(cx, cy, r, 4.0, 0.0)])
    c = 49
    a = np.arange(-35, 36)
    # exact point symmetry about the integer centre and the diagonal mirror
    assert np.array_equal(naz[c + a][:, c + a], naz[c - a][:, c - a])
    assert np.array_equal(naz[c + a][:, c + a], naz[c + a][:, c + a].T)
    # lossless deep interior: 1/naz = eps_r to rounding, nbz = 0
    assert abs(1 / naz[c, c] - 4.0) < 1e-14 and nbz[c, c] == 0
    # outside the cylinder and inside the PML band: vacuum
    assert naz[0, 0] == 1 and naz[c, c + r + 2] == 1
    # two cylinders: the first containing one wins the sub-sample
    naz2, _ = oracle.fd2d_dielectric(nx, ny, npml, dt, [(cx, cy, r, 4.0, 0.0), (cx, cy, r + 5, 9.0, 0.0)])
    assert naz2[c, c] == naz[c, c]
    assert abs(1 / naz2[c, c + r + 3] - 9.0) < 1e-14


# ------------------------------------------------------------- updates

@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("with_pml", [True, False])
def test_numpy_restatement_agrees_with_c_oracle(oracle, dtype, with_pml):
    nx, ny, ns = 41, 57, 70
    pml = None
    if with_pml:
        pml = {k: v.astype(dtype) for k, v in oracle.fd2d_pmlparam(nx, ny, 6).items()}
    naz, nbz = oracle.fd2d_dielectric(nx, ny, 6, 1.6666666e-11, [(19, 27, 9, 30.0, 0.3)])
    naz, nbz = naz.astype(dtype), nbz.astype(dtype)
    w = oracle.waveform("gaussian", 0, ns, dtype, center=20, spread=6.0)
    st = oracle.TMState(nx, ny, dtype)
    pr = oracle.fd2d_tm_run(st, 0, ns, pml, naz, nbz, [(20, 28)], w[None, :], [(5, 5), (30, 40)])
    nst = {k: np.zeros((nx, ny), dtype) for k in ("dz", "iz", "ez", "hx", "hy", "ihx", "ihy")}
    npr = NO.tm_run(nst, ns, pml, naz, nbz, [(20, 28)], [w], [(5, 5), (30, 40)])
    for k in nst:
        assert np.array_equal(nst[k], getattr(st, k)), k
    assert np.array_equal(npr, pr)
    assert np.abs(st.ez).max() > 1e-3


def test_free_space_pec_edges_and_transpose_symmetry(oracle):
    """Free space, square grid, source on the diagonal: PEC edges never change
    and Ez is symmetric about the diagonal to rounding (FP addition order
    differs between (i,j) and (j,i))."""
    n, ns = 64, 80
    st = oracle.TMState(n, n, np.float64)
    w = oracle.waveform("gaussian", 0, ns, np.float64, center=20, spread=6.0)
    oracle.fd2d_tm_run(st, 0, ns, None, np.ones((n, n)), None, [(32, 32)], w[None, :])
    assert np.all(st.dz[0, :] == 0) and np.all(st.dz[:, 0] == 0)
    assert np.all(st.hx[n - 1, :] == 0) and np.all(st.hy[:, n - 1] == 0)
    assert np.abs(st.ez).max() > 1e-2
    assert np.allclose(st.ez, st.ez.T, rtol=0, atol=1e-13)
    assert np.allclose(st.hx, -st.hy.T, rtol=0, atol=1e-13)


def test_pml_form_with_default_profiles_equals_free_space_form(oracle):
    """With default coefficients (f*1 = 0, the rest 1) the PML update bodies
    reduce to the free-space bodies of fd2d_3_1 exactly (multiplying by 1 and
    adding 0*ihx are exact)."""
    n, ns = 48, 60
    dflt = {k: (np.zeros(n) if k.endswith("1") else np.ones(n)) for k in oracle.PML_FIELDS}
    w = oracle.waveform("gaussian", 0, ns, np.float32, center=20, spread=6.0)
    a = oracle.TMState(n, n, np.float32)
    b = oracle.TMState(n, n, np.float32)
    oracle.fd2d_tm_run(a, 0, ns, dflt, np.ones((n, n)), None, [(24, 24)], w[None, :])
    oracle.fd2d_tm_run(b, 0, ns, None, np.ones((n, n)), None, [(24, 24)], w[None, :])
    for k in ("dz", "ez", "hx", "hy"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


def test_fd1d_mirror_symmetry(oracle):
    nx, kc = 401, 200
    ex, hy, bc = np.zeros(nx, np.float32), np.zeros(nx, np.float32), np.zeros(4, np.float32)
    w = oracle.waveform("gaussian", 0, 350, np.float32, center=40, spread=12.0)
    oracle.fd1d_run(nx, 0, 350, ex, hy, bc, None, None, 1, kc, w)
    m = np.arange(1, 200)
    assert np.array_equal(ex[kc - m], ex[kc + m])
    assert np.array_equal(hy[kc + m], -hy[kc - m - 1])


def test_fd1d_abc_absorbs(oracle):
    """After the pulse leaves, the two-step boundary leaves a residual ~1e-4
    of the pulse (what numerical dispersion lets through), while PEC walls
    (no boundary) keep the pulse on the grid at full amplitude."""
    nx = 200
    w = oracle.waveform("gaussian", 0, 1200, np.float64, center=40, spread=12.0)
    ex, hy, bc = np.zeros(nx), np.zeros(nx), np.zeros(4)
    oracle.fd1d_run(nx, 0, 1200, ex, hy, bc, None, None, 1, 100, w)
    ex2, hy2 = np.zeros(nx), np.zeros(nx)
    oracle.fd1d_run(nx, 0, 1200, ex2, hy2, None, None, None, 1, 100, w)
    assert np.abs(ex).max() < 1e-4
    assert np.abs(ex2).max() > 0.3


def test_energy_matches_numpy(oracle):
    g = np.random.default_rng(3)
    a, b, c = (g.standard_normal((50, 70)).astype(np.float32) for _ in range(3))
    ref = float(np.sum(a.astype(np.float64) ** 2 + b.astype(np.float64) ** 2 + c.astype(np.float64) ** 2))
    assert abs(oracle.fd2d_energy(a, b, c) - ref) <= 1e-12 * ref
