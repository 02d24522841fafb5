"""Pins of the learned-boundary restatement (oracle/lbc_oracle.py), CPU only.

bf16 rounding against torch's bfloat16 conversion; the backward pass against
central differences of the FP64 forward; the canonical boundary sample
against the fourfold symmetry of a centred pulse; the learned step against
the plain update with the outer three cells held at zero; Adam's first step
against its closed form."""
import numpy as np
import pytest
import torch

from oracle import lbc_oracle as L
from oracle import numpy_oracle as NO


def test_bf16_matches_torch_rounding():
    rng = np.random.default_rng(1)
    x = rng.standard_normal(20000).astype(np.float32) * np.float32(10.0) ** rng.integers(-8, 8, 20000)
    # exact ties: low 16 bits = 0x8000 with even and odd upper halves
    ties = (np.arange(64, dtype=np.uint32) << 16 | 0x3F800000 | 0x8000).view(np.float32)
    x = np.concatenate([x, ties, -ties, np.float32([0.0, -0.0, 1.0, 3.0e38])])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(L.bf16(x), want)


def _problem(rng, B=24):
    X = rng.standard_normal((B, L.IN)).astype(np.float32) * np.float32(0.3)
    X[3] = 0.0                                        # an all-zero row (weight 0)
    T = rng.standard_normal((B, L.OUT)).astype(np.float32) * np.float32(0.1)
    p = L.init_params(3)
    p[L.P_W3:L.P_L] *= 5.0                             # non-trivial output layer
    p[L.P_L:] = rng.standard_normal(L.NPARAM - L.P_L).astype(np.float32) * np.float32(0.2)
    return p.astype(np.float32), X, T


def test_backward_matches_central_differences():
    rng = np.random.default_rng(7)
    p, X, T = _problem(rng)
    loss, g = L.loss_and_grad(p, X, T, rounding=False)

    def f(q):
        y, _ = L.forward(q, X, rounding=False)
        th, w = L.targets(X, T, L.pad_samples(X[:, 27:54]))
        e = (y - th) * w[:, None]
        return float(np.sum(e * e) / len(X))
    assert abs(f(p) - loss) < 1e-12
    # the linear path is fitted by least squares, not by gradient
    assert not np.any(g[L.P_L:])
    idx = np.concatenate([rng.choice(L.P_L, 40, replace=False),
                          [0, L.P_L - 1, L.P_L - 3, 64 * 108 + 5]])
    p64 = p.astype(np.float64)
    for k in idx:
        h = 1e-3 * max(1.0, abs(p64[k]))
        a, b = p64.copy(), p64.copy()
        a[k] += h
        b[k] -= h
        fd = (f(a) - f(b)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-4 * abs(fd), (k, fd, g[k])


def test_rounded_gradient_close_to_exact():
    rng = np.random.default_rng(8)
    p, X, T = _problem(rng, 256)
    _, g0 = L.loss_and_grad(p, X, T, rounding=False)
    _, g1 = L.loss_and_grad(p, X, T, rounding=True)
    cos = g0 @ g1 / np.linalg.norm(g0) / np.linalg.norm(g1)
    assert cos > 0.999
    assert np.linalg.norm(g1 - g0) / np.linalg.norm(g0) < 0.05


def test_canonical_sample_fourfold_symmetry():
    """A pulse from the centre of a square free-space grid is symmetric under
    the x-mirror, the y-mirror and the transpose; the canonical samples of
    the four edges of a centred window must then agree."""
    n, c = 61, 30
    st = {k: np.zeros((n, n), np.float32) for k in ("dz", "iz", "ez", "hx", "hy")}
    naz = np.ones((n, n), np.float32)
    t = np.arange(1, 26, dtype=np.float64)
    w = np.exp(-0.5 * ((t - 8.0) / 3.0) ** 2).astype(np.float32)
    for s in range(25):
        NO.dfield(n, n, None, st["dz"], st["hx"], st["hy"])
        st["dz"][c, c] = w[s]
        NO.efield(naz, None, st["dz"], st["iz"], st["ez"])
        NO.hfield(n, n, None, st["ez"], None, None, st["hx"], st["hy"])
    a, b = 20, 41                                   # 21 x 21 window centred on (30, 30)
    win = [st[k][a:b, a:b] for k in ("ez", "hx", "hy")]
    S = L.sample(*win, 21, 21)
    pos = L.positions(21, 21)
    blocks = {e: np.array([S[p] for p, (ee, tt) in enumerate(pos) if ee == e and 3 <= tt < 18])
              for e in range(4)}
    scale = np.abs(blocks[0]).max()
    assert scale > 1e-3
    for e in (1, 2, 3):
        assert np.abs(blocks[e] - blocks[0]).max() <= 1e-5 * scale, e
    # the tangential-coupled entries at offset +1 are zero by definition
    assert not np.any(S[:, [L.sidx(2, d, 1) for d in range(3)]])


def test_zero_network_is_the_two_step_boundary():
    """With every parameter zero the prediction is the baseline: depth d of
    each edge takes the ez of depth d+1 from two steps back (the reference's
    two-step boundary, fd1d/pycuda/fd1d_1_2.py:53-61, on three layers).
    Restated here directly on the ez arrays."""
    nx, ny, ns = 41, 47, 90
    rng = np.random.default_rng(0)
    naz = rng.uniform(0.5, 1.0, (nx, ny)).astype(np.float32)
    nbz = rng.uniform(0.0, 0.03, (nx, ny)).astype(np.float32)
    naz[:8], naz[-8:], naz[:, :8], naz[:, -8:] = 1, 1, 1, 1
    nbz[:8], nbz[-8:], nbz[:, :8], nbz[:, -8:] = 0, 0, 0, 0
    w = np.exp(-0.5 * ((np.arange(1, ns + 1) - 10.0) / 4.0) ** 2).astype(np.float32)
    a = {k: np.zeros((nx, ny), np.float32) for k in ("dz", "iz", "ez", "hx", "hy")}
    L.lbc_run(a, naz, nbz, np.zeros(L.NPARAM, np.float32), ns, sources=[(20, 24)], src_wave=[w])
    b = {k: np.zeros((nx, ny), np.float32) for k in ("dz", "iz", "ez", "hx", "hy")}
    old = [b["ez"].copy(), b["ez"].copy()]           # ez after steps t and t-1
    for s in range(ns):
        NO.dfield(nx, ny, None, b["dz"], b["hx"], b["hy"])
        b["dz"][20, 24] = w[s]
        NO.efield(naz, nbz, b["dz"], b["iz"], b["ez"])
        e1 = old[1]
        for d in range(3):
            b["ez"][d, :] = e1[d + 1, :]
            b["ez"][nx - 1 - d, :] = e1[nx - 2 - d, :]
            b["ez"][3:nx - 3, d] = e1[3:nx - 3, d + 1]
            b["ez"][3:nx - 3, ny - 1 - d] = e1[3:nx - 3, ny - 2 - d]
        NO.hfield(nx, ny, None, b["ez"], None, None, b["hx"], b["hy"])
        old = [b["ez"].copy(), old[0]]
    for k in ("ez", "hx", "hy"):
        assert np.array_equal(a[k], b[k]), k
    assert np.abs(b["ez"][:3]).max() > 1e-3


def test_adam_first_step_closed_form():
    g = np.array([0.5, -2.0, 1e-3, 0.0], np.float32)
    p = np.zeros(4, np.float32)
    m = np.zeros(4, np.float32)
    v = np.zeros(4, np.float32)
    L.adam(p, m, v, g, 1, 0.01)
    want = -0.01 * np.sign(g) * np.abs(g) / (np.abs(g) + 1e-8)
    assert np.allclose(p, want, rtol=1e-5, atol=1e-9)


def _planted(rng, B=400):
    """Rows with targets that the linear path reproduces exactly: T = base +
    s * (L xn) for a known L."""
    X = rng.standard_normal((B, L.IN)).astype(np.float32)
    X[:, rng.choice(L.IN, 10, replace=False)] *= np.float32(5.0)
    X[7] = 0.0
    S1 = L.pad_samples(X[:, L.NFEAT:2 * L.NFEAT])
    Lt = (rng.standard_normal((L.OUT, L.IN)) * 0.3).astype(np.float64)
    s, _, _ = L.scale(X)
    xn = L.normalised(X).astype(np.float64)
    T = (L.base(S1).astype(np.float64) + s[:, None] * (xn @ Lt.T)).astype(np.float32)
    return X, T, S1, Lt


def test_fit_linear_recovers_a_planted_map():
    """Normal equations + solve give back the L that generated the targets
    (to the FP32 rounding of the targets); the all-zero row is left out."""
    rng = np.random.default_rng(11)
    X, T, S1, Lt = _planted(rng)
    G, R, n = L.gram(X, T, S1)
    assert n == len(X) - 1
    Lf = L.fit_linear(G, R, n, 0.0)
    assert np.abs(Lf - Lt).max() < 1e-4
    p = np.zeros(L.NPARAM, np.float32)
    p[L.P_L:] = Lf.astype(np.float32).reshape(-1)
    th, w = L.targets(X, T, S1)
    y, _ = L.forward(p, X)
    assert np.abs((y - th) * w[:, None]).max() < 1e-4


def test_ridge_fit_is_stationary():
    """The ridge solution zeroes the gradient of
    sum_rows |t - L xn|^2 + rho n |L|^2, evaluated row by row."""
    rng = np.random.default_rng(12)
    X, T, S1, _ = _planted(rng, 300)
    T = T + rng.standard_normal(T.shape).astype(np.float32) * np.float32(0.05)
    G, R, n = L.gram(X, T, S1)
    rho = 1e-2
    Lf = L.fit_linear(G, R, n, rho)
    th, w = L.targets(X, T, S1)
    xn = L.normalised(X).astype(np.float64)
    grad = np.zeros_like(Lf)
    for b in range(len(X)):
        if w[b]:
            grad += -2.0 * np.outer(th[b] - Lf @ xn[b], xn[b])
    grad += 2.0 * rho * n * Lf
    assert np.abs(grad).max() < 1e-8 * max(1.0, np.abs(G).max())


def test_linear_path_is_scale_free_and_odd():
    rng = np.random.default_rng(13)
    X = rng.standard_normal((32, L.IN)).astype(np.float32)
    p = np.zeros(L.NPARAM, np.float32)
    p[L.P_L:] = rng.standard_normal(L.NPARAM - L.P_L).astype(np.float32)
    y, _ = L.forward(p, X)
    y4, _ = L.forward(p, X * np.float32(4.0))        # power-of-two scale: exact
    ym, _ = L.forward(p, -X)
    assert np.array_equal(y, y4)
    assert np.array_equal(y, -ym)


def test_predict_f64_pins():
    """The FP64 prediction (the FP64 error reference of the learned form)
    against the FP32 one: a zero network gives the baseline exactly in both;
    a network whose hidden branch is silenced (W3 = 0) differs only by the
    FP32 rounding of the linear path; a generic network by the bf16 rounding
    points of the hidden layers."""
    rng = np.random.default_rng(3)
    hist = [rng.standard_normal((64, L.NSAMP)).astype(np.float32) for _ in range(4)]
    z = np.zeros(L.NPARAM, np.float32)
    b = L.base(hist[1])
    assert np.array_equal(L.predict(z, hist), b)
    assert np.array_equal(L.predict_f64(z, hist), b.astype(np.float64))
    p = L.init_params(5)
    d = L.unpack(p)
    d["W3"][:] = 0.0
    d["L"][:] = rng.standard_normal(d["L"].shape).astype(np.float32) * 0.05
    q = L.pack(d)
    a32, a64 = L.predict(q, hist), L.predict_f64(q, hist)
    assert np.allclose(a32, a64, rtol=1e-5, atol=1e-6)
    assert not np.array_equal(a64, b.astype(np.float64))
    g32, g64 = L.predict(p, hist), L.predict_f64(p, hist)
    assert np.abs(g32 - g64).max() < 0.05 * np.abs(g64).max()
