"""Learned absorbing boundary -- CPU restatement (TEST INFRASTRUCTURE ONLY).

The reference has no learned boundary; this file states the exact semantics
that the HIP path (simulation_amd/csrc/lbc*.hip) implements, so the GPU
kernels have a checker:

* the canonical boundary sample of an edge position (27 features: ez, the
  normal-coupled H and the tangential-coupled H at grid depths 3..5 x 3
  tangential offsets; plus the ez of depths 2, 1 and 0, targets only), the four
  edges mapped onto one frame by the exact symmetries of the Yee update
  (x-mirror, y-mirror, transpose);
* the baseline (the reference's two-step absorbing boundary,
  fd1d/pycuda/fd1d_1_2.py:53-61, on the three outer layers) that the network
  corrects;
* the network y = L xf + (f(xh) - f(-xh))/2: a linear path on the FP32
  normalised inputs xf plus the odd part of f: 108 -> 64 -> 64 -> 3 (GELU in
  the sigmoid form z sigma(1.702 z)) on their bf16 rounding xh, with the
  GPU's rounding points (inputs and hidden activations rounded to bf16,
  weights bf16, products accumulated in FP32 on the GPU and FP64 here), and
  its backward pass (hidden layers only);
* the least-squares fit of the linear path (normal equations with a ridge
  term, FP64);
* one learned-boundary time step: the free-space/lossy update of
  fd2d/program/fd2d_3_3.py:68-98 (as oracle/numpy_oracle.py) without a PML,
  with ez of the outer three cells of every edge replaced by the network's
  prediction between the E and H updates.

Only tests/ import it; the product path never does.
"""
from __future__ import annotations

import math

import numpy as np

from oracle import numpy_oracle as NO

IN, HID, OUT = 108, 64, 3
S_MIN = np.float32(1e-30)      # rows with max|x| <= S_MIN count as all-zero
NFEAT = 27                     # features per sample
NSAMP = 30                     # stored sample: features, then ez of grid depths 2, 1, 0
S_EZ2, S_EZ1, S_EZ0 = 27, 28, 29
FD0 = 3                        # grid depth of feature depth 0
HIST = 4
NPARAM = HID * IN + HID + HID * HID + HID + OUT * HID + OUT + OUT * IN
P_W3 = HID * IN + HID + HID * HID + HID      # offsets of W3 and of the linear path L
P_L = P_W3 + OUT * HID + OUT


def sidx(f: int, d: int, o: int) -> int:
    """Index in the 27-value sample of field f (0 ez, 1 normal-coupled H,
    2 tangential-coupled H), depth d (0 = outermost), tangential offset o."""
    return (f * 3 + d) * 3 + (o + 1)


# ------------------------------------------------------------------ bf16

def bf16(x) -> np.ndarray:
    """Round FP32 values to bfloat16 (nearest, ties to even), kept as FP32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


# ------------------------------------------------------------------ params

def unpack(p: np.ndarray) -> dict:
    """Flat FP32 parameter vector -> W1 (64,108), b1, W2 (64,64), b2, W3 (3,64),
    b3, L (3,108) (the linear path from the inputs to the outputs)."""
    p = np.asarray(p, dtype=np.float32)
    o = 0
    out = {}
    for name, shape in (("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
                        ("W3", (OUT, HID)), ("b3", (OUT,)), ("L", (OUT, IN))):
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    assert o == NPARAM
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[k], np.float32).reshape(-1)
                           for k in ("W1", "b1", "W2", "b2", "W3", "b3", "L")])


def init_params(seed: int = 0) -> np.ndarray:
    """Scaled-uniform initial weights (fan-in), zero biases, last layer small."""
    rng = np.random.default_rng(seed)
    w1 = rng.uniform(-1, 1, (HID, IN)) * math.sqrt(3.0 / IN)
    w2 = rng.uniform(-1, 1, (HID, HID)) * math.sqrt(3.0 / HID)
    w3 = rng.uniform(-1, 1, (OUT, HID)) * math.sqrt(3.0 / HID) * 0.1
    return pack(dict(W1=w1, b1=np.zeros(HID), W2=w2, b2=np.zeros(HID), W3=w3, b3=np.zeros(OUT),
                     L=np.zeros((OUT, IN))))


# ------------------------------------------------------------------ GELU

GELU_A = 1.702


def _sig(u):
    return 1.0 / (1.0 + np.exp(-u))


def gelu(z):
    """GELU in the sigmoid form of Hendrycks & Gimpel (2016): z sigma(1.702 z)."""
    z = np.asarray(z, np.float64)
    return z * _sig(GELU_A * z)


def gelu_grad(z):
    z = np.asarray(z, np.float64)
    s = _sig(GELU_A * z)
    return s * (1.0 + GELU_A * z * (1.0 - s))


# ------------------------------------------------------------------ network

def scale(X: np.ndarray):
    """Per-row scale s = max|x| and normalised bf16 inputs x*(1/s) (FP32
    reciprocal and product, as the kernels compute them); rows with
    s <= S_MIN give zero inputs."""
    X = np.asarray(X, np.float32)
    s = np.max(np.abs(X), axis=1).astype(np.float32)
    big = s > S_MIN
    inv = np.where(big, np.float32(1.0) / np.where(big, s, np.float32(1.0)), np.float32(0.0))
    inv = inv.astype(np.float32)
    xh = bf16((X * inv[:, None]).astype(np.float32))
    return s, inv, xh


def normalised(X, rounding: bool = True) -> np.ndarray:
    """FP32 normalised inputs x * (1/s) (the linear path's operand), FP64
    products with rounding=False."""
    s, inv, _ = scale(X)
    if rounding:
        return (np.asarray(X, np.float32) * inv[:, None]).astype(np.float32)
    return np.asarray(X, np.float64) * inv[:, None].astype(np.float64)


def _branch(P, xh, r, rounding):
    W1, W2, W3 = (r(P[k]).astype(np.float64) for k in ("W1", "W2", "W3"))
    z1 = xh @ W1.T + P["b1"].astype(np.float64)
    a1 = r(gelu(z1).astype(np.float32) if rounding else gelu(z1)).astype(np.float64)
    z2 = a1 @ W2.T + P["b2"].astype(np.float64)
    a2 = r(gelu(z2).astype(np.float32) if rounding else gelu(z2)).astype(np.float64)
    y = a2 @ W3.T + P["b3"].astype(np.float64)
    return y, dict(xh=xh, z1=z1, a1=a1, z2=z2, a2=a2, W1=W1, W2=W2, W3=W3)


def forward(params, X, rounding: bool = True):
    """Network on raw features X (B,108) -> (y (B,3) FP64, cache).

    The network is odd, y(-x) = -y(x), as the exact boundary map is linear:
    y = L xf + (f(xh) - f(-xh)) / 2 with f the 108-64-64-3 GELU network (a
    constant or even part would pump a one-signed field into the grid through
    the scale s).  L acts on the FP32 normalised inputs with FP32 weights.
    With rounding=True the bf16 rounding points of the kernels are applied
    (FP64 sums in between; bf16(-x) = -bf16(x)); with False everything is
    FP64 (the exact function of the FP32 master weights)."""
    P = unpack(params)
    s, inv, xh = scale(X)
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    if not rounding:
        xh = (np.asarray(X, np.float64) * inv[:, None].astype(np.float64))
    xh = np.asarray(xh, np.float64)
    yp, cp = _branch(P, xh, r, rounding)
    ym, cm = _branch(P, -xh, r, rounding)
    lin = normalised(X, rounding).astype(np.float64) @ P["L"].astype(np.float64).T
    return lin + 0.5 * (yp - ym), dict(s=s, inv=inv, plus=cp, minus=cm)


def base(S1) -> np.ndarray:
    """Baseline next-step ez (B,3) of grid depths 0..2, FP32, from the full
    samples S1 (B,30) of age 1 (two steps before the predicted one): the
    two-step absorbing boundary of fd1d/pycuda/fd1d_1_2.py:53-61 on the three
    outer layers, ez_d(t+1) = ez_{d+1}(t-1)."""
    S1 = np.asarray(S1, np.float32)
    return np.stack([S1[:, S_EZ1], S1[:, S_EZ2], S1[:, sidx(0, 0, 0)]], axis=1)


def targets(X, T, S1):
    """Normalised targets (T - base) / s for features X (B,108), next-step
    ez T (B,3) of grid depths 0..2 and age-1 samples S1 (B,30), FP32 as the
    kernels compute them, and the sample weights (0 for rows with
    s <= S_MIN)."""
    X = np.asarray(X, np.float32)
    s, inv, _ = scale(X)
    th = ((np.asarray(T, np.float32) - base(S1)) * inv[:, None]).astype(np.float32)
    w = (s > S_MIN).astype(np.float32)
    return th, w


def loss_and_grad(params, X, T, S1=None, nrm: int | None = None, rounding: bool = True):
    """Mean-square loss sum_b w_b |y_b - th_b|^2 / nrm and its gradient with
    respect to the flat FP32 parameters (straight through the bf16 weight
    rounding).  S1: age-1 samples for the baseline (default: the age-1
    features padded with zero outer ez).  The odd network's gradient is the
    f(xh) branch's with dy/2 plus the f(-xh) branch's with -dy/2.  The linear
    path is fitted by least squares (fit_linear), not by gradient: its
    gradient entries are zero.  rounding=True rounds the backward MFMA
    operands (dy, dz2, dz1) to bf16 as the kernels do."""
    y, c = forward(params, X, rounding)
    if S1 is None:
        S1 = pad_samples(np.asarray(X, np.float32)[:, NFEAT:2 * NFEAT])
    th, w = targets(X, T, S1)
    B = len(X) if nrm is None else nrm
    e = (y - th) * w[:, None]
    loss = float(np.sum(e * e) / B)
    r = (lambda a: bf16(np.asarray(a, np.float32)).astype(np.float64)) if rounding else (lambda a: a)
    g = np.zeros(NPARAM)
    for sgn, cc in ((0.5, c["plus"]), (-0.5, c["minus"])):
        dy = sgn * 2.0 * e / B
        dyr = r(dy)
        gW3 = dyr.T @ cc["a2"]
        gb3 = dy.sum(0)
        da2 = dyr @ cc["W3"]
        dz2 = da2 * gelu_grad(cc["z2"])
        dz2r = r(dz2)
        gW2 = dz2r.T @ cc["a1"]
        gb2 = dz2.sum(0)
        da1 = dz2r @ cc["W2"]
        dz1 = da1 * gelu_grad(cc["z1"])
        dz1r = r(dz1)
        gW1 = dz1r.T @ cc["xh"]
        gb1 = dz1.sum(0)
        g += np.concatenate([gW1.reshape(-1), gb1, gW2.reshape(-1), gb2, gW3.reshape(-1), gb3,
                             np.zeros(OUT * IN)])
    return loss, g


def gram(X, T, S1, round_bf16: bool = False):
    """Normal-equation sums of the linear path over rows with s > S_MIN:
    G = sum xn xn^T (108 x 108), R = sum xn t^T (108 x 3), count; xn the
    FP32 normalised inputs (bf16-rounded with round_bf16), t the normalised
    residual targets (targets()); FP64 sums."""
    xn = normalised(X).astype(np.float32)
    if round_bf16:
        xn = bf16(xn)
    th, w = targets(X, T, S1)
    xn = xn.astype(np.float64) * w[:, None]
    th = th.astype(np.float64) * w[:, None]
    return xn.T @ xn, xn.T @ th, float(w.sum())


def fit_linear(G, R, count: float, rho: float) -> np.ndarray:
    """Ridge solution L (3 x 108) of (G + rho * count * I) L^T = R."""
    A = np.asarray(G, np.float64) + rho * count * np.eye(IN)
    return np.linalg.solve(A, np.asarray(R, np.float64)).T


def adam(params, m, v, g, step: int, lr: float, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update in FP32 (the kernel's arithmetic order)."""
    f = np.float32
    g = np.asarray(g, f)
    m[:] = f(b1) * m + (f(1) - f(b1)) * g
    v[:] = f(b2) * v + (f(1) - f(b2)) * g * g
    c1 = f(1.0 - b1 ** step)
    c2 = f(1.0 - b2 ** step)
    params[:] = params - f(lr) * (m / c1) / (np.sqrt(v / c2) + f(eps))


# ------------------------------------------------------------------ boundary samples

def pad_samples(F) -> np.ndarray:
    """(B,27) features -> (B,30) samples with zero outer ez."""
    F = np.asarray(F, np.float32)
    return np.concatenate([F, np.zeros((len(F), NSAMP - NFEAT), np.float32)], axis=1)


def positions(nx: int, ny: int):
    """(edge, tau) of every boundary position in kernel order: top and bottom
    tau in [0, ny) (they own the corners), left and right tau in [3, nx-3)."""
    out = [(0, t) for t in range(ny)] + [(1, t) for t in range(ny)]
    out += [(2, t) for t in range(3, nx - 3)] + [(3, t) for t in range(3, nx - 3)]
    return out


def cell(e: int, d: int, t: int, nx: int, ny: int):
    """Grid cell of canonical depth d, tangential index t on edge e."""
    return [(d, t), (nx - 1 - d, t), (t, d), (t, ny - 1 - d)][e]


def sample(ez, hx, hy, nx: int, ny: int, dtype=np.float32) -> np.ndarray:
    """Canonical samples (npos, 30) of the current fields: 27 features of
    grid depths D = 3, 4, 5 (feature depth d = D - 3; the cells just inside
    the three predicted ones, so the network never reads its own output):
    top:    ez[D,t],      hn =  hy[D,t],        ht =  hx[D,t]
    bottom: ez[nx-1-D,t], hn = -hy[nx-2-D,t],   ht =  hx[nx-1-D,t]
    left:   ez[t,D],      hn = -hx[t,D],        ht = -hy[t,D]
    right:  ez[t,ny-1-D], hn =  hx[t,ny-2-D],   ht = -hy[t,ny-1-D]
    ht is taken only at offsets -1 and 0 -- the two H points beside the cell
    (ht is staggered by half a cell along the edge), so a prediction reaches
    one cell to either side; H in the last row or column reads as +0,
    whatever its sign (the H update never runs there, so a window of a
    larger grid must not see it either); tangential indices are clamped to
    the edge.  Then ez of grid
    depths 2, 1, 0 at offset 0 (targets and baseline, not network inputs)."""
    pos = positions(nx, ny)
    out = np.zeros((len(pos), NSAMP), dtype)

    def H(a, i, j, sign=1):
        # the last row or column reads as +0 whatever the sign; any other
        # cell is negated as it is (a stored +0 gives -0)
        if i == nx - 1 or j == ny - 1:
            return dtype(0.0)
        return -a[i, j] if sign < 0 else a[i, j]

    def ezc(e, D, t):
        i, j = cell(e, D, t, nx, ny)
        return ez[i, j]
    for p, (e, t) in enumerate(pos):
        L = ny if e < 2 else nx
        for o in (-1, 0, 1):
            tt = min(max(t + o, 0), L - 1)
            for d in range(3):
                D = d + FD0
                if e == 0:
                    v = (H(hy, D, tt), H(hx, D, tt))
                elif e == 1:
                    v = (H(hy, nx - 2 - D, tt, -1), H(hx, nx - 1 - D, tt))
                elif e == 2:
                    v = (H(hx, tt, D, -1), H(hy, tt, D, -1))
                else:
                    v = (H(hx, tt, ny - 2 - D), H(hy, tt, ny - 1 - D, -1))
                out[p, sidx(0, d, o)] = ezc(e, D, tt)
                out[p, sidx(1, d, o)] = v[0]
                out[p, sidx(2, d, o)] = v[1] if o < 1 else 0.0
        out[p, S_EZ2] = ezc(e, 2, t)
        out[p, S_EZ1] = ezc(e, 1, t)
        out[p, S_EZ0] = ezc(e, 0, t)
    return out


def features(hist) -> np.ndarray:
    """108 features from the last four samples, newest first."""
    return np.concatenate([h[:, :NFEAT] for h in hist], axis=1).astype(np.float32)


def predict(params, hist):
    """Predicted next-step ez (npos, 3) of grid depths 0..2 in FP32:
    base + s * y."""
    X = features(hist)
    y, c = forward(params, X, True)
    s = c["s"]
    b = base(hist[1])
    yf = y.astype(np.float32)
    pred = (b + (s[:, None] * yf).astype(np.float32)).astype(np.float32)
    small = s <= S_MIN
    pred[small] = b[small]
    return pred


def predict_f64(params, hist):
    """The same prediction evaluated in FP64 throughout (no bf16 or FP32
    rounding points: the exact function of the FP32 master weights), for the
    error of the FP32 path: base + s * (L xf + (f(xf) - f(-xf))/2)."""
    X = np.concatenate([np.asarray(h[:, :NFEAT], np.float64) for h in hist], axis=1)
    s = np.max(np.abs(X), axis=1)
    big = s > float(S_MIN)
    inv = np.where(big, 1.0 / np.where(big, s, 1.0), 0.0)
    xf = X * inv[:, None]
    P = unpack(params)
    ident = lambda a: np.asarray(a, np.float64)  # noqa: E731
    yp, _ = _branch(P, xf, ident, False)
    ym, _ = _branch(P, -xf, ident, False)
    y = xf @ P["L"].astype(np.float64).T + 0.5 * (yp - ym)
    S1 = np.asarray(hist[1], np.float64)
    b = np.stack([S1[:, S_EZ1], S1[:, S_EZ2], S1[:, sidx(0, 0, 0)]], axis=1)
    pred = b + s[:, None] * y
    pred[~big] = b[~big]
    return pred


def lbc_run(state: dict, naz, nbz, params, nsteps: int, hist=None, sources=(), src_wave=None,
            t0: int = 0):
    """Advance a learned-boundary grid (no PML) by nsteps in place.

    state {dz, iz, ez, hx, hy} FP32 (or FP64: the whole step, samples and
    prediction in FP64, predict_f64); hist: list of 4 samples (newest first)
    or None for a start from rest; returns the updated history."""
    dz, iz, ez, hx, hy = (state[k] for k in ("dz", "iz", "ez", "hx", "hy"))
    nx, ny = ez.shape
    f64 = ez.dtype == np.float64
    dt = np.float64 if f64 else np.float32
    pos = positions(nx, ny)
    if hist is None:
        z = np.zeros((len(pos), NSAMP), dt)
        hist = [z, z.copy(), z.copy(), z.copy()]
    hist = list(hist)
    for s in range(nsteps):
        pred = predict_f64(params, hist) if f64 else predict(params, hist)
        NO.dfield(nx, ny, None, dz, hx, hy)
        for k, (i, j) in enumerate(sources):
            dz[i, j] = src_wave[k][t0 + s]
        NO.efield(naz, nbz, dz, iz, ez)
        for p, (e, t) in enumerate(pos):
            for d in range(3):
                ez[cell(e, d, t, nx, ny)] = pred[p, d]
        NO.hfield(nx, ny, None, ez, None, None, hx, hy)
        hist = [sample(ez, hx, hy, nx, ny, dt)] + hist[:3]
    return hist
