"""Learned absorbing boundary -- CPU restatement (TEST INFRASTRUCTURE ONLY).

The reference has no learned boundary; this file states the exact semantics
that the HIP path (simulation_amd/csrc/lbc*.hip) implements, so the GPU
kernels have a checker:

* the canonical boundary sample of an edge position (27 values: ez, the
  normal-coupled H and the tangential-coupled H at 3 depths x 3 tangential
  offsets), the four edges mapped onto one frame by the exact symmetries of
  the Yee update (x-mirror, y-mirror, transpose);
* the network 108 -> 64 -> 64 -> 3 (GELU), with the GPU's rounding points
  (inputs and hidden activations rounded to bf16, weights bf16, products
  accumulated in FP32 on the GPU and FP64 here), and its backward pass;
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
NSAMP = 27
HIST = 4
NPARAM = HID * IN + HID + HID * HID + HID + OUT * HID + OUT


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
    """Flat FP32 parameter vector -> W1 (64,108), b1, W2 (64,64), b2, W3 (3,64), b3."""
    p = np.asarray(p, dtype=np.float32)
    o = 0
    out = {}
    for name, shape in (("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
                        ("W3", (OUT, HID)), ("b3", (OUT,))):
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    assert o == NPARAM
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[k], np.float32).reshape(-1)
                           for k in ("W1", "b1", "W2", "b2", "W3", "b3")])


def init_params(seed: int = 0) -> np.ndarray:
    """Scaled-uniform initial weights (fan-in), zero biases, last layer small."""
    rng = np.random.default_rng(seed)
    w1 = rng.uniform(-1, 1, (HID, IN)) * math.sqrt(3.0 / IN)
    w2 = rng.uniform(-1, 1, (HID, HID)) * math.sqrt(3.0 / HID)
    w3 = rng.uniform(-1, 1, (OUT, HID)) * math.sqrt(3.0 / HID) * 0.1
    return pack(dict(W1=w1, b1=np.zeros(HID), W2=w2, b2=np.zeros(HID), W3=w3, b3=np.zeros(OUT)))


# ------------------------------------------------------------------ GELU

_erf = np.vectorize(math.erf, otypes=[np.float64])


def gelu(z):
    z = np.asarray(z, np.float64)
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    z = np.asarray(z, np.float64)
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi)


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


def forward(params, X, rounding: bool = True):
    """Network on raw features X (B,108) -> (y (B,3) FP64, cache).  With
    rounding=True the bf16 rounding points of the kernels are applied (FP64
    sums in between); with False everything is FP64 (the exact function of
    the FP32 master weights)."""
    P = unpack(params)
    s, inv, xh = scale(X)
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    if not rounding:
        xh = (np.asarray(X, np.float64) * inv[:, None].astype(np.float64))
    W1, W2, W3 = (r(P[k]).astype(np.float64) for k in ("W1", "W2", "W3"))
    xh = np.asarray(xh, np.float64)
    z1 = xh @ W1.T + P["b1"].astype(np.float64)
    a1 = r(gelu(z1).astype(np.float32) if rounding else gelu(z1)).astype(np.float64)
    z2 = a1 @ W2.T + P["b2"].astype(np.float64)
    a2 = r(gelu(z2).astype(np.float32) if rounding else gelu(z2)).astype(np.float64)
    y = a2 @ W3.T + P["b3"].astype(np.float64)
    return y, dict(s=s, inv=inv, xh=xh, z1=z1, a1=a1, z2=z2, a2=a2, W1=W1, W2=W2, W3=W3)


def targets(X, T):
    """Normalised targets (T - ez_now) / s for raw features X and next-step
    ez T (B,3), FP32 as the kernels compute them, and the sample weights
    (0 for rows with s <= S_MIN)."""
    X = np.asarray(X, np.float32)
    s, inv, _ = scale(X)
    ez0 = X[:, [sidx(0, d, 0) for d in range(OUT)]]
    th = ((np.asarray(T, np.float32) - ez0) * inv[:, None]).astype(np.float32)
    w = (s > S_MIN).astype(np.float32)
    return th, w


def loss_and_grad(params, X, T, nrm: int | None = None, rounding: bool = True):
    """Mean-square loss sum_b w_b |y_b - th_b|^2 / nrm and its gradient with
    respect to the flat FP32 parameters (straight through the bf16 weight
    rounding).  rounding=True rounds the backward MFMA operands (dy, dz2, dz1)
    to bf16 as the kernels do."""
    y, c = forward(params, X, rounding)
    th, w = targets(X, T)
    B = len(X) if nrm is None else nrm
    e = (y - th) * w[:, None]
    loss = float(np.sum(e * e) / B)
    r = (lambda a: bf16(np.asarray(a, np.float32)).astype(np.float64)) if rounding else (lambda a: a)
    dy = 2.0 * e / B
    dyr = r(dy)
    gW3 = dyr.T @ c["a2"]
    gb3 = dy.sum(0)
    da2 = dyr @ c["W3"]
    dz2 = da2 * gelu_grad(c["z2"])
    dz2r = r(dz2)
    gW2 = dz2r.T @ c["a1"]
    gb2 = dz2.sum(0)
    da1 = dz2r @ c["W2"]
    dz1 = da1 * gelu_grad(c["z1"])
    dz1r = r(dz1)
    gW1 = dz1r.T @ c["xh"]
    gb1 = dz1.sum(0)
    g = np.concatenate([gW1.reshape(-1), gb1, gW2.reshape(-1), gb2, gW3.reshape(-1), gb3])
    return loss, g


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

def positions(nx: int, ny: int):
    """(edge, tau) of every boundary position in kernel order: top and bottom
    tau in [0, ny) (they own the corners), left and right tau in [3, nx-3)."""
    out = [(0, t) for t in range(ny)] + [(1, t) for t in range(ny)]
    out += [(2, t) for t in range(3, nx - 3)] + [(3, t) for t in range(3, nx - 3)]
    return out


def cell(e: int, d: int, t: int, nx: int, ny: int):
    """Grid cell of canonical depth d, tangential index t on edge e."""
    return [(d, t), (nx - 1 - d, t), (t, d), (t, ny - 1 - d)][e]


def sample(ez, hx, hy, nx: int, ny: int) -> np.ndarray:
    """Canonical 27-value samples (npos, 27) of the current fields.
    top:    ez[d,t],      hn =  hy[d,t],        ht =  hx[d,t]
    bottom: ez[nx-1-d,t], hn = -hy[nx-2-d,t],   ht =  hx[nx-1-d,t]
    left:   ez[t,d],      hn = -hx[t,d],        ht = -hy[t,d]
    right:  ez[t,ny-1-d], hn =  hx[t,ny-2-d],   ht = -hy[t,ny-1-d]
    ht at depth 0 is 0 on every edge (never updated on the bottom and right
    edges, never read by an update on the other two), and ht is taken only at
    offsets -1 and 0 -- the two H points beside the cell (ht is staggered by
    half a cell along the edge), so a prediction reaches one cell to either
    side; tangential indices are clamped to the edge."""
    pos = positions(nx, ny)
    out = np.zeros((len(pos), NSAMP), np.float32)
    for p, (e, t) in enumerate(pos):
        L = ny if e < 2 else nx
        for o in (-1, 0, 1):
            tt = min(max(t + o, 0), L - 1)
            for d in range(3):
                if e == 0:
                    v = (ez[d, tt], hy[d, tt], hx[d, tt])
                elif e == 1:
                    v = (ez[nx - 1 - d, tt], -hy[nx - 2 - d, tt], hx[nx - 1 - d, tt])
                elif e == 2:
                    v = (ez[tt, d], -hx[tt, d], -hy[tt, d])
                else:
                    v = (ez[tt, ny - 1 - d], hx[tt, ny - 2 - d], -hy[tt, ny - 1 - d])
                out[p, sidx(0, d, o)] = v[0]
                out[p, sidx(1, d, o)] = v[1]
                out[p, sidx(2, d, o)] = v[2] if (d > 0 and o < 1) else 0.0
    return out


def features(hist) -> np.ndarray:
    """108 features from the last four samples, newest first."""
    return np.concatenate(list(hist), axis=1).astype(np.float32)


def predict(params, hist):
    """Predicted next-step ez (npos, 3) in FP32: ez_now + s * y."""
    X = features(hist)
    y, c = forward(params, X, True)
    s = c["s"]
    ez0 = X[:, [sidx(0, d, 0) for d in range(OUT)]]
    yf = y.astype(np.float32)
    pred = (ez0 + (s[:, None] * yf).astype(np.float32)).astype(np.float32)
    small = s <= S_MIN
    pred[small] = ez0[small]
    return pred


def lbc_run(state: dict, naz, nbz, params, nsteps: int, hist=None, sources=(), src_wave=None,
            t0: int = 0):
    """Advance a learned-boundary grid (no PML) by nsteps in place.

    state {dz, iz, ez, hx, hy} FP32; hist: list of 4 samples (newest first)
    or None for a start from rest; returns the updated history."""
    dz, iz, ez, hx, hy = (state[k] for k in ("dz", "iz", "ez", "hx", "hy"))
    nx, ny = ez.shape
    pos = positions(nx, ny)
    if hist is None:
        z = np.zeros((len(pos), NSAMP), np.float32)
        hist = [z, z.copy(), z.copy(), z.copy()]
    hist = list(hist)
    for s in range(nsteps):
        pred = predict(params, hist)
        NO.dfield(nx, ny, None, dz, hx, hy)
        for k, (i, j) in enumerate(sources):
            dz[i, j] = src_wave[k][t0 + s]
        NO.efield(naz, nbz, dz, iz, ez)
        for p, (e, t) in enumerate(pos):
            for d in range(3):
                ez[cell(e, d, t, nx, ny)] = pred[p, d]
        NO.hfield(nx, ny, None, ez, None, None, hx, hy)
        hist = [sample(ez, hx, hy, nx, ny)] + hist[:3]
    return hist
