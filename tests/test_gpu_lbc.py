"""Learned-boundary kernels against the oracle (oracle/lbc_oracle.py).

Bars: boundary samples are copies and negations -- bit-exact.  Network
forward and gradients use the oracle's bf16 rounding points with FP64 sums;
the GPU sums bf16 products in FP32 in the MFMA's own order, so a hidden
activation can round to the neighbouring bf16 value now and then:
forward rel L2 <= 2e-4, gradient rel L2 <= 2e-3 per tensor (cosine >= 0.9995
against the unrounded FP64 gradient).  Adam
is the same FP32 arithmetic: within 1e-4 of the step size."""
import numpy as np
import pytest
import torch

from oracle import lbc_oracle as L
from conftest import same_bits

pytestmark = pytest.mark.gpu


def _rows(X, dev):
    """Rows of raw features as (src, offs, tstride): sample k of row r (27
    features and three zero outer ez) at src[120 r + 30 k]."""
    X = np.asarray(X, np.float32)
    S = np.zeros((len(X), 4, L.NSAMP), np.float32)
    S[:, :, :L.NFEAT] = X.reshape(len(X), 4, L.NFEAT)
    src = torch.from_numpy(S.ravel()).to(dev)
    offs = torch.arange(len(X), dtype=torch.int64, device=dev) * (4 * L.NSAMP)
    return src, offs, -L.NSAMP


def _problem(seed, B):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((B, L.IN)).astype(np.float32)
    X *= (10.0 ** rng.uniform(-6, 1, B)).astype(np.float32)[:, None]   # row scales 1e-6..10
    X[5] = 0.0
    T = X[:, [L.sidx(0, d, 0) for d in range(3)]] + (
        rng.standard_normal((B, 3)) * np.abs(X).max(1, keepdims=True) * 0.05).astype(np.float32)
    p = L.init_params(seed)
    p[L.P_W3:L.P_L] *= 8.0
    p[L.P_L:] = (rng.standard_normal(L.NPARAM - L.P_L) * 0.2).astype(np.float32)
    return p.astype(np.float32), X, T.astype(np.float32)


def test_sample_matches_oracle_bit_exact(gpu):
    from simulation_amd import learned
    rng = np.random.default_rng(2)
    F = [rng.standard_normal((53, 61)).astype(np.float32) for _ in range(3)]
    T = [torch.from_numpy(f).to(gpu) for f in F]
    got = learned.sample(*T).cpu().numpy()
    assert same_bits(got, L.sample(*F, 53, 61))
    win = (5, 42, 7, 50)
    got = learned.sample(*T, window=win).cpu().numpy()
    want = L.sample(*[np.ascontiguousarray(f[5:42, 7:50]) for f in F], 37, 43)
    assert same_bits(got, want)


def test_forward_matches_oracle(gpu):
    from simulation_amd import learned
    p, X, _ = _problem(1, 1000)
    net = learned.Net(p, gpu)
    y, s = net.forward(*_rows(X, gpu))
    y, s = y.cpu().numpy(), s.cpu().numpy()
    yo, c = L.forward(p, X, rounding=True)
    assert same_bits(s, c["s"])
    rel = np.linalg.norm(y - yo) / np.linalg.norm(yo)
    print(f"[lbc forward] rel L2 {rel:.2e}, max abs {np.abs(y - yo).max():.2e}", flush=True)
    assert rel < 2e-4
    assert not np.any(y[5])              # zero row: the odd network gives exactly 0


def test_gradient_matches_oracle(gpu):
    from simulation_amd import learned
    p, X, T = _problem(2, 1000)
    net = learned.Net(p, gpu)
    src, offs, ts = _rows(X, gpu)
    loss, g = net.grad(src, offs, ts, tgt=torch.from_numpy(T).to(gpu), nblocks=7)
    g = g.cpu().numpy().astype(np.float64)
    lo, go = L.loss_and_grad(p, X, T, rounding=True)
    _, ge = L.loss_and_grad(p, X, T, rounding=False)
    assert abs(loss - lo) <= 1e-3 * lo
    P = L.unpack(np.arange(L.NPARAM, dtype=np.float32))
    for name, idx in P.items():
        idx = idx.astype(np.int64).ravel()
        if name in ("b3", "L"):
            # the odd network does not depend on b3, and L is fitted by least
            # squares, not by gradient: both gradients are 0
            assert not np.any(go[idx]) and not np.any(g[idx])
            continue
        r = np.linalg.norm(g[idx] - go[idx]) / np.linalg.norm(go[idx])
        print(f"[lbc grad] {name}: rel L2 vs oracle {r:.2e}", flush=True)
        assert r < 2e-3, name
    cos = g @ ge / np.linalg.norm(g) / np.linalg.norm(ge)
    assert cos > 0.9995
    # deterministic: same blocks, same bits
    _, g2 = net.grad(src, offs, ts, tgt=torch.from_numpy(T).to(gpu), nblocks=7)
    assert same_bits(g2.cpu().numpy().astype(np.float64), g)


def test_next_sample_targets(gpu):
    """tgt=None takes the targets from the following sample (offset +tstride)."""
    from simulation_amd import learned
    rng = np.random.default_rng(4)
    nt, npos = 9, 50
    G = rng.standard_normal((nt, npos, L.NSAMP)).astype(np.float32)
    src = torch.from_numpy(G.ravel()).to(gpu)
    ts = npos * L.NSAMP
    rows = [(t, q) for t in range(3, nt - 1) for q in range(npos)]
    offs = torch.tensor([t * ts + q * L.NSAMP for t, q in rows], dtype=torch.int64, device=gpu)
    # the same rows as full samples side by side, targets given explicitly:
    # ez of grid depths 0, 1, 2 of the next sample
    S = np.stack([np.stack([G[t - k, q] for k in range(4)]) for t, q in rows])
    T = np.stack([G[t + 1, q, [L.S_EZ0, L.S_EZ1, L.S_EZ2]] for t, q in rows])
    src2 = torch.from_numpy(S.ravel()).to(gpu)
    offs2 = torch.arange(len(rows), dtype=torch.int64, device=gpu) * (4 * L.NSAMP)
    net = learned.Net(L.init_params(5), gpu)
    l1, g1 = net.grad(src, offs, ts, nblocks=3)
    l2, g2 = net.grad(src2, offs2, -L.NSAMP, tgt=torch.from_numpy(T).to(gpu), nblocks=3)
    assert l1 == l2
    assert torch.equal(g1, g2)


def test_adam_step_matches_oracle(gpu):
    from simulation_amd import learned
    p, X, T = _problem(3, 512)
    net = learned.Net(p, gpu)
    src, offs, ts = _rows(X, gpu)
    tg = torch.from_numpy(T).to(gpu)
    _, g = net.grad(src, offs, ts, tgt=tg, nblocks=4)
    net.train_step(src, offs, ts, 1e-3, tgt=tg, nblocks=4)
    po, m, v = p.copy(), np.zeros_like(p), np.zeros_like(p)
    L.adam(po, m, v, g.cpu().numpy(), 1, 1e-3)
    got = net.numpy()
    d = np.abs(got - po)
    print(f"[lbc adam] max |diff| {d.max():.2e} (step size 1e-3)", flush=True)
    # same FP32 operations; the GPU's sqrt/divide may differ in the last ulp,
    # which shows where an update nearly cancels a parameter: 1e-4 of a step
    assert d.max() <= 1e-7
    # the bf16 operand copies follow the parameters
    w = net.wbuf.cpu().numpy().view(np.uint16)
    W1 = L.unpack(got)["W1"]
    b = (w[:64 * 128].reshape(64, 128)[:, :108].astype(np.uint32) << 16).view(np.float32)
    assert same_bits(b, L.bf16(W1))


def _next_rows(X, S1, T, dev):
    """Rows laid out as [next sample, age 0 .. age 3] (150 floats each), the
    targets in the next sample, full age-1 samples S1: (src, offs, tstride)."""
    B = len(X)
    S = np.zeros((B, 5, L.NSAMP), np.float32)
    S[:, 1:, :L.NFEAT] = np.asarray(X, np.float32).reshape(B, 4, L.NFEAT)
    S[:, 2] = S1
    S[:, 0, [L.S_EZ0, L.S_EZ1, L.S_EZ2]] = T
    src = torch.from_numpy(S.ravel()).to(dev)
    offs = torch.arange(B, dtype=torch.int64, device=dev) * (5 * L.NSAMP) + L.NSAMP
    return src, offs, -L.NSAMP


def _planted(seed, B, noise):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((B, L.IN)).astype(np.float32)
    X *= (10.0 ** rng.uniform(-4, 1, B)).astype(np.float32)[:, None]
    X[9] = 0.0
    S1 = rng.standard_normal((B, L.NSAMP)).astype(np.float32)
    S1[:, :L.NFEAT] = X[:, L.NFEAT:2 * L.NFEAT]
    Lt = rng.standard_normal((L.OUT, L.IN)) * 0.3
    s, _, _ = L.scale(X)
    xn = L.normalised(X).astype(np.float64)
    T = L.base(S1) + s[:, None] * (xn @ Lt.T + noise * rng.standard_normal((B, L.OUT)))
    return X, S1, T.astype(np.float32), Lt


def test_gram_and_ridge_fit_match_oracle(gpu):
    """Normal-equation sums: the same FP32 inputs, exact FP64 products, FP64
    sums in another order -- rel 1e-12.  The FP64 Cholesky solve against the
    oracle's FP64 solve: rel 1e-9 before the FP32 store.  Noise-free planted
    targets are recovered to 1e-4."""
    from simulation_amd import _lib, learned
    X, S1, T, Lt = _planted(21, 3000, 0.05)
    src, offs, ts = _next_rows(X, S1, T, gpu)
    net = learned.Net(np.zeros(L.NPARAM, np.float32), gpu)
    part = net.gram(src, offs, ts, nblocks=5)
    Go, Ro, no = L.gram(X, T, S1)
    for rho in (0.0, 1e-3):
        g = net.fit_linear(part, rho).cpu().numpy()
        G = g[:L.IN * (L.IN + 3)].reshape(L.IN, L.IN + 3)
        assert g[-1] == no == len(X) - 1
        assert np.abs(G[:, :L.IN] - Go).max() <= 1e-12 * np.abs(Go).max()
        assert np.abs(G[:, L.IN:] - Ro).max() <= 1e-12 * np.abs(Ro).max()
        want = L.fit_linear(Go, Ro, no, rho)
        got = L.unpack(net.numpy())["L"].astype(np.float64)
        rel = np.abs(got - want).max() / np.abs(want).max()
        print(f"[lbc fit] rho {rho}: rel max {rel:.2e}", flush=True)
        assert rel < 1e-6
    X, S1, T, Lt = _planted(22, 2000, 0.0)
    net.fit_linear(net.gram(*_next_rows(X, S1, T, gpu)), 0.0)
    assert np.abs(L.unpack(net.numpy())["L"] - Lt).max() < 1e-4
    # singular system (no usable rows): refused, L unchanged
    before = net.numpy()
    Z = np.zeros_like(X[:64])
    with pytest.raises(_lib.FdtdError, match="pivot 0"):
        net.fit_linear(net.gram(*_next_rows(Z, S1[:64] * 0, T[:64] * 0, gpu)), 0.0)
    assert same_bits(net.numpy(), before)


def test_training_reduces_loss(gpu):
    """Linear path by least squares, then Adam on the hidden layers (linear
    path fixed) for an odd nonlinear remainder: the loss left by the linear
    fit drops 10x in 400 steps (the oracle's own Adam run drops it 500x)."""
    from simulation_amd import learned
    rng = np.random.default_rng(6)
    B = 8192
    X = (rng.standard_normal((B, L.IN)) * (10.0 ** rng.uniform(-3, 0, B))[:, None]).astype(np.float32)
    A = (rng.standard_normal((L.IN, 3)) * 0.05).astype(np.float32)
    Cm = (rng.standard_normal((L.IN, 3)) * 0.15).astype(np.float32)
    s = np.abs(X).max(1, keepdims=True)
    xn = X / s
    S1 = L.pad_samples(X[:, L.NFEAT:2 * L.NFEAT])
    T = (L.base(S1) + s * (xn @ A + np.tanh(xn @ Cm) ** 3)).astype(np.float32)
    p = L.init_params(0)
    p[L.P_L:] = 0.0
    net = learned.Net(p, gpu)
    net.fit_linear(net.gram(*_next_rows(X, S1, T, gpu)), 0.0)
    src, offs, ts = _rows(X, gpu)
    tg = torch.from_numpy(T).to(gpu)
    l0, _ = net.grad(src, offs, ts, tgt=tg)
    fitted = net.numpy()[L.P_L:]
    for _ in range(400):
        net.train_step(src, offs, ts, 3e-3, tgt=tg, only=(0, L.P_L))
    l1, _ = net.grad(src, offs, ts, tgt=tg)
    print(f"[lbc train] loss after the linear fit {l0:.3e} -> {l1:.3e}", flush=True)
    assert l1 < 0.1 * l0
    assert same_bits(net.numpy()[L.P_L:], fitted)


def _frame_operands(p):
    """Per-lane A operands of the time-stepping kernel's transposed network,
    restated from their definition: lane (c, g), op mt*4+ks -> W1[16mt+c]
    [27g + 8ks + v] (zero past 27), op 16+2mt+ks -> W2[16mt+c][u], op 24+ks ->
    W3[c][u] (zero for c >= 3), u = (2ks + v//4)*16 + 4g + v%4."""
    P = L.unpack(p)
    out = np.zeros((64, 26, 8), np.float32)
    for lane in range(64):
        c, g = lane % 16, lane // 16
        for v in range(8):
            for ks in range(4):
                i = 8 * ks + v
                for mt in range(4):
                    if i < L.NFEAT:
                        out[lane, mt * 4 + ks, v] = P["W1"][16 * mt + c, 27 * g + i]
            for ks in range(2):
                u = (2 * ks + v // 4) * 16 + 4 * g + v % 4
                for mt in range(4):
                    out[lane, 16 + 2 * mt + ks, v] = P["W2"][16 * mt + c, u]
                if c < L.OUT:
                    out[lane, 24 + ks, v] = P["W3"][c, u]
    return L.bf16(out).reshape(-1)


def test_frame_operand_layout(gpu):
    """The bf16 buffer's per-lane operands: written by lbc_pack, kept current
    by lbc_adam (the whole buffer equals a fresh pack after a step); the
    flag after them (8 elements) says whether W3/b3 hold a non-zero value."""
    from simulation_amd import learned
    p, X, T = _problem(9, 256)
    net = learned.Net(p, gpu)
    w = net.wbuf.cpu().numpy().view(np.uint16)
    nfrm = 64 * 26 * 8
    assert w.size == net.wbuf.numel() and w.size >= nfrm + 8
    got = (w[-nfrm - 8:-8].astype(np.uint32) << 16).view(np.float32)
    assert same_bits(got, _frame_operands(p))
    assert w[-8] == 1 and not w[-7:].any()
    src, offs, ts = _rows(X, gpu)
    net.train_step(src, offs, ts, 1e-2, tgt=torch.from_numpy(T).to(gpu), nblocks=3)
    fresh = learned.Net(net.numpy(), gpu)
    assert torch.equal(net.wbuf, fresh.wbuf)
    got = (net.wbuf.cpu().numpy().view(np.uint16)[-nfrm - 8:-8].astype(np.uint32) << 16).view(np.float32)
    assert same_bits(got, _frame_operands(net.numpy()))
    # a zero output layer: the flag is clear, and an Adam step that moves W3
    # off zero raises it
    q = p.copy()
    q[learned.P_W3:learned.P_L] = 0.0
    z = learned.Net(q, gpu)
    assert z.wbuf.cpu().numpy().view(np.uint16)[-8] == 0
    z.train_step(src, offs, ts, 1e-2, tgt=torch.from_numpy(T).to(gpu), nblocks=3)
    assert z.wbuf.cpu().numpy().view(np.uint16)[-8] == 1
