"""The proposer's HIP kernels against float32/float64 references of the same ops.

Inputs and weights are BF16 values (rounded on the host first), so the references see
exactly the operands the MFMA kernels see; what remains is FP32 accumulation order and
the final BF16 rounding of stored activations (relative 2^-8).
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest
from scipy.special import erf

pytestmark = pytest.mark.gpu

L, C = 4096, 64


def _torch():
    import torch

    return torch


def _bf(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(torch.bfloat16)


def _dev(t):
    return t.to("cuda")


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(None)


def _s():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _lib():
    from cuperiod_amd import _lib as lib

    return lib.load_library()


def _gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def _gelu_grad(z):
    cdf = 0.5 * (1.0 + erf(z / math.sqrt(2.0)))
    return cdf + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def _conv_ref(h, w):
    """z[b][x][o] = sum_{k,c} w[o][k][c] h[b][x+k-3][c] (float64)."""
    b = h.shape[0]
    pad = np.zeros((b, L + 6, h.shape[2]))
    pad[:, 3:L + 3] = h
    out = np.zeros((b, L, w.shape[0]))
    for k in range(7):
        out += pad[:, k:k + L] @ w[:, k, :].T
    return out


def _rng(seed):
    return np.random.default_rng(seed)


def test_conv_forward_epilogues():
    torch = _torch()
    lib = _lib()
    rng = _rng(1)
    n = 2
    h = _bf(rng.normal(0, 1, (n, L, C)))
    w = _bf(rng.normal(0, 0.05, (C, 7, C)))
    bias = torch.from_numpy(rng.normal(0, 0.1, C).astype(np.float32))
    zsrc = _bf(rng.normal(0, 1, (n, L, C)))
    ref_acc = _conv_ref(h.float().numpy().astype(np.float64), w.float().numpy().astype(np.float64))
    hd, wd, bd, zd = _dev(h), _dev(w), _dev(bias), _dev(zsrc)
    for epi in (0, 1, 2):
        out = torch.empty((n, L, C), dtype=torch.bfloat16, device="cuda")
        zs = torch.empty_like(out)
        st = lib.cpp_conv_forward(_s(), _p(hd), _p(wd), _p(bd), _p(zd), _p(zs), _p(out), n, epi)
        assert st == 0
        torch.cuda.synchronize()
        got = out.float().cpu().numpy()
        if epi == 0:
            z = ref_acc + bias.numpy()
            np.testing.assert_allclose(zs.float().cpu().numpy(), z, rtol=1e-2, atol=1e-2)
            ref = _gelu(z)
        elif epi == 1:
            ref = ref_acc * _gelu_grad(zsrc.float().numpy().astype(np.float64))
        else:
            ref = ref_acc + bias.numpy()
        np.testing.assert_allclose(got, ref, rtol=1e-2, atol=1e-2)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err < 8e-3, (epi, err)


def test_conv_weight_gradient_and_flip():
    torch = _torch()
    lib = _lib()
    rng = _rng(2)
    n = 2
    dz = _bf(rng.normal(0, 1, (n, L, C)))
    h = _bf(rng.normal(0, 1, (n, L, C)))
    dw = torch.zeros((C, 7, C), dtype=torch.float32, device="cuda")
    db = torch.zeros(C, dtype=torch.float32, device="cuda")
    dzd, hd = _dev(dz), _dev(h)  # held: a freed temporary could be reused before the launch
    assert lib.cpp_conv_wgrad(_s(), _p(dzd), _p(hd), n, _p(dw), _p(db)) == 0
    torch.cuda.synchronize()
    dzn = dz.float().numpy().astype(np.float64)
    hn = h.float().numpy().astype(np.float64)
    pad = np.zeros((n, L + 6, C))
    pad[:, 3:L + 3] = hn
    ref = np.stack([np.einsum("bxo,bxc->oc", dzn, pad[:, k:k + L]) for k in range(7)], axis=1)
    np.testing.assert_allclose(dw.cpu().numpy(), ref, rtol=1e-4, atol=1e-3)
    np.testing.assert_allclose(db.cpu().numpy(), dzn.sum(axis=(0, 1)), rtol=1e-4, atol=1e-3)
    w = _bf(rng.normal(0, 1, (C, 7, C)))
    wt = torch.empty((C, 7, C), dtype=torch.bfloat16, device="cuda")
    wd = _dev(w)
    assert lib.cpp_flip_weights(_s(), _p(wd), _p(wt)) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(wt.cpu().float().numpy(),
                                  w.float().numpy()[:, ::-1, :].transpose(2, 1, 0))


def test_layer1_forward_and_gradient():
    torch = _torch()
    lib = _lib()
    rng = _rng(3)
    n = 2
    x = rng.normal(0, 1, (n, L, 2)).astype(np.float32)
    w1 = rng.normal(0, 0.3, (C, 7, 2)).astype(np.float32)
    b1 = rng.normal(0, 0.1, C).astype(np.float32)
    xd, wd, bd = (_dev(torch.from_numpy(a)) for a in (x, w1, b1))
    z1 = torch.empty((n, L, C), dtype=torch.bfloat16, device="cuda")
    h1 = torch.empty_like(z1)
    assert lib.cpp_layer1_forward(_s(), _p(xd), _p(wd), _p(bd), n, _p(z1), _p(h1)) == 0
    torch.cuda.synchronize()
    z = _conv_ref(x.astype(np.float64), w1.astype(np.float64)) + b1
    np.testing.assert_allclose(z1.float().cpu().numpy(), z, rtol=8e-3, atol=8e-3)
    np.testing.assert_allclose(h1.float().cpu().numpy(), _gelu(z), rtol=8e-3, atol=8e-3)
    dz = _bf(rng.normal(0, 1, (n, L, C)))
    dw = torch.zeros((C, 7, 2), dtype=torch.float32, device="cuda")
    db = torch.zeros(C, dtype=torch.float32, device="cuda")
    dzd = _dev(dz)
    assert lib.cpp_layer1_wgrad(_s(), _p(dzd), _p(xd), n, _p(dw), _p(db)) == 0
    torch.cuda.synchronize()
    dzn = dz.float().numpy().astype(np.float64)
    pad = np.zeros((n, L + 6, 2))
    pad[:, 3:L + 3] = x
    ref = np.stack([np.einsum("bxo,bxc->oc", dzn, pad[:, k:k + L]) for k in range(7)], axis=1)
    np.testing.assert_allclose(dw.cpu().numpy(), ref, rtol=1e-4, atol=1e-3)
    np.testing.assert_allclose(db.cpu().numpy(), dzn.sum(axis=(0, 1)), rtol=1e-4, atol=1e-3)


def test_head_train_and_windows():
    torch = _torch()
    lib = _lib()
    rng = _rng(4)
    n = 3
    h4 = _bf(rng.normal(0, 1, (n, L, C)))
    z4 = _bf(rng.normal(0, 1, (n, L, C)))
    wh = rng.normal(0, 0.3, C).astype(np.float32)
    bh = np.array([0.2], dtype=np.float32)
    labels = np.array([[5, 5], [3990, 4010], [2000, 2100]], dtype=np.int32)  # [1] is clipped
    whd, bhd = _dev(torch.from_numpy(wh)), _dev(torch.from_numpy(bh))
    dz4 = torch.empty((n, L, C), dtype=torch.bfloat16, device="cuda")
    dwh = torch.zeros(C, dtype=torch.float32, device="cuda")
    dbh = torch.zeros(1, dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    scale = 1.0 / n
    h4d, z4d, labd = _dev(h4), _dev(z4), _dev(torch.from_numpy(labels))
    st = lib.cpp_head_train(_s(), _p(h4d), _p(z4d), _p(whd), _p(bhd), _p(labd), n, scale,
                            _p(dz4), _p(dwh), _p(dbh), _p(loss))
    assert st == 0
    torch.cuda.synchronize()
    hn = h4.float().numpy().astype(np.float64)
    zn = z4.float().numpy().astype(np.float64)
    logit = hn @ wh.astype(np.float64) + bh[0]
    p = np.exp(logit - logit.max(axis=1, keepdims=True))
    p /= p.sum(axis=1, keepdims=True)
    target = np.zeros_like(p)
    for b, (lo, hi) in enumerate(labels):
        hi = min(hi, L - 1)
        target[b, lo:hi + 1] = 1.0 / (hi - lo + 1)
    dl = scale * (p - target)
    ref_loss = -scale * (target * np.log(p)).sum()
    assert abs(loss.item() - ref_loss) < 1e-4 * max(1.0, abs(ref_loss))
    np.testing.assert_allclose(dwh.cpu().numpy(), np.einsum("bx,bxc->c", dl, hn), rtol=1e-3,
                               atol=1e-5)
    assert abs(dbh.item()) < 1e-5
    ref_dz = dl[:, :, None] * wh[None, None, :] * _gelu_grad(zn)
    np.testing.assert_allclose(dz4.float().cpu().numpy(), ref_dz, rtol=1e-2,
                               atol=1e-2 * np.abs(ref_dz).max())

    n_win, radius = 6, 2
    f_lo = np.array([0.01, 0.002, 0.5])
    dlog = np.log(2.0 / f_lo) / (L - 1)
    half = np.array([0.0, 0.002, 0.01])
    geo = [_dev(torch.from_numpy(a)) for a in (f_lo, dlog, half)]
    centre = torch.empty((n, n_win), dtype=torch.int32, device="cuda")
    mass = torch.empty((n, n_win), dtype=torch.float32, device="cuda")
    span = torch.empty((n, n_win, 2), dtype=torch.int32, device="cuda")
    assert lib.cpp_head_windows(_s(), _p(h4d), _p(whd), _p(bhd), *(_p(g) for g in geo), n,
                                n_win, radius, _p(centre), _p(mass), _p(span)) == 0
    torch.cuda.synchronize()
    for b in range(n):
        q = p[b].copy()
        avail = np.ones(L, bool)
        for w in range(n_win):
            c = int(np.argmax(np.where(avail, q, -1.0)))
            assert centre[b, w].item() == c
            fx = f_lo[b] * np.exp(c * dlog[b])
            lo, hi = c - radius, c + radius
            if fx - half[b] > f_lo[b]:
                lo = min(lo, int(np.ceil(np.log((fx - half[b]) / f_lo[b]) / dlog[b])))
            else:
                lo = min(lo, 0)
            hi = max(hi, int(np.floor(np.log((fx + half[b]) / f_lo[b]) / dlog[b])))
            lo, hi = max(lo, 0), min(hi, L - 1)
            assert abs(mass[b, w].item() - q[lo:hi + 1].sum()) < 1e-5
            assert span[b, w, 0].item() == lo and span[b, w, 1].item() == hi
            q[lo:hi + 1] = 0.0
            avail[lo:hi + 1] = False


def test_stacked_features():
    # chunks' coarse powers summed as squares per cell, standardised per curve; the
    # cadence histogram of the parent curve
    torch = _torch()
    lib = _lib()
    rng = _rng(9)
    n_chunks = 5
    power = rng.normal(3.0, 1.5, n_chunks * L + 7)
    power[11] = -np.inf
    power[L + 40] = np.nan
    power[2 * L + 9] = -2.0
    base = np.array([0, L, 2 * L + 3, 3 * L + 5, 4 * L + 7], dtype=np.int64)
    chunk_start = np.array([0, 3, 5], dtype=np.int32)
    tau = np.concatenate([np.sort(rng.uniform(0, 80.0, 400)), np.sort(rng.uniform(0, 9.0, 50))])
    pbase = np.array([0, 400], dtype=np.int64)
    npts = np.array([400, 50], dtype=np.int32)
    baseline = np.array([80.0, 9.0])
    x = torch.empty((2, L, 2), dtype=torch.float32, device="cuda")
    args = [_dev(torch.from_numpy(a)) for a in (power, chunk_start, base, tau, pbase, npts,
                                                  baseline)]
    assert lib.cpp_stack_features(_s(), *(_p(a) for a in args), 2, 30.0, _p(x)) == 0
    torch.cuda.synchronize()
    got = x.cpu().numpy()
    for c in range(2):
        stack = np.zeros(L)
        for k in range(chunk_start[c], chunk_start[c + 1]):
            row = power[base[k]:base[k] + L][::-1]
            ok = np.isfinite(row) & (row > 0)
            stack += np.where(ok, row, 0.0) ** 2
        stack32 = stack.astype(np.float32).astype(np.float64)
        sd = np.sqrt(np.mean((stack32 - stack.mean()) ** 2))
        want = np.clip((stack32 - stack.mean()) / sd, -30, 30)
        np.testing.assert_allclose(got[c, :, 0], want, rtol=1e-5, atol=1e-5)
        t = tau[pbase[c]:pbase[c] + npts[c]]
        hist = np.bincount(np.clip((t * (512 / baseline[c])).astype(int), 0, 511), minlength=512)
        np.testing.assert_allclose(got[c, :, 1], np.repeat(hist, 8) * 512 / npts[c], rtol=1e-6)


def test_features_and_adam():
    torch = _torch()
    lib = _lib()
    rng = _rng(5)
    n_out = 2 * L + 10
    power = rng.normal(5.0, 2.0, n_out)
    power[7] = np.inf
    base = np.array([3, L + 6], dtype=np.int64)
    summary = np.array([[5.0, 2.0, 0, 0], [4.0, 0.0, 0, 0]], dtype=np.float64)
    tau = np.concatenate([np.sort(rng.uniform(0, 10.0, 300)), np.sort(rng.uniform(0, 50.0, 77))])
    pbase = np.array([0, 300], dtype=np.int64)
    npts = np.array([300, 77], dtype=np.int32)
    baseline = np.array([10.0, 50.0])
    x = torch.empty((2, L, 2), dtype=torch.float32, device="cuda")
    args = [_dev(torch.from_numpy(a)) for a in (power, base, summary.ravel(), tau, pbase, npts,
                                                  baseline)]
    assert lib.cpp_features(_s(), *(_p(a) for a in args), 2, 30.0, _p(x)) == 0
    torch.cuda.synchronize()
    got = x.cpu().numpy()
    for c in range(2):
        row = power[base[c]:base[c] + L][::-1]
        row = np.where(np.isfinite(row), row, 0.0)
        sd = summary[c, 1]
        s0 = np.clip((row - summary[c, 0]) / sd, -30, 30) if sd > 0 else np.zeros(L)
        np.testing.assert_allclose(got[c, :, 0], s0, rtol=1e-6, atol=1e-6)
        t = tau[pbase[c]:pbase[c] + npts[c]]
        hist = np.bincount(np.clip((t * (512 / baseline[c])).astype(int), 0, 511), minlength=512)
        np.testing.assert_allclose(got[c, :, 1], np.repeat(hist * 512.0 / npts[c], 8), rtol=1e-6)

    k = 1000
    p = rng.normal(0, 1, k).astype(np.float32)
    g = rng.normal(0, 1, k).astype(np.float32)
    pd, gd = _dev(torch.from_numpy(p.copy())), _dev(torch.from_numpy(g.copy()))
    md, vd = torch.zeros(k, device="cuda"), torch.zeros(k, device="cuda")
    pb = torch.empty(k, dtype=torch.bfloat16, device="cuda")
    for t in (1, 2):
        gd.copy_(torch.from_numpy(g))
        assert lib.cpp_adam(_s(), _p(pd), _p(gd), _p(md), _p(vd), _p(pb), k, 1e-2, 0.9, 0.999,
                            1e-8, t) == 0
    torch.cuda.synchronize()
    m = v = np.zeros(k)
    q = p.astype(np.float64)
    for t in (1, 2):
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        q = q - 1e-2 * (m / (1 - 0.9**t)) / (np.sqrt(v / (1 - 0.999**t)) + 1e-8)
    np.testing.assert_allclose(pd.cpu().numpy(), q, rtol=1e-5, atol=1e-6)
    assert float(gd.abs().max()) == 0.0
    np.testing.assert_array_equal(pb.cpu().float().numpy(),
                                  torch.from_numpy(pd.cpu().numpy()).to(torch.bfloat16).float().numpy())


def test_training_reduces_loss():
    """A learnable toy task: the label is the position of the largest input bump."""
    torch = _torch()
    from cuperiod_amd.learned.network import ProposerNet, init_parameters

    rng = _rng(6)
    n = 32
    x = np.zeros((n, L, 2), dtype=np.float32)
    x[:, :, 0] = rng.normal(0, 0.3, (n, L))
    x[:, :, 1] = 1.0
    labels = rng.integers(10, L - 10, n).astype(np.int32)
    for b, lab in enumerate(labels):
        x[b, lab - 2:lab + 3, 0] += np.array([1.0, 3.0, 6.0, 3.0, 1.0])
    xd = _dev(torch.from_numpy(x))
    ld = _dev(torch.from_numpy(np.stack([labels, labels], axis=1).astype(np.int32)))
    net = ProposerNet(init_parameters(0))
    losses = []
    for _ in range(60):
        net.loss.zero_()
        net.train_step(xd, ld, 3e-3)
        torch.cuda.synchronize()
        losses.append(net.loss.item())
    assert all(np.isfinite(losses))
    assert losses[-1] < 0.5 * losses[0], losses[::10]
    geo = [_dev(torch.from_numpy(a)) for a in (np.full(n, 0.01), np.full(n, 0.001),
                                                np.zeros(n))]
    centre, mass, _ = net.windows(xd, 4, 1, *geo)
    torch.cuda.synchronize()
    hit = np.mean([labels[b] in set(centre[b].cpu().numpy().tolist()) for b in range(n)])
    assert hit > 0.5
