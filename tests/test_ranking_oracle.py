"""Pins the CPU oracles of the learned pair ranking (features, solo gains,
BF16 rounding, the MLP's forward / gradient / Adam) without a GPU."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import pdfgrade as pg
from oracle import brute
from oracle import mlp as omlp
from oracle import oracle as orc
from oracle import ranking as orank
from helpers import netlist_lists, pack, small_config


# ------------------------------------------------------------------ BF16

def test_bf16_round_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 20000),
                        np.array([0.0, -0.0, 1.0, -1.0, 3.4e38, 1e-40], np.float32)])
    # exact ties (low half 0x8000) with even and odd upper halves
    ties = (rng.integers(0x3F00, 0x4100, 2000).astype(np.uint32) << 16 | 0x8000).view(np.float32)
    x = np.concatenate([x, ties]).astype(np.float32)
    ours = orank.bf16_round(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    np.testing.assert_array_equal(ours, ref)


# ---------------------------------------------------------------- features

def _feat(nl, a_bits, P, max_level, b_bits=None, robust=True):
    net = orc.OracleNet(nl.arrays())
    codes = orc.sim_chunk(net, a_bits, P, 0, P, b_bits=b_bits)
    _, fwd = orc.mark_chunk(net, codes, robust)
    return orank.features(nl.arrays(), codes, fwd, max_level)


def test_features_closed_forms():
    nl = pg.Netlist.generate(small_config(7, num_pi=6, num_gates=30, levels=5))
    arrs = nl.arrays()
    P = 128
    rng = np.random.default_rng(1)
    vals = rng.integers(0, 2, (P, 6)).astype(np.uint8)
    a = pack(vals)
    # stable pairs: every feature is 0
    f = _feat(nl, a, P, 10, b_bits=a)
    assert not f.any()
    # every PI toggles: PI buckets that hold a PI read 1.0, as do level 0's
    # transitions and forward marks
    f = _feat(nl, a, P, 10, b_bits=pack(1 - vals))
    one = orank.bf16_round(np.float32([1.0]))[0]
    assert (f[:, 0] == one).all() and (f[:, 128] == one).all()
    for b in range(orank.PI_BUCKETS):
        lo, hi = orank.pi_bucket_bounds(6, b)
        expect = one if hi > lo else 0
        assert (f[:, 256 + b] == expect).all()
        assert (f[:, 384 + b] == expect).all()   # every generated PI has a fanout
    # buckets partition the PIs
    assert sum(orank.pi_bucket_bounds(6, b)[1] - orank.pi_bucket_bounds(6, b)[0] for b in range(128)) == 6
    assert sum(orank.pi_bucket_bounds(300, b)[1] - orank.pi_bucket_bounds(300, b)[0] for b in range(128)) == 300


@pytest.mark.parametrize("max_level,robust", [(0, True), (2, False), (99, True)])
def test_features_match_python_loop(max_level, robust):
    """The numpy features against plain loops over brute-force frames and
    forward marks."""
    nl = pg.Netlist.generate(small_config(11, num_pi=9, num_gates=40, levels=6))
    arrs = nl.arrays()
    types, fanins, po = netlist_lists(arrs)
    P, I = 64, 9
    rng = np.random.default_rng(2)
    va = rng.integers(0, 2, (P, I)).astype(np.uint8)
    vb = rng.integers(0, 2, (P, I)).astype(np.uint8)
    got = _feat(nl, pack(va), P, max_level, b_bits=pack(vb), robust=robust)
    fanouts = [[] for _ in range(arrs["n"])]
    for g in range(arrs["n"]):
        for f in fanins[g]:
            fanouts[f].append(g)
    level = arrs["level"]
    L = int(level.max())
    nfo = [0] * arrs["n"]
    for g in range(arrs["n"]):
        for f in fanins[g]:
            nfo[f] += 1
    for p in range(P):
        fr = brute.frames(types, fanins, I, va[p].tolist(), vb[p].tolist())
        t = [a ^ b for a, b in fr]
        fm = brute.marks_for_pattern(types, fanins, fanouts, po, fr, robust, remove=False)
        want = np.zeros(512, np.float32)
        for f in range(min(max_level, L) + 1):
            nodes = [g for g in range(arrs["n"]) if level[g] == f]
            inv = np.float32(1) / np.float32(len(nodes))
            want[f] = np.float32(sum(t[g] for g in nodes)) * inv
            want[128 + f] = np.float32(sum(fm[g] for g in nodes)) * inv
        for b in range(128):
            members = [i for i in range(I) if b * I // 128 <= i < (b + 1) * I // 128]
            if members:
                want[256 + b] = np.float32(sum(t[i] for i in members)) * (np.float32(1) / np.float32(len(members)))
                wsum = sum(nfo[i] for i in members)
                want[384 + b] = np.float32(sum(nfo[i] * t[i] for i in members)) * (np.float32(1) / np.float32(wsum))
        np.testing.assert_array_equal(got[p], orank.bf16_round(want), err_msg=f"pair {p}")


# ------------------------------------------------------------------- gains

def _brute_gains(nl, va, vb, robust):
    arrs = nl.arrays()
    types, fanins, po = netlist_lists(arrs)
    n = arrs["n"]
    fanouts = [[] for _ in range(n)]
    for g in range(n):
        for f in fanins[g]:
            fanouts[f].append(g)
    sp = brute.suffix_paths(fanouts, po)
    out = []
    for p in range(len(va)):
        fr = brute.frames(types, fanins, arrs["num_pi"], va[p], vb[p])
        m = brute.marks_for_pattern(types, fanins, fanouts, po, fr, robust)
        out.append(sum(1 for g in range(arrs["num_pi"]) for path in sp[g] if all(m[v] for v in path)))
    return np.array(out, np.uint32)


@pytest.mark.parametrize("seed,robust", [(3, True), (4, False), (5, True)])
def test_pair_gains_match_path_enumeration(seed, robust):
    nl = pg.Netlist.generate(small_config(seed, num_pi=6, num_gates=26, levels=5))
    P = 96
    rng = np.random.default_rng(seed)
    va = rng.integers(0, 2, (P, 6)).astype(np.uint8)
    vb = va.copy()
    flip = rng.random((P, 6)) < 0.4
    vb[flip] ^= 1
    got = orank.pair_gains(nl.arrays(), pack(va), P, robust, b_bits=pack(vb))
    want = _brute_gains(nl, va.tolist(), vb.tolist(), robust)
    np.testing.assert_array_equal(got, want)
    assert want.any()
    # a pair's solo gain bounds what it adds in the in-order grading
    seq = brute.grade(*netlist_lists(nl.arrays())[:2], netlist_lists(nl.arrays())[2], 6, va.tolist(),
                      vb.tolist(), robust)
    assert all(s <= g for s, g in zip(seq["per_pattern"], want))


def test_pair_gains_planted_mistake_fails():
    """A gain counted with the detection history of the in-order grading (new =
    first detected at p) instead of an empty history differs from the path
    enumeration: the check can fail."""
    nl = pg.Netlist.generate(small_config(3, num_pi=6, num_gates=26, levels=5))
    P = 96
    rng = np.random.default_rng(3)
    va = rng.integers(0, 2, (P, 6)).astype(np.uint8)
    vb = va.copy()
    vb[rng.random((P, 6)) < 0.4] ^= 1
    want = _brute_gains(nl, va.tolist(), vb.tolist(), True)
    res = orc.grade(orc.OracleNet(nl.arrays()), pack(va), P, P, True, b_bits=pack(vb))
    net = orc.OracleNet(nl.arrays())
    codes = orc.sim_chunk(net, pack(va), P, 0, P, b_bits=pack(vb))
    marks, _ = orc.mark_chunk(net, codes, True)
    wrong = []
    for p in range(P):
        tot, _ = orc.cover_chunk(net, np.ascontiguousarray(marks[:, p:p + 1]), p, res["first_t0"].copy(),
                                 res["first_t1"].copy(), np.zeros(net.n, np.uint64))
        wrong.append(tot)
    assert not np.array_equal(np.array(wrong, np.uint32), want)


# --------------------------------------------------------------------- MLP

def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    parts = dict(W1=rng.uniform(-1, 1, (256, 512)) * np.sqrt(6 / 512) * scale,
                 b1=rng.uniform(-0.1, 0.1, 256), W2=rng.uniform(-1, 1, (256, 256)) * np.sqrt(6 / 256) * scale,
                 b2=rng.uniform(-0.1, 0.1, 256), w3=rng.uniform(-1, 1, 256) * np.sqrt(6 / 256),
                 b3=np.array([0.3]))
    return omlp.flatten(parts)


def _torch_net(flat):
    p = omlp.unflatten(flat)
    net = torch.nn.Sequential(torch.nn.Linear(512, 256), torch.nn.ReLU(), torch.nn.Linear(256, 256),
                              torch.nn.ReLU(), torch.nn.Linear(256, 1)).double()
    with torch.no_grad():
        net[0].weight.copy_(torch.from_numpy(p["W1"]))
        net[0].bias.copy_(torch.from_numpy(p["b1"]))
        net[2].weight.copy_(torch.from_numpy(p["W2"]))
        net[2].bias.copy_(torch.from_numpy(p["b2"]))
        net[4].weight.copy_(torch.from_numpy(p["w3"][None, :]))
        net[4].bias.copy_(torch.from_numpy(p["b3"]))
    return net


def test_mlp_layout_size():
    assert omlp.NUM_PARAMS == pg.MLP_PARAMS == 197377


def test_mlp_forward_and_gradient_match_torch():
    flat = _params(0)
    rng = np.random.default_rng(1)
    X = rng.random((40, 512))
    t = rng.standard_normal(40)
    y, _ = omlp.forward(flat, X, quant=False)
    net = _torch_net(flat)
    xt = torch.from_numpy(X)
    yt = net(xt)[:, 0]
    np.testing.assert_allclose(y, yt.detach().numpy(), rtol=1e-12, atol=1e-12)
    loss = torch.mean((yt - torch.from_numpy(t)) ** 2)
    loss.backward()
    tg = omlp.flatten(dict(W1=net[0].weight.grad.numpy(), b1=net[0].bias.grad.numpy(),
                           W2=net[2].weight.grad.numpy(), b2=net[2].bias.grad.numpy(),
                           w3=net[4].weight.grad.numpy()[0], b3=net[4].bias.grad.numpy()))
    l2, g = omlp.loss_and_grad(flat, X, t, quant=False)
    assert abs(l2 - loss.item()) < 1e-12
    np.testing.assert_allclose(g, tg, rtol=1e-9, atol=1e-12)


def test_mlp_gradient_finite_differences():
    flat = _params(2)
    rng = np.random.default_rng(3)
    X = rng.random((16, 512))
    t = rng.standard_normal(16)
    _, g = omlp.loss_and_grad(flat, X, t, quant=False)
    h = 1e-6
    idx = np.concatenate([rng.choice(n, 12, replace=False) + off for off, n in
                          ((0, 256 * 512), (131072, 256), (131328, 65536), (196864, 256), (197120, 256))]
                         + [np.array([197376])])
    for i in idx:
        e = np.zeros_like(flat)
        e[i] = h
        fd = (omlp.loss_and_grad(flat + e, X, t, quant=False)[0] -
              omlp.loss_and_grad(flat - e, X, t, quant=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_mlp_quantized_close_to_exact():
    flat = _params(4)
    rng = np.random.default_rng(5)
    X = orank.bf16_to_f32(orank.bf16_round(rng.random((64, 512)).astype(np.float32)))
    t = rng.standard_normal(64)
    lq, gq = omlp.loss_and_grad(flat, X, t, quant=True)
    le, ge = omlp.loss_and_grad(flat, X, t, quant=False)
    assert abs(lq - le) <= 0.02 * le
    # BF16 rounding errors of single rows partly survive the cancellation of
    # the batch sum: compare directions per tensor
    a, b = omlp.unflatten(gq), omlp.unflatten(ge)
    for k in a:
        cos = float(np.dot(a[k].ravel(), b[k].ravel()) / (np.linalg.norm(a[k]) * np.linalg.norm(b[k])))
        assert cos > 0.99, (k, cos)


def test_adam_matches_torch():
    rng = np.random.default_rng(6)
    w = rng.standard_normal(1000).astype(np.float32)
    grads = [rng.standard_normal(1000).astype(np.float32) for _ in range(3)]
    p = torch.nn.Parameter(torch.from_numpy(w.copy()))
    opt = torch.optim.Adam([p], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    ours = w.copy()
    for step, g in enumerate(grads, 1):
        p.grad = torch.from_numpy(g.copy())
        opt.step()
        ours, m, v = omlp.adam_step(ours, g, m, v, step, 1e-2)
    np.testing.assert_allclose(ours, p.detach().numpy(), rtol=2e-6, atol=1e-7)


def test_gradient_check_detects_planted_mistake():
    """A backward pass without the ReLU mask of H1 fails the finite-difference
    check: the check can fail."""
    flat = _params(7)
    rng = np.random.default_rng(8)
    X = rng.random((8, 512))
    t = rng.standard_normal(8)

    def wrong_grad(flat):
        y, c = omlp.forward(flat, X, quant=False)
        dy = 2.0 * (y - t) / len(t)
        dZ2 = dy[:, None] * c["p"]["w3"][None, :] * (c["H2"] > 0)
        dZ1 = dZ2 @ c["W2"]          # mask of H1 dropped
        return dZ1.T @ c["X"]

    gw = wrong_grad(flat)
    h = 1e-6
    bad = 0
    for i in range(0, 256 * 512, 4099):
        e = np.zeros_like(flat)
        e[i] = h
        fd = (omlp.loss_and_grad(flat + e, X, t, quant=False)[0] -
              omlp.loss_and_grad(flat - e, X, t, quant=False)[0]) / (2 * h)
        bad += abs(fd - gw.reshape(-1)[i]) > 1e-6 + 1e-5 * abs(fd)
    assert bad > 0
