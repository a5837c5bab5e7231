"""GPU parity of the learned pair ranking: features, solo gains, ranking,
coverage curve and the MFMA MLP, each against its CPU oracle."""
from __future__ import annotations

import numpy as np
import pytest

import pdfgrade as pg
from oracle import mlp as omlp
from oracle import oracle as orc
from oracle import ranking as orank
from helpers import small_config

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = "cuda:0"


def _say(*a):
    print(*a, flush=True)


def _explicit(num_pi, P, seed):
    return pg.gen_vectors(seed, num_pi, P), pg.gen_vectors(seed + 1, num_pi, P)


def _oracle_features(nl, a_bits, P, max_level, b_bits=None, robust=True):
    net = orc.OracleNet(nl.arrays())
    codes = orc.sim_chunk(net, a_bits, P, 0, P, b_bits=b_bits)
    _, fwd = orc.mark_chunk(net, codes, robust)
    return orank.features(nl.arrays(), codes, fwd, max_level)


def _features_gpu(g, P, max_level, pad=0, robust=True):
    """GPU features as (P, 512) uint16, written into a feature-major matrix
    with `pad` extra columns on each side (left untouched)."""
    ld = (P + 2 * pad + 63) // 64 * 64 + 64 * (pad > 0)
    feat = torch.full((pg.FEATURES, ld), 0x7F7F, dtype=torch.int16, device=DEV)
    g.pair_features(max_level, robust, feat[:, pad:pad + P] if pad else feat)
    full = feat.cpu().numpy().view(np.uint16)
    if pad:
        assert (full[:, :pad] == 0x7F7F).all() and (full[:, pad + P:] == 0x7F7F).all()
    return full[:, pad:pad + P].T.copy()


# ----------------------------------------------------------------- features

@pytest.mark.parametrize("mode,P,chunk,max_level,pad,robust", [("seq", 1024, 1024, 99, 0, True),
                                                               ("seq", 2048, 512, 2, 64, False),
                                                               ("pairs", 1000, 256, 0, 0, True),
                                                               ("pairs", 777, 777, 5, 8, False)])
def test_pair_features_match_oracle(mode, P, chunk, max_level, pad, robust):
    _say("features", mode, P, chunk, max_level, pad, robust)
    nl = pg.Netlist.generate(small_config(21, num_pi=40, num_gates=300, levels=9, alpha=2.2))
    I = 40
    g = pg.Grader(nl)
    if mode == "seq":
        a = pg.gen_vectors(31, I, P)
        b = None
        g.set_vectors(a, P)
    else:
        a, b = _explicit(I, P, 33)
        g.set_pairs(a, b, P)
    g.set_chunk(chunk)
    got = _features_gpu(g, P, max_level, pad, robust)
    want = _oracle_features(nl, a, P, max_level, b, robust)
    np.testing.assert_array_equal(got, want)
    assert got[:, :256].any() and got[:, 256:].any()
    # the planes now hold a partial simulation; grading still starts afresh
    tot = g.grade(True)
    assert tot == orc.grade(orc.OracleNet(nl.arrays()), a, P, chunk, True, b_bits=b)["total"]
    g.close()


def test_pair_features_wide_netlist_buckets():
    """300 PIs (uneven buckets) and a 2000-gate netlist."""
    _say("features wide")
    nl = pg.Netlist.generate(pg.GenConfig(seed=41, num_pi=300, num_po=64, num_gates=2000, levels=20,
                                          fanin_p=(0.1, 0.6, 0.2, 0.1), prev_level_frac=0.7, fanout_alpha=2.2,
                                          fanout_cap=64))
    P = 512
    a, b = _explicit(300, P, 43)
    g = pg.Grader(nl)
    g.set_pairs(a, b, P)
    got = _features_gpu(g, P, 255)
    np.testing.assert_array_equal(got, _oracle_features(nl, a, P, 255, b))
    g.close()


# -------------------------------------------------------------------- gains

@pytest.mark.parametrize("seed,mode,robust,chunk", [(51, "pairs", True, 128), (52, "seq", False, 256),
                                                    (53, "pairs", False, 1000)])
def test_pair_gains_match_oracle(seed, mode, robust, chunk):
    _say("gains", seed, mode, robust, chunk)
    nl = pg.Netlist.generate(small_config(seed, num_pi=24, num_gates=400, levels=10, alpha=2.2))
    P = 1000 if mode == "pairs" else 1024
    g = pg.Grader(nl)
    if mode == "seq":
        a, b = pg.gen_vectors(seed, 24, P), None
        g.set_vectors(a, P)
    else:
        a, b = _explicit(24, P, seed)
        g.set_pairs(a, b, P)
    g.set_chunk(chunk)
    gains = torch.zeros(P, dtype=torch.int32, device=DEV)
    g.pair_gains(robust, gains)
    got = gains.cpu().numpy().view(np.uint32)
    want = orank.pair_gains(nl.arrays(), a, P, robust, b_bits=b)
    np.testing.assert_array_equal(got, want)
    assert want.any()
    g.close()


def test_pair_gains_5k_gates():
    _say("gains 5k")
    nl = pg.Netlist.generate(pg.GenConfig(seed=61, num_pi=128, num_po=64, num_gates=5000, levels=30,
                                          fanin_p=(0.1, 0.6, 0.2, 0.1), prev_level_frac=0.7, fanout_alpha=2.2,
                                          fanout_cap=64))
    P = 512
    a, b = _explicit(128, P, 63)
    g = pg.Grader(nl)
    g.set_pairs(a, b, P)
    gains = torch.zeros(P, dtype=torch.int32, device=DEV)
    g.pair_gains(True, gains)
    np.testing.assert_array_equal(gains.cpu().numpy().view(np.uint32),
                                  orank.pair_gains(nl.arrays(), a, P, True, b_bits=b))
    g.close()


# ------------------------------------------------------------ rank + curve

def test_rank_pairs_then_grade_matches_oracle():
    _say("rank")
    nl = pg.Netlist.generate(small_config(71, num_pi=30, num_gates=500, levels=10, alpha=2.2))
    P = 3000
    a, b = _explicit(30, P, 73)
    rng = np.random.default_rng(5)
    scores = rng.standard_normal(P).astype(np.float32)
    scores[rng.integers(0, P, 400)] = 0.75      # ties keep their pair order
    g = pg.Grader(nl)
    g.set_pairs(a, b, P)
    g.set_chunk(1024)
    perm = torch.zeros(P, dtype=torch.int32, device=DEV)
    g.rank_pairs(torch.from_numpy(scores).to(DEV), perm)
    order = np.argsort(-scores, kind="stable")
    np.testing.assert_array_equal(perm.cpu().numpy(), order)
    # a second ranking with reversed scores (reusing the scratch buffers)
    # composes with the first
    g.rank_pairs(torch.from_numpy(-scores[order]).to(DEV), perm)
    order2 = np.argsort(scores[order], kind="stable")
    np.testing.assert_array_equal(perm.cpu().numpy(), order2)
    order = order[order2]
    # the ranked pairs graded on the GPU = the oracle on host-permuted pairs
    from helpers import pack, unpack
    pa = pack(unpack(a, P)[order])
    pb = pack(unpack(b, P)[order])
    want = orc.grade(orc.OracleNet(nl.arrays()), pa, P, 1024, True, b_bits=pb)
    assert g.grade(True) == want["total"]
    f0, f1 = g.first()
    np.testing.assert_array_equal(f0, want["first_t0"])
    np.testing.assert_array_equal(f1, want["first_t1"])
    np.testing.assert_array_equal(g.gate_counts(), want["gate_counts"])
    g.close()


def test_rank_pairs_rejects_sequence_mode():
    nl = pg.Netlist.generate(small_config(72, num_pi=8, num_gates=40, levels=5))
    g = pg.Grader(nl)
    g.set_vectors(pg.gen_vectors(1, 8, 256), 256)
    with pytest.raises(pg.PdfError):
        g.rank_pairs(torch.zeros(256, dtype=torch.float32, device=DEV))
    g.close()


@pytest.mark.parametrize("mode", ["seq", "pairs"])
def test_grade_curve_matches_chunk_totals(mode):
    _say("curve", mode)
    nl = pg.Netlist.generate(small_config(81, num_pi=20, num_gates=300, levels=9, alpha=2.2))
    P = 4096
    g = pg.Grader(nl)
    if mode == "seq":
        a, b = pg.gen_vectors(82, 20, P), None
        g.set_vectors(a, P)
    else:
        a, b = _explicit(20, P, 82)
        g.set_pairs(a, b, P)
    g.set_chunk(512)
    pt = torch.full((P,), 7, dtype=torch.int32, device=DEV)
    cum, ms = g.grade_curve(True, pt)
    want = orc.grade(orc.OracleNet(nl.arrays()), a, P, 512, True, b_bits=b)
    per_pair = np.cumsum(pt.cpu().numpy().astype(np.uint64))
    np.testing.assert_array_equal(per_pair[511::512], cum)
    # the oracle's per-pair totals: each single-pair cover against the final first[]
    net = orc.OracleNet(nl.arrays())
    marks, _ = orc.mark_chunk(net, orc.sim_chunk(net, a, P, 0, P, b_bits=b), True)
    some = np.random.default_rng(0).choice(P, 64, replace=False)
    for p in some:
        tot, _ = orc.cover_chunk(net, np.ascontiguousarray(marks[:, p:p + 1]), int(p), want["first_t0"].copy(),
                                 want["first_t1"].copy(), np.zeros(net.n, np.uint64))
        assert int(pt[int(p)]) == tot
    np.testing.assert_array_equal(np.diff(np.concatenate([[0], cum])).astype(np.uint64), want["chunk_totals"])
    assert cum[-1] == want["total"] == g.total()
    assert (np.diff(ms) >= 0).all() and ms[0] > 0
    np.testing.assert_array_equal(g.gate_counts(), want["gate_counts"])
    g.close()


# --------------------------------------------------------------------- MLP

def _rand_params(seed):
    rng = np.random.default_rng(seed)
    parts = dict(W1=rng.uniform(-1, 1, (256, 512)) * np.sqrt(6 / 512), b1=rng.uniform(-0.1, 0.1, 256),
                 W2=rng.uniform(-1, 1, (256, 256)) * np.sqrt(6 / 256), b2=rng.uniform(-0.1, 0.1, 256),
                 w3=rng.uniform(-1, 1, 256) * np.sqrt(6 / 256), b3=np.array([0.3]))
    return omlp.flatten(parts).astype(np.float32)


def _rand_feat(rows, seed):
    rng = np.random.default_rng(seed)
    return orank.bf16_round(rng.random((rows, 512)).astype(np.float32))


def _dev16(x, col0=0):
    """(rows, 512) uint16 features -> a feature-major device column slice
    starting at column col0 of a zero matrix with a row stride multiple of 64."""
    rows = x.shape[0]
    ld = (col0 + rows + 8 + 63) // 64 * 64
    m = torch.zeros((512, ld), dtype=torch.int16, device=DEV)
    m[:, col0:col0 + rows] = torch.from_numpy(np.ascontiguousarray(x.T).view(np.int16)).to(DEV)
    return m[:, col0:col0 + rows]


@pytest.mark.parametrize("rows,col0", [(1, 0), (200, 8), (1000, 0), (4096, 64)])
def test_mlp_forward_matches_oracle(rows, col0):
    _say("mlp forward", rows, col0)
    m = pg.Mlp(seed=3)
    flat = _rand_params(rows)
    m.set_params(flat)
    np.testing.assert_array_equal(m.params(), flat)
    feat = _rand_feat(rows, rows + 1)
    out = torch.zeros(rows, dtype=torch.float32, device=DEV)
    m.forward(_dev16(feat, col0), out)
    got = out.cpu().numpy()
    want, _ = omlp.forward(flat.astype(np.float64), orank.bf16_to_f32(feat), quant=True)
    np.testing.assert_allclose(got, want, rtol=2e-3, atol=2e-3)
    m.close()


@pytest.mark.parametrize("rows", [256, 1000, 4093])
def test_mlp_gradients_match_oracle(rows):
    _say("mlp gradients", rows)
    m = pg.Mlp(seed=4)
    flat = _rand_params(100 + rows)
    m.set_params(flat)
    feat = _rand_feat(rows, rows + 7)
    rng = np.random.default_rng(rows)
    t = rng.standard_normal(rows).astype(np.float32)
    loss, grad = m.gradients(_dev16(feat, 24), torch.from_numpy(t).to(DEV))
    wl, wg = omlp.loss_and_grad(flat.astype(np.float64), orank.bf16_to_f32(feat), t, quant=True)
    assert abs(loss - wl) <= 1e-4 * wl
    a, b = omlp.unflatten(grad.astype(np.float64)), omlp.unflatten(wg)
    for k in a:
        err = np.linalg.norm(a[k] - b[k]) / np.linalg.norm(b[k])
        assert err < 1e-2, (k, err)
    m.close()


def test_mlp_adam_step_matches_oracle():
    _say("mlp adam")
    m = pg.Mlp(seed=5)
    flat = _rand_params(9)
    m.set_params(flat)
    feat = _rand_feat(512, 10)
    t = np.random.default_rng(11).standard_normal(512).astype(np.float32)
    w, mm, vv = flat.copy(), np.zeros_like(flat), np.zeros_like(flat)
    for step in (1, 2, 3):
        _, grad = m.gradients(_dev16(feat), torch.from_numpy(t).to(DEV))
        m.adam(1e-3)
        w, mm, vv = omlp.adam_step(w, grad, mm, vv, step, 1e-3)
        np.testing.assert_allclose(m.params(), w, rtol=1e-5, atol=1e-7)
    m.close()


def test_mlp_training_reduces_loss():
    """A learnable target (a fixed random projection of the features through
    a ReLU) is fitted by train_step alone."""
    _say("mlp train")
    rows = 8192
    feat = _rand_feat(rows, 12)
    x = orank.bf16_to_f32(feat)
    proj = np.random.default_rng(13).standard_normal((512, 8)).astype(np.float32) / 8
    t = np.maximum(x @ proj, 0).sum(axis=1)
    t = ((t - t.mean()) / t.std()).astype(np.float32)
    m = pg.Mlp(seed=6)
    f_dev, t_dev = _dev16(feat), torch.from_numpy(t).to(DEV)
    first = m.train_step(f_dev, t_dev, 1e-3, want_loss=True)
    for _ in range(150):
        m.train_step(f_dev, t_dev, 1e-3)
    last = m.train_step(f_dev, t_dev, 1e-3, want_loss=True)
    _say("loss", first, last)
    assert last < 0.2 * first
    m.close()


@pytest.mark.parametrize("lanes,chunk,P", [(2, 1024, 5000), (1, 2048, 4096), (3, 512, 3000)])
def test_score_pairs_equals_features_then_forward(lanes, chunk, P):
    """The pipelined feature + scoring pass writes the same features and the
    same scores (bit for bit) as pair_features followed by Mlp.forward."""
    _say("score pairs", lanes, chunk, P)
    nl = pg.Netlist.generate(small_config(61, num_pi=40, num_gates=900, levels=14))
    a, b = _explicit(40, P, 610)
    m = pg.Mlp(seed=61)
    ld = (P + 63) // 64 * 64
    with pg.Grader(nl) as g:
        g.set_pairs(a, b, P)
        g.set_tuning(pg.TUNE_LANES, lanes)
        g.set_chunk(chunk)
        f1 = torch.full((pg.FEATURES, ld), 0x7F7F, dtype=torch.int16, device=DEV)
        s1 = torch.zeros(P, dtype=torch.float32, device=DEV)
        g.pair_features(4, False, f1)
        m.forward(f1[:, :P], s1)
        f2 = torch.full((pg.FEATURES, ld), 0x7F7F, dtype=torch.int16, device=DEV)
        s2 = torch.zeros(P, dtype=torch.float32, device=DEV)
        g.score_pairs(m, 4, False, f2, s2)
        torch.cuda.synchronize()
        assert torch.equal(f1[:, :P], f2[:, :P])
        assert torch.equal(s1, s2)
        assert float(s1.abs().sum()) > 0
    m.close()
