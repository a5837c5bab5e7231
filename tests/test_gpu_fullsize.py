"""Parity at the workloads' full sizes.

* 2,000,000-gate / 200-level netlist, all 1,048,576 pairs: the recorded
  tuning (2 lanes x 131,072-pair chunks x 2 words per lane, the compacted
  reverse pass) and 32,768-pair
  blocks, each against the bit-sliced CPU grader on every host thread
  (total, first-detection pairs, per-gate counts: u64 accumulation across
  all blocks).
* A held-out circuit of the learned-ranking experiment (50,000 gates,
  1,048,576 explicit candidate pairs, the experiment's chunking): features
  and solo gains of two 2,048-pair slices of the full-size GPU run against
  the CPU oracle, and the grade of all candidates in a GPU-ranked order
  against the bit-sliced grader on the same order applied to the host pairs.
"""
import sys

import numpy as np
import pytest

import pdfgrade as pg
from pdfgrade.configs import GATES_2M
from oracle import oracle as orc
from oracle import ranking as orank

from helpers import bitsliced_grade, pack, unpack

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def state(g, total):
    f0, f1 = g.first()
    return dict(total=total, first_t0=f0, first_t1=f1, gate_counts=g.gate_counts())


def assert_same(got, ref, ctx):
    assert got["total"] == ref["total"], (ctx, got["total"], ref["total"])
    assert np.array_equal(got["first_t0"], ref["first_t0"]), ctx
    assert np.array_equal(got["first_t1"], ref["first_t1"]), ctx
    assert np.array_equal(got["gate_counts"], ref["gate_counts"]), ctx


def test_2m_gates_all_pairs_against_bitsliced():
    say("2m gates: generating")
    nl = pg.Netlist.generate(GATES_2M.netlist)
    P = GATES_2M.pairs
    assert P == 1_048_576
    bits = pg.gen_vectors(GATES_2M.vector_seed, nl.info()["num_pi"], P)
    say("2m gates: recorded tuning on the GPU")
    with pg.Grader(nl) as g:
        g.set_vectors(bits, P)
        g.set_chunk(GATES_2M.chunk)
        g.set_tuning(pg.TUNE_LANES, GATES_2M.lanes)
        g.set_tuning(pg.TUNE_WORDS_PER_LANE, GATES_2M.words_per_lane)
        g.set_tuning(pg.TUNE_REV_COMPACT, GATES_2M.rev_compact)
        tuned = state(g, g.grade(True))
        # the reference's merge launches reach 1,310,700 of the 2,004,096
        # nodes (mergekernel.cuh:243)
        flags, unmerged = g.ref_flags()
        assert flags & pg.REF_SEGMENT_STRIDE and unmerged == 693_396
    say("2m gates: 32,768-pair blocks on the GPU")
    with pg.Grader(nl) as g:
        g.set_vectors(bits, P)
        g.set_chunk(32768)
        blocks = state(g, g.grade(True))
    say("2m gates: bit-sliced CPU grader (all host threads)")
    ref = bitsliced_grade(nl, bits, P, 32768, True)
    say("2m gates: total", ref["total"], "detected", int((ref["first_t0"] != pg.NONE).sum()))
    assert_same(tuned, ref, "recorded tuning")
    assert_same(blocks, ref, "32k blocks")


def test_learned_heldout_circuit_full_size():
    torch = pytest.importorskip("torch")
    from pdfgrade import learned

    cfg = learned.LearnedConfig()
    # held-out circuit 5102: 20 robust and 30 non-robust paths over its
    # candidates (5100 has none)
    seed = cfg.heldout_seeds[2]
    nl = pg.Netlist.generate(learned.circuit(seed, cfg.num_gates, cfg.levels, cfg.num_pi, cfg.num_po))
    I = nl.info()["num_pi"]
    P = cfg.candidates
    assert P == 1 << 20
    a, b = learned.candidate_bits(I, P, cfg.candidate_seed + (seed - cfg.heldout_seeds[0]), cfg.q_lo, cfg.q_hi)
    dev = "cuda:0"
    g = pg.Grader(nl)
    g.set_pairs(a, b, P)
    g.set_chunk(cfg.chunk)
    g.set_tuning(pg.TUNE_LANES, cfg.lanes)
    say("learned: features and gains of all candidates")
    feat = torch.zeros((pg.FEATURES, P), dtype=torch.int16, device=dev)
    g.pair_features(cfg.max_level, cfg.robust, feat)
    gains = torch.zeros(P, dtype=torch.int32, device=dev)
    g.pair_gains(cfg.robust, gains)
    feat_h = feat.cpu().numpy().view(np.uint16)
    gains_h = gains.cpu().numpy().view(np.uint32)
    arrs = nl.arrays()
    net = orc.OracleNet(arrs)
    S = 2048
    for p0 in (0, P // 2 + 64 * 7):
        say("learned: oracle slice", p0)
        w0 = p0 // 64
        sa = np.ascontiguousarray(a[:, w0:w0 + S // 64])
        sb = np.ascontiguousarray(b[:, w0:w0 + S // 64])
        codes = orc.sim_chunk(net, sa, S, 0, S, b_bits=sb)
        _, fwd = orc.mark_chunk(net, codes, cfg.robust)
        np.testing.assert_array_equal(feat_h[:, p0:p0 + S].T, orank.features(arrs, codes, fwd, cfg.max_level))
        np.testing.assert_array_equal(gains_h[p0:p0 + S], orank.pair_gains(arrs, sa, S, cfg.robust, b_bits=sb))
    del feat
    say("learned: rank all candidates on the GPU, then grade")
    gen = torch.Generator(device=dev)
    gen.manual_seed(7)
    # the solo gain plus uniform noise in [0, 1): pairs with gains first, the rest shuffled
    scores = gains.to(torch.float32) + torch.rand(P, generator=gen, device=dev)
    perm = torch.zeros(P, dtype=torch.int32, device=dev)
    g.rank_pairs(scores, perm)
    g.set_chunk(cfg.curve_chunk)
    got = {robust: state(g, g.grade(robust)) for robust in (True, False)}
    g.close()
    order = perm.cpu().numpy().astype(np.int64)
    assert np.array_equal(np.sort(order), np.arange(P))
    assert not np.array_equal(order, np.arange(P))
    pa = pack(unpack(a, P)[order])
    pb = pack(unpack(b, P)[order])
    for robust in (True, False):
        say("learned: bit-sliced CPU grader on the ranked pairs", robust)
        ref = bitsliced_grade(nl, pa, P, cfg.curve_chunk, robust, b_bits=pb)
        assert ref["total"] > 0
        assert_same(got[robust], ref, ("ranked order", robust))
