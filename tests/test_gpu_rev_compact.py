"""The compacted reverse pass (PDF_TUNE_REV_COMPACT on a grader): the forward
pass appends each level's marked (gate, 64-word tile) items to a list and the
reverse launches of levels >= 1 run over those lists only.  Integer work, so
every output must equal the oracle's (small cases) or the dense reverse
pass's and the bit-sliced grader's (full sizes)."""
import sys

import numpy as np
import pytest

import pdfgrade as pg
from oracle import oracle as orc
from pdfgrade.configs import GATES_2M, GATES_50K, TINY

from helpers import biased_vectors, bitsliced_grade, edge_netlist, oracle_grade, pack, small_config, unpack

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def state(g, total, per_chunk=None):
    f0, f1 = g.first()
    return dict(total=total, per_chunk=per_chunk, first_t0=f0, first_t1=f1, gate_counts=g.gate_counts())


def assert_same(got, ref, ctx=""):
    assert got["total"] == ref["total"], (ctx, got["total"], ref["total"])
    assert np.array_equal(got["first_t0"], ref["first_t0"]), ctx
    assert np.array_equal(got["first_t1"], ref["first_t1"]), ctx
    assert np.array_equal(got["gate_counts"], ref["gate_counts"]), ctx
    if got.get("per_chunk") is not None and "chunk_totals" in ref:
        assert np.array_equal(np.array(got["per_chunk"], np.uint64), ref["chunk_totals"]), ctx


def compact_grades(nl, bits, P, chunk, robust, b_bits=None, lanes=1, graph=1, wpl=1, stages=True, reps=1):
    """Fused grades (reps of them) and, with stages, the staged chunk loop,
    with the compacted reverse pass on."""
    out = []
    with pg.Grader(nl) as g:
        if b_bits is None:
            g.set_vectors(bits, P)
        else:
            g.set_pairs(bits, b_bits, P)
        g.set_chunk(chunk)
        g.set_tuning(pg.TUNE_REV_COMPACT, 1)
        g.set_tuning(pg.TUNE_LANES, lanes)
        g.set_tuning(pg.TUNE_USE_GRAPH, graph)
        g.set_tuning(pg.TUNE_WORDS_PER_LANE, wpl)
        for _ in range(reps):
            out.append(("fused", state(g, g.grade(robust))))
        if stages:
            total, per_chunk = g.grade_stages(robust)
            out.append(("stages", state(g, total, per_chunk)))
    return out


@pytest.mark.parametrize("robust", [True, False])
def test_compact_ragged_chunks_against_oracle(robust):
    """1,500 gates over a 65,536-pair chunk (16 tiles) and a ragged
    40,859-pair last chunk (639 words: 9 full tiles and a partial one), one
    and two lanes, graph replay and direct launches, one and two words per
    lane in the forward pass, fused and stage by stage."""
    cfg = pg.GenConfig(seed=77, num_pi=64, num_po=64, num_gates=1500, levels=20, fanin_p=(0.1, 0.6, 0.2, 0.1),
                       prev_level_frac=0.7, fanout_alpha=2.2, fanout_cap=64)
    nl = pg.Netlist.generate(cfg)
    P = 65536 + 40896 - 37
    bits = biased_vectors(77, nl.info()["num_pi"], P, p_hold=0.6)
    ref = oracle_grade(nl, bits, P, 65536, robust)
    assert ref["total"] > 0
    for lanes, graph, wpl in ((1, 1, 1), (1, 0, 2), (2, 1, 2), (2, 0, 1)):
        say("compact ragged", robust, lanes, graph, wpl)
        for kind, got in compact_grades(nl, bits, P, 65536, robust, lanes=lanes, graph=graph, wpl=wpl, reps=2):
            assert_same(got, ref, (kind, lanes, graph, wpl))


@pytest.mark.parametrize("seed", range(6))
def test_compact_random_circuits(seed):
    """Random small circuits, both modes, chunks of 64 words and more (the
    lists), 40 words and 1 word (dense reverse launches: the lists need a
    whole 64-word tile per wave)."""
    nl = pg.Netlist.generate(small_config(seed + 300, num_pi=10, num_po=12, num_gates=400, levels=8, wide=True,
                                          alpha=2.2))
    P = 9000
    bits = biased_vectors(seed + 300, nl.info()["num_pi"], P, p_hold=0.5)
    for robust in (True, False):
        for chunk in (4096, 8192, 2560, 64):
            say("compact random", seed, robust, chunk)
            ref = oracle_grade(nl, bits, P, chunk, robust)
            for kind, got in compact_grades(nl, bits, P, chunk, robust):
                assert_same(got, ref, (kind, seed, robust, chunk))


@pytest.mark.parametrize("robust", [True, False])
def test_compact_edge_netlist_explicit_pairs(robust):
    """Every gate type, 12-input gates, explicit pairs (two planes)."""
    nl = edge_netlist()
    P = 5000
    va = biased_vectors(11, nl.info()["num_pi"], P, p_hold=0.3)
    vb = biased_vectors(12, nl.info()["num_pi"], P, p_hold=0.3)
    ref = oracle_grade(nl, va, P, 4096, robust, b_bits=vb)
    say("compact edge", robust)
    for kind, got in compact_grades(nl, va, P, 4096, robust, b_bits=vb):
        assert_same(got, ref, (kind, robust))


def test_compact_tiny_and_mark_planes():
    """The 200-gate workload at 1,024 pairs in one 16-word chunk (dense) and
    on 4,096 pairs in one 64-word chunk (lists): totals against the oracle,
    and the stage calls' mark plane against the oracle's final marks."""
    nl = pg.Netlist.generate(TINY.netlist)
    net = orc.OracleNet(nl.arrays())
    for P in (1024, 4096):
        bits = pg.gen_vectors(TINY.vector_seed, nl.info()["num_pi"], P)
        for robust in (True, False):
            say("compact tiny", P, robust)
            ref = oracle_grade(nl, bits, P, P, robust)
            for kind, got in compact_grades(nl, bits, P, P, robust):
                assert_same(got, ref, (kind, P, robust))
            with pg.Grader(nl) as g:
                g.set_vectors(bits, P)
                g.set_chunk(P)
                g.set_tuning(pg.TUNE_REV_COMPACT, 1)
                g.sim(0, 0)
                g.mark(0, 0, robust)
                codes = orc.sim_chunk(net, bits, P, 0, P)
                marks, _ = orc.mark_chunk(net, codes, robust)
                assert np.array_equal(pg.plane_to_bytes(g.mark_plane(P), P), marks), (P, robust)
    with pg.Grader(nl) as g, pytest.raises(pg.PdfError):
        g.set_tuning(pg.TUNE_REV_COMPACT, 2)


@pytest.mark.parametrize("robust", [True, False])
def test_compact_50k_against_bitsliced(robust):
    """The 50k-gate workload's 65,536 pairs with per-level launches on 2
    lanes x 32,768 pairs, against the bit-sliced CPU grader."""
    nl = pg.Netlist.generate(GATES_50K.netlist)
    P = GATES_50K.pairs
    bits = pg.gen_vectors(GATES_50K.vector_seed, nl.info()["num_pi"], P)
    say("compact 50k", robust)
    ref = bitsliced_grade(nl, bits, P, 32768, robust)
    for kind, got in compact_grades(nl, bits, P, 32768, robust, lanes=2, stages=False, reps=2):
        assert_same(got, ref, (kind, robust))


def test_compact_2m_against_dense():
    """The 2,000,000-gate netlist on 262,144 pairs with its recorded tuning
    (2 lanes x 131,072-pair chunks x 2 words per lane), the compacted reverse
    pass against the dense one: same total, first detections and counts."""
    nl = pg.Netlist.generate(GATES_2M.netlist)
    P = 262144
    bits = pg.gen_vectors(GATES_2M.vector_seed, nl.info()["num_pi"], P)
    res = {}
    for compact in (0, 1):
        say("compact 2m", compact)
        with pg.Grader(nl) as g:
            g.set_vectors(bits, P)
            g.set_chunk(GATES_2M.chunk)
            g.set_tuning(pg.TUNE_LANES, GATES_2M.lanes)
            g.set_tuning(pg.TUNE_WORDS_PER_LANE, GATES_2M.words_per_lane)
            g.set_tuning(pg.TUNE_REV_COMPACT, compact)
            res[compact] = [state(g, g.grade(True)) for _ in range(2)]
    assert res[0][0]["total"] > 0
    for rep in range(2):
        assert_same(res[1][rep], res[0][0], rep)


@pytest.mark.parametrize("order", [0, 1])
def test_forward_gate_order(order):
    """The per-level forward launches in id order (0, the default) and in
    fanin-locality order (1): random circuits with both reverse passes, fused
    and stage by stage, against the oracle."""
    for seed in range(3):
        nl = pg.Netlist.generate(small_config(seed + 500, num_pi=12, num_po=16, num_gates=600, levels=9, wide=True,
                                              alpha=2.2))
        P = 8192
        bits = biased_vectors(seed + 500, nl.info()["num_pi"], P, p_hold=0.5)
        for robust in (True, False):
            ref = oracle_grade(nl, bits, P, 4096, robust)
            for compact in (0, 1):
                say("forward order", order, seed, robust, compact)
                with pg.Grader(nl) as g:
                    g.set_vectors(bits, P)
                    g.set_chunk(4096)
                    g.set_tuning(pg.TUNE_FWD_ORDER, order)
                    g.set_tuning(pg.TUNE_REV_COMPACT, compact)
                    assert_same(state(g, g.grade(robust)), ref, (order, seed, robust, compact, "fused"))
                    total, per_chunk = g.grade_stages(robust)
                    assert_same(state(g, total, per_chunk), ref, (order, seed, robust, compact, "stages"))
    with pg.Grader(nl) as g, pytest.raises(pg.PdfError):
        g.set_tuning(pg.TUNE_FWD_ORDER, 2)


def test_compact_lists_follow_their_forward_pass():
    """The lists belong to the forward pass that filled them: an event-timed
    grade (one lane, lists built and used per chunk) followed by direct
    launches on two lanes (levels interleaved across the lanes' chunks, so
    dense reverse launches) and by graph replays, in every order, each equal
    to the oracle."""
    cfg = pg.GenConfig(seed=91, num_pi=32, num_po=32, num_gates=900, levels=12, fanin_p=(0.1, 0.6, 0.2, 0.1),
                       prev_level_frac=0.7, fanout_alpha=2.2, fanout_cap=64)
    nl = pg.Netlist.generate(cfg)
    P = 4 * 8192 - 640
    bits = biased_vectors(91, nl.info()["num_pi"], P, p_hold=0.6)
    for robust in (True, False):
        ref = oracle_grade(nl, bits, P, 8192, robust)
        with pg.Grader(nl) as g:
            g.set_vectors(bits, P)
            g.set_chunk(8192)
            g.set_tuning(pg.TUNE_REV_COMPACT, 1)
            g.set_tuning(pg.TUNE_LANES, 2)
            for step, graph, timed in ((0, 0, True), (1, 0, False), (2, 1, True), (3, 1, False), (4, 0, True),
                                       (5, 1, False), (6, 0, False)):
                say("compact lists order", robust, step)
                g.set_tuning(pg.TUNE_USE_GRAPH, graph)
                total = g.grade_timed(robust)[0] if timed else g.grade(robust)
                assert_same(state(g, total), ref, (robust, step))



def test_compact_more_tiles_than_sublists():
    """Chunks of 524,288 pairs (128 64-word tiles, two tiles to each of the
    64 sub-lists) and a ragged last chunk, both modes, against the bit-sliced
    CPU grader."""
    cfg = pg.GenConfig(seed=95, num_pi=48, num_po=48, num_gates=800, levels=14, fanin_p=(0.1, 0.6, 0.2, 0.1),
                       prev_level_frac=0.7, fanout_alpha=2.2, fanout_cap=64)
    nl = pg.Netlist.generate(cfg)
    P = 524288 + 70000
    bits = biased_vectors(95, nl.info()["num_pi"], P, p_hold=0.5)
    for robust in (True, False):
        say("compact 128 tiles", robust)
        ref = bitsliced_grade(nl, bits, P, 524288, robust)
        assert ref["total"] > 0
        for kind, got in compact_grades(nl, bits, P, 524288, robust, wpl=2, stages=False, reps=2):
            assert_same(got, ref, (kind, robust))
