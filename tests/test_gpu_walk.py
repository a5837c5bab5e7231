"""The level walk (PDF_TUNE_WALK: one forward and one reverse launch per chunk,
a workgroup per tile of 1/2/4/8 words walking every level) and the reverse
walk alone after per-level forward launches (PDF_TUNE_REV_WALK) against the
CPU oracle on small circuits and the bit-sliced CPU grader at full size;
integer work, so every output must be equal."""
import sys

import numpy as np
import pytest

import pdfgrade as pg
from pdfgrade.configs import GATES_50K, TINY

from helpers import biased_vectors, bitsliced_grade, edge_netlist, oracle_grade, pack, small_config, unpack

pytestmark = pytest.mark.gpu

WALKS = (1, 2, 4, 8)
# (forward and reverse walk, reverse walk only): every tile width either way
MODES = [(tw, 0) for tw in WALKS] + [(0, tw) for tw in WALKS]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def walk_grade(nl, bits, P, chunk, robust, tw, b_bits=None, lanes=1, graph=1, reps=1, rev=0):
    with pg.Grader(nl) as g:
        if b_bits is None:
            g.set_vectors(bits, P)
        else:
            g.set_pairs(bits, b_bits, P)
        g.set_chunk(chunk)
        g.set_tuning(pg.TUNE_WALK, tw)
        g.set_tuning(pg.TUNE_REV_WALK, rev)
        g.set_tuning(pg.TUNE_LANES, lanes)
        g.set_tuning(pg.TUNE_USE_GRAPH, graph)
        out = []
        for _ in range(reps):
            total = g.grade(robust)
            f0, f1 = g.first()
            out.append(dict(total=total, first_t0=f0, first_t1=f1, gate_counts=g.gate_counts()))
        return out


def assert_same(got, ref, ctx=""):
    assert got["total"] == ref["total"], (ctx, got["total"], ref["total"])
    assert np.array_equal(got["first_t0"], ref["first_t0"]), ctx
    assert np.array_equal(got["first_t1"], ref["first_t1"]), ctx
    assert np.array_equal(got["gate_counts"], ref["gate_counts"]), ctx


@pytest.mark.parametrize("tw,rev", MODES)
def test_walk_tiny_workload(tw, rev):
    nl = pg.Netlist.generate(TINY.netlist)
    bits = pg.gen_vectors(TINY.vector_seed, nl.info()["num_pi"], TINY.pairs)
    for robust in (True, False):
        for chunk in (64, 320, 1024):
            say("walk tiny", tw, robust, chunk)
            ref = oracle_grade(nl, bits, TINY.pairs, chunk, robust)
            for got in walk_grade(nl, bits, TINY.pairs, chunk, robust, tw, reps=2, rev=rev):
                assert_same(got, ref, (tw, rev, robust, chunk))


@pytest.mark.parametrize("tw,rev", MODES)
def test_walk_random_circuits(tw, rev):
    for seed in range(10):
        cfg = small_config(100 + seed, num_pi=6 + seed % 5, num_gates=40 + 9 * seed, levels=4 + seed % 7,
                           wide=seed % 2 == 0, alpha=2.2 if seed % 3 == 0 else 0.0)
        nl = pg.Netlist.generate(cfg)
        P = 64 * (3 + seed % 5) + (seed * 13) % 64  # ragged: not a multiple of 64 for most seeds
        bits = biased_vectors(seed, nl.info()["num_pi"], P, p_hold=0.3 + 0.1 * (seed % 5))
        for robust in (True, False):
            say("walk random", tw, seed, robust)
            chunk = 64 * (1 + seed % 3)
            ref = oracle_grade(nl, bits, P, chunk, robust)
            assert_same(walk_grade(nl, bits, P, chunk, robust, tw, rev=rev)[0], ref, (tw, rev, seed, robust))


@pytest.mark.parametrize("tw,rev", MODES)
def test_walk_edge_netlist_and_explicit_pairs(tw, rev):
    nl = edge_netlist()
    P = 704  # 11 words: a partial last tile for every tw > 1
    bits = biased_vectors(9, 8, P, p_hold=0.5)
    I = 8
    b = pack(np.where(np.random.default_rng(3).random((P, I)) < 0.3, 1 - unpack(bits, P), unpack(bits, P)))
    for robust in (True, False):
        say("walk edge", tw, robust)
        ref = oracle_grade(nl, bits, P, 448, robust)
        assert_same(walk_grade(nl, bits, P, 448, robust, tw, rev=rev)[0], ref, ("edge", tw, rev, robust))
        ref = oracle_grade(nl, bits, P, 448, robust, b_bits=b)
        assert_same(walk_grade(nl, bits, P, 448, robust, tw, b_bits=b, rev=rev)[0], ref,
                    ("explicit", tw, rev, robust))


@pytest.mark.parametrize("tw,rev", [(2, 0), (4, 0), (0, 2), (0, 4)])
def test_walk_lanes_and_graphs(tw, rev):
    """Several chunks on 2 lanes, with and without graph replay, repeated."""
    nl = pg.Netlist.generate(small_config(7, num_pi=16, num_po=24, num_gates=900, levels=14, wide=True,
                                          alpha=2.2))
    P = 8192
    bits = biased_vectors(7, nl.info()["num_pi"], P, p_hold=0.6)
    for robust in (True, False):
        ref = oracle_grade(nl, bits, P, 2048, robust)
        for graph in (0, 1):
            say("walk lanes", tw, robust, graph)
            for got in walk_grade(nl, bits, P, 2048, robust, tw, lanes=2, graph=graph, reps=2, rev=rev):
                assert_same(got, ref, (tw, rev, robust, graph))


@pytest.mark.parametrize("robust", [True, False])
def test_walk_50k_full_size_against_bitsliced(robust):
    """All 65,536 pairs of the 50k-gate workload in one chunk, every tile
    width, against the bit-sliced CPU grader."""
    nl = pg.Netlist.generate(GATES_50K.netlist)
    P = GATES_50K.pairs
    bits = pg.gen_vectors(GATES_50K.vector_seed, nl.info()["num_pi"], P)
    ref = bitsliced_grade(nl, bits, P, P, robust)
    for tw, rev in MODES:
        say("walk 50k vs bit-sliced", robust, tw, rev)
        for got in walk_grade(nl, bits, P, P, robust, tw, reps=2, rev=rev):
            assert_same(got, ref, (robust, tw, rev))
    say("walk 50k vs bit-sliced, 2 lanes x 32,768, reverse walk", robust)
    for got in walk_grade(nl, bits, P, 32768, robust, 0, lanes=2, reps=2, rev=2):
        assert_same(got, ref, (robust, "lanes"))


def test_walk_50k_hold_biased_against_bitsliced():
    """Hold-biased vectors (many more marks and paths than uniform ones) on
    the 50k-gate netlist, both modes, against the bit-sliced CPU grader."""
    nl = pg.Netlist.generate(GATES_50K.netlist)
    P = 16384
    bits = biased_vectors(41, nl.info()["num_pi"], P, p_hold=0.9)
    for robust in (True, False):
        ref = bitsliced_grade(nl, bits, P, P, robust)
        assert ref["total"] > 0
        for tw in (2, 4):
            say("walk 50k hold-biased", robust, tw)
            assert_same(walk_grade(nl, bits, P, P, robust, tw)[0], ref, (robust, tw))


@pytest.mark.parametrize("tw", [1, 4])
def test_walk_sim_planes_after_a_grade(tw):
    """The walk writes back only the records a later level reads outside its
    LDS window, so after a fused walk grade the sim-plane getter refuses;
    with a debug flag set before the grade every record is kept and the
    last chunk's planes equal the oracle's codes.  Results are the same
    either way.  The final mark plane (masked by the walk's per-tile
    mark-present bytes) equals the oracle's marks in both cases."""
    from oracle import oracle as orc

    nl = pg.Netlist.generate(small_config(21, num_pi=12, num_po=10, num_gates=700, levels=9, wide=True,
                                          alpha=2.2))
    P, chunk = 1024, 512
    bits = biased_vectors(21, nl.info()["num_pi"], P, p_hold=0.4)
    ref = oracle_grade(nl, bits, P, chunk, True)
    net = orc.OracleNet(nl.arrays())
    with pg.Grader(nl) as g:
        g.set_vectors(bits, P)
        g.set_chunk(chunk)
        g.set_tuning(pg.TUNE_WALK, tw)
        say("walk planes", tw)
        codes = orc.sim_chunk(net, bits, P, P - chunk, chunk)
        marks, _ = orc.mark_chunk(net, codes, True)
        assert g.grade(True) == ref["total"]
        with pytest.raises(pg.PdfError):
            g.sim_planes(chunk)
        assert np.array_equal(pg.plane_to_bytes(g.mark_plane(chunk), chunk), marks)
        g.set_debug(pg.DEBUG_FORWARD_MARKS)
        assert g.grade(True) == ref["total"]
        a, b = g.sim_planes(chunk)
        assert np.array_equal(pg.planes_to_codes(a, b, chunk), codes)
        assert np.array_equal(pg.plane_to_bytes(g.mark_plane(chunk), chunk), marks)
