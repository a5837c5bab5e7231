"""The cover DP's global-memory paths, forced by lowering how much of the DP
stays in LDS (PDF_TUNE_COVER_NODES, PDF_TUNE_COVER_POOL): pairs with more
marked nodes than the cap keep c/h in global scratch, and fanout ids past
the pool are read from the netlist.  Every result must equal the oracle's
(and the default LDS path's).  The circuits are chosen so that pairs do
exceed the caps, which the oracle's marks confirm."""
import sys

import numpy as np
import pytest

import pdfgrade as pg
from oracle import oracle as orc
from oracle import ranking as orank

from helpers import biased_vectors, edge_netlist, oracle_grade, small_config

pytestmark = pytest.mark.gpu

CAPS = [(1280, 4096), (4, 4096), (1280, 8), (16, 0), (1, 0)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def grade_with_caps(nl, bits, P, chunk, robust, caps, segs=1):
    with pg.Grader(nl) as g:
        g.set_vectors(bits, P)
        g.set_chunk(chunk)
        g.set_tuning(pg.TUNE_COVER_NODES, caps[0])
        g.set_tuning(pg.TUNE_COVER_POOL, caps[1])
        total = g.grade(robust) if segs == 1 else g.grade_segs(robust, segs)
        f0, f1 = g.first()
        return dict(total=total, first_t0=f0, first_t1=f1, gate_counts=g.gate_counts())


def assert_same(got, ref, ctx):
    assert got["total"] == ref["total"], (ctx, got["total"], ref["total"])
    assert np.array_equal(got["first_t0"], ref["first_t0"]), ctx
    assert np.array_equal(got["first_t1"], ref["first_t1"]), ctx
    assert np.array_equal(got["gate_counts"], ref["gate_counts"]), ctx


def max_marked_per_pair(nl, bits, P, robust):
    net = orc.OracleNet(nl.arrays())
    codes = orc.sim_chunk(net, bits, P, 0, P)
    marks, _ = orc.mark_chunk(net, codes, robust)
    return int(marks.astype(np.int64).sum(axis=0).max())


def dense_circuit(seed):
    # mostly NOT/BUFF (a fan-in-1 gate passes its input's mark on) with some
    # 2- and 3-input gates, few levels, held vectors: up to ~30 nodes and
    # ~90 fanout ids are marked at one pair
    return pg.GenConfig(seed=seed, num_pi=24, num_po=60, num_gates=600, levels=8, fanin_p=(0.7, 0.25, 0.05, 0.0),
                        prev_level_frac=0.7, fanout_alpha=2.2, fanout_cap=16)


@pytest.mark.parametrize("caps", CAPS)
def test_cover_paths_random_circuits(caps):
    most = 0
    for seed in range(4):
        nl = pg.Netlist.generate(dense_circuit(300 + seed))
        P = 1024
        bits = biased_vectors(seed, nl.info()["num_pi"], P, p_hold=0.85)
        for robust in (True, False):
            say("cover paths", caps, seed, robust)
            most = max(most, max_marked_per_pair(nl, bits, P, robust))
            ref = oracle_grade(nl, bits, P, 512, robust)
            assert ref["total"] > 0
            assert_same(grade_with_caps(nl, bits, P, 512, robust, caps), ref, (caps, seed, robust))
    # the global-memory DP ran: some pair had more marked nodes than the cap
    assert caps[0] == 1280 or most > caps[0], (most, caps)


@pytest.mark.parametrize("caps", CAPS[1:])
def test_cover_paths_edge_netlist_and_segments(caps):
    nl = edge_netlist()
    P = 512
    bits = biased_vectors(9, 8, P, p_hold=0.5)
    for robust in (True, False):
        say("cover paths edge", caps, robust)
        ref = oracle_grade(nl, bits, P, 256, robust)
        assert_same(grade_with_caps(nl, bits, P, 256, robust, caps), ref, (caps, robust))
    # segments of length 2: the DP reads the new-node plane
    nl = pg.Netlist.generate(dense_circuit(310))
    P = 768
    bits = biased_vectors(11, nl.info()["num_pi"], P, p_hold=0.85)
    ref = orc.grade_segs(orc.OracleNet(nl.arrays()), bits, P, 256, False, 2)
    got = grade_with_caps(nl, bits, P, 256, False, caps, segs=2)
    assert got["total"] == ref["total"]
    assert np.array_equal(got["gate_counts"], ref["gate_counts"])


@pytest.mark.parametrize("caps", [(1, 0), (8, 16)])
def test_cover_paths_pair_gains(caps):
    """Solo gains (every marked node new) through the global paths."""
    torch = pytest.importorskip("torch")
    nl = pg.Netlist.generate(dense_circuit(320))
    I = nl.info()["num_pi"]
    P = 640
    a = biased_vectors(21, I, P, p_hold=0.0)
    b = biased_vectors(22, I, P, p_hold=0.0)
    want = orank.pair_gains(nl.arrays(), a, P, False, b_bits=b)
    assert want.max() > 0
    with pg.Grader(nl) as g:
        g.set_pairs(a, b, P)
        g.set_tuning(pg.TUNE_COVER_NODES, caps[0])
        g.set_tuning(pg.TUNE_COVER_POOL, caps[1])
        gains = torch.zeros(P, dtype=torch.int32, device="cuda:0")
        g.pair_gains(False, gains)
        got = gains.cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(got, want)


def test_cover_knob_limits():
    nl = pg.Netlist.generate(small_config(5))
    with pg.Grader(nl) as g:
        for key, bad in ((pg.TUNE_COVER_NODES, 0), (pg.TUNE_COVER_NODES, 1281), (pg.TUNE_COVER_POOL, -1),
                         (pg.TUNE_COVER_POOL, 4097)):
            with pytest.raises(pg.PdfError):
                g.set_tuning(key, bad)
