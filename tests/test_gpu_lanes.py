"""Chunk lanes: consecutive chunks graded concurrently on several streams give
exactly the one-lane results."""
from __future__ import annotations

import numpy as np
import pytest

import pdfgrade as pg
from oracle import oracle as orc
from helpers import small_config

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["seq", "pairs"])
def test_lanes_match_oracle(mode):
    nl = pg.Netlist.generate(small_config(501, num_pi=24, num_gates=500, levels=12, alpha=2.2))
    P = 5000 if mode == "pairs" else 4096
    a = pg.gen_vectors(502, 24, P)
    b = pg.gen_vectors(503, 24, P) if mode == "pairs" else None
    g = pg.Grader(nl)
    if b is None:
        g.set_vectors(a, P)
    else:
        g.set_pairs(a, b, P)
    for chunk in (512, 1024):
        want = orc.grade(orc.OracleNet(nl.arrays()), a, P, chunk, True, b_bits=b)
        for lanes in (1, 2, 3, 4):
            for graph in (0, 1):
                print("chunk", chunk, "lanes", lanes, "graph", graph, flush=True)
                g.set_tuning(pg.TUNE_LANES, lanes)
                g.set_chunk(chunk)
                g.set_tuning(pg.TUNE_USE_GRAPH, graph)
                for rep in range(2):
                    assert g.grade(True) == want["total"], (chunk, lanes, graph, rep)
                f0, f1 = g.first()
                np.testing.assert_array_equal(f0, want["first_t0"])
                np.testing.assert_array_equal(f1, want["first_t1"])
                np.testing.assert_array_equal(g.gate_counts(), want["gate_counts"])
    # the stage API and the timed grade keep working on lane 0
    tot, kt = g.grade_timed(True)
    assert tot == want["total"]
    g.close()


def test_lanes_50k_equal_one_lane():
    from pdfgrade.configs import GATES_50K as W
    nl = pg.Netlist.generate(W.netlist)
    I = nl.info()["num_pi"]
    bits = pg.gen_vectors(W.vector_seed, I, W.pairs)
    g = pg.Grader(nl)
    g.set_vectors(bits, W.pairs)
    g.set_chunk(W.pairs)
    ref = g.grade(True)
    ref_gc = g.gate_counts()
    for lanes, chunk in ((2, 32768), (4, 16384), (2, 16384)):
        g.set_tuning(pg.TUNE_LANES, lanes)
        g.set_chunk(chunk)
        for graph in (1, 0):
            g.set_tuning(pg.TUNE_USE_GRAPH, graph)
            assert g.grade(True) == ref, (lanes, chunk, graph)
            np.testing.assert_array_equal(g.gate_counts(), ref_gc)
    g.close()
