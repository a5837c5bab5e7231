"""Two independent CPU graders must agree: the literal byte-per-pattern oracle
(oracle.c) and the bit-sliced OpenMP grader (bitsliced.c), on random circuits,
the every-gate-type edge netlist, ragged pattern counts, both modes, explicit
pairs and several thread counts."""
from __future__ import annotations

import numpy as np
import pytest

import pdfgrade as pg
from helpers import biased_vectors, edge_netlist, small_config
from oracle import oracle as orc


def _agree(nl, bits, P, chunk, robust, b_bits=None, threads=1):
    net = orc.OracleNet(nl.arrays())
    ref = orc.grade(net, bits, P, chunk, robust, b_bits=b_bits)
    got = orc.grade_bitsliced(net, bits, P, chunk, robust, b_bits=b_bits, threads=threads)
    assert got["total"] == ref["total"]
    assert got["flags"] == ref["flags"] & orc.FLAG_U64_OVERFLOW
    for k in ("first_t0", "first_t1", "gate_counts"):
        assert np.array_equal(got[k], ref[k]), k
    return ref["total"]


def test_bitsliced_matches_oracle_random():
    totals = []
    for seed in range(10):
        gates, levels = (24, 5) if seed % 2 else (60, 7)
        nl = pg.Netlist.generate(small_config(seed, num_gates=gates, levels=levels, wide=seed % 4 < 2,
                                              alpha=2.2 if seed % 3 == 0 else 0.0))
        P = 64 * 5 + 37 * (seed % 3)  # ragged tails on two seeds in three
        bits = biased_vectors(seed, nl.info()["num_pi"], P, p_hold=0.5)
        totals += [_agree(nl, bits, P, chunk, robust, threads=1 + seed % 4)
                   for robust in (True, False) for chunk in (64, 128, P)]
    assert sum(t > 0 for t in totals) >= len(totals) // 2, totals


@pytest.mark.parametrize("robust", [True, False])
def test_bitsliced_edge_netlist_and_explicit_pairs(robust):
    nl = edge_netlist()
    P = 300
    a = biased_vectors(9, 8, P, p_hold=0.5)
    b = biased_vectors(10, 8, P, p_hold=0.3)
    assert _agree(nl, a, P, 128, robust, threads=3) > 0
    _agree(nl, a, P, 64, robust, b_bits=b, threads=2)


def test_bitsliced_larger_circuit_threads():
    cfg = pg.GenConfig(seed=21, num_pi=64, num_po=64, num_gates=1500, levels=24, fanin_p=(0.1, 0.5, 0.3, 0.1),
                       prev_level_frac=0.7, fanout_alpha=2.2, fanout_cap=16)
    nl = pg.Netlist.generate(cfg)
    P = 4096 + 19
    bits = biased_vectors(21, 64, P, p_hold=0.6)
    for threads in (1, 8):
        assert _agree(nl, bits, P, 1024, True, threads=threads) > 0


def test_bitsliced_rejects_unaligned_chunks():
    nl = pg.Netlist.generate(small_config(1))
    bits = biased_vectors(1, nl.info()["num_pi"], 200)
    with pytest.raises(RuntimeError):
        orc.grade_bitsliced(orc.OracleNet(nl.arrays()), bits, 200, 100, True)
