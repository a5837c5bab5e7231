"""Segments of length K >= 2 in the oracle (orc_segments, orc_grade_segs),
pinned against the brute-force grader (oracle/brute.py: its own walk
enumerator, per-pattern segment detection, and coverage by enumerating every
marked PI->PO path), closed-form segment counts, and K = 1 reducing to the
per-node path."""
from __future__ import annotations

import numpy as np
import pytest

import pdfgrade as pg
from helpers import biased_vectors, edge_netlist, netlist_lists, small_config, unpack
from oracle import brute
from oracle import oracle as orc

NONE = 0xFFFFFFFF


def _fanouts(arrs):
    fo, f = arrs["fout_off"], arrs["fout"]
    return [[int(x) for x in f[fo[g]:fo[g + 1]]] for g in range(arrs["n"])]


def _walk_count(fanouts, K):
    """count_K(g) = 1 if K == 1 or g has no fanouts, else sum over fanouts of
    count_{K-1}: the number of segments starting at g."""
    n = len(fanouts)
    cnt = [1] * n
    for _ in range(K - 1):
        cnt = [1 if not fanouts[g] else sum(cnt[f] for f in fanouts[g]) for g in range(n)]
    return sum(cnt)


@pytest.mark.parametrize("seed", range(6))
def test_segment_enumeration(seed):
    nl = pg.Netlist.generate(small_config(seed, num_gates=40, levels=6, wide=seed % 2 == 0,
                                          alpha=2.2 if seed % 3 == 0 else 0.0))
    arrs = nl.arrays()
    net = orc.OracleNet(arrs)
    fanouts = _fanouts(arrs)
    assert np.array_equal(orc.segments(net, 1)[:, 0], np.arange(arrs["n"], dtype=np.uint32))
    sinks = sum(1 for f in fanouts if not f)
    assert orc.segments(net, 2).shape[0] == len(arrs["fout"]) + sinks
    for K in (2, 3, 5, 8):
        got = orc.segments(net, K)
        want = [[NONE if v is None else v for v in w] for w in brute.segment_walks(fanouts, K)]
        assert got.tolist() == want, K
        assert got.shape[0] == _walk_count(fanouts, K)
        # every entry after the first padding is padding; consecutive nodes are edges
        for row in got.tolist()[:200]:
            k = row.index(NONE) if NONE in row else K
            assert all(v == NONE for v in row[k:])
            assert all(row[i + 1] in fanouts[row[i]] for i in range(k - 1))
            assert k == K or not fanouts[row[k - 1]]


def test_segment_length_bounds():
    net = orc.OracleNet(pg.Netlist.generate(small_config(1)).arrays())
    for K in (0, 9):
        with pytest.raises(ValueError):
            orc.segments(net, K)


def test_k1_equals_node_path():
    nl = pg.Netlist.generate(small_config(4, num_gates=60, levels=8))
    P = 256
    bits = biased_vectors(4, nl.info()["num_pi"], P, p_hold=0.5)
    net = orc.OracleNet(nl.arrays())
    for robust in (True, False):
        ref = orc.grade(net, bits, P, 64, robust)
        seg = orc.grade_segs(net, bits, P, 64, robust, 1)
        assert seg["total"] == ref["total"] and seg["flags"] == ref["flags"]
        assert np.array_equal(seg["seg_t0"], ref["first_t0"]) and np.array_equal(seg["seg_t1"], ref["first_t1"])
        assert np.array_equal(seg["gate_counts"], ref["gate_counts"])
        assert np.array_equal(seg["chunk_totals"], ref["chunk_totals"])


def _brute(nl, bits, P, robust, K):
    arrs = nl.arrays()
    types, fanins, po = netlist_lists(arrs)
    va = unpack(bits, P)
    vb = np.roll(va, -1, axis=0)
    return brute.grade(types, fanins, po, arrs["num_pi"], va.tolist(), vb.tolist(), robust, segs=K)


@pytest.mark.parametrize("seed", range(8))
def test_segments_match_brute_force(seed):
    nl = pg.Netlist.generate(small_config(seed, wide=seed % 2 == 0, alpha=2.2 if seed % 3 == 0 else 0.0))
    P = 128
    bits = biased_vectors(seed, nl.info()["num_pi"], P, p_hold=0.5)
    net = orc.OracleNet(nl.arrays())
    for K in (2, 3, 4):
        for robust in (True, False):
            ref = _brute(nl, bits, P, robust, K)
            got = orc.grade_segs(net, bits, P, 64 if seed % 2 else 128, robust, K)
            ctx = f"K={K} robust={robust}"
            assert got["segments"].tolist() == ref["segments"], ctx
            assert got["seg_t0"].tolist() == ref["seg_t0"], ctx
            assert got["seg_t1"].tolist() == ref["seg_t1"], ctx
            assert got["gate_counts"].tolist() == ref["gate_counts"], ctx
            assert got["total"] == ref["total"], ctx


@pytest.mark.parametrize("robust", [True, False])
def test_segments_edge_netlist(robust):
    nl = edge_netlist()
    P = 192
    bits = biased_vectors(9, 8, P, p_hold=0.5)
    net = orc.OracleNet(nl.arrays())
    for K in (2, 3):
        ref = _brute(nl, bits, P, robust, K)
        got = orc.grade_segs(net, bits, P, 64, robust, K)
        assert got["seg_t0"].tolist() == ref["seg_t0"] and got["seg_t1"].tolist() == ref["seg_t1"]
        assert got["gate_counts"].tolist() == ref["gate_counts"] and got["total"] == ref["total"]


def test_segments_have_coverage_and_differ_from_k1():
    """The cases above must exercise non-zero coverage, and K = 2 and K = 3
    must change results somewhere (otherwise the comparison could not tell
    them from K = 1)."""
    totals, differs = [], 0
    for seed in range(8):
        nl = pg.Netlist.generate(small_config(seed, wide=seed % 2 == 0))
        bits = biased_vectors(seed, nl.info()["num_pi"], 128, p_hold=0.5)
        net = orc.OracleNet(nl.arrays())
        k1 = orc.grade(net, bits, 128, 128, True)
        for K in (2, 3):
            kk = orc.grade_segs(net, bits, 128, 128, True, K)
            totals.append(kk["total"])
            differs += not np.array_equal(kk["gate_counts"], k1["gate_counts"])
    assert sum(t > 0 for t in totals) >= 10, totals
    assert differs >= 4


def test_planted_polarity_mistake_is_caught():
    """Taking a segment's polarity from its last node instead of node 0 must
    give different first patterns than the oracle on these cases."""
    caught = 0
    for seed in range(8):
        nl = pg.Netlist.generate(small_config(seed, wide=seed % 2 == 0))
        P = 128
        bits = biased_vectors(seed, nl.info()["num_pi"], P, p_hold=0.5)
        ref = _brute(nl, bits, P, False, 3)
        wrong0 = []
        for wk in ref["segments"]:
            last = max(i for i, v in enumerate(wk) if v != NONE)
            first = NONE
            for p in range(P):
                if all(ref["marks"][p][v] for v in wk if v != NONE) and ref["frames"][p][wk[last]][1] == 0:
                    first = p
                    break
            wrong0.append(first)
        got = orc.grade_segs(orc.OracleNet(nl.arrays()), bits, P, P, False, 3)
        caught += wrong0 != got["seg_t0"].tolist()
    assert caught >= 1
