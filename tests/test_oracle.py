"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

The reference ships no tests, fixtures or golden vectors for this path, so the
oracle is pinned by: closed forms of its tables, an independent brute-force
restatement (oracle/brute.py) on many small random circuits, a hand-derived
c17 case, invariants, and planted mistakes that those checks must catch.
"""
import itertools

import numpy as np
import pytest

import pdfgrade as pg
from oracle import brute
from oracle import oracle as orc

from helpers import C17_BENCH, biased_vectors, brute_grade, edge_netlist, oracle_grade, pack, small_config

CODE_FRAMES = {pg.S0: (0, 0), pg.S1: (1, 1), pg.T0: (1, 0), pg.T1: (0, 1)}


def frames_to_code(a, b):
    return b | ((a ^ b) << 1)


@pytest.mark.parametrize("which,fn", [
    (0, lambda x, y: x & y), (1, lambda x, y: 1 - (x & y)), (2, lambda x, y: x | y),
    (3, lambda x, y: 1 - (x | y)), (4, lambda x, y: x ^ y), (5, lambda x, y: 1 - (x ^ y))])
def test_gate_tables_are_two_frame_boolean(which, fn):
    """simkernel.cuh:158-163: every entry equals the 2-input function applied to
    each frame, read as tex2D(lut, val, r) = lut[r*4 + val]."""
    t = orc.table(which, 16)
    for val, r in itertools.product(range(4), range(4)):
        (va, vb), (ra, rb) = CODE_FRAMES[val], CODE_FRAMES[r]
        assert t[r * 4 + val] == frames_to_code(fn(va, ra), fn(vb, rb)), (which, val, r)
        assert t[r * 4 + val] == t[val * 4 + r]  # symmetric: operand order of tex2D is moot


def test_stable_and_not_tables():
    st = orc.table(6, 4)
    for v1, v2 in itertools.product((0, 1), (0, 1)):
        assert st[v2 * 2 + v1] == frames_to_code(v1, v2)
    nt = orc.table(7, 4)
    for c in range(4):
        a, b = CODE_FRAMES[c]
        assert nt[c] == frames_to_code(1 - a, 1 - b)


@pytest.mark.parametrize("robust", [True, False])
def test_mark_luts_match_verbal_definition(robust):
    """markkernel.cuh:36-41 bit (on*4+off) against brute.sensitizes()."""
    for idx, ty in enumerate((pg.AND, pg.NAND, pg.OR, pg.NOR, pg.XOR, pg.XNOR)):
        lut = orc.mark_lut(robust, idx)
        for on, off in itertools.product(range(4), range(4)):
            bit = (lut >> (on * 4 + off)) & 1
            assert bit == brute.sensitizes(robust, ty, CODE_FRAMES[on], CODE_FRAMES[off]), (robust, ty, on, off)


def c17():
    import tempfile, os
    with tempfile.NamedTemporaryFile("w", suffix=".bench", delete=False) as f:
        f.write(C17_BENCH)
        path = f.name
    try:
        return pg.Netlist.read_bench(path)
    finally:
        os.unlink(path)


def c17_pairs(first_a, first_b):
    """64 explicit pairs: pair 0 = (first_a, first_b), the rest hold first_b."""
    a = np.array([first_a] + [first_b] * 63, np.uint8)
    b = np.array([first_b] * 64, np.uint8)
    return pack(a), pack(b)


@pytest.mark.parametrize("robust", [True, False])
def test_c17_known_answer(robust):
    """Hand-derived.  Ids: inputs 1,2,3,6,7 -> 0..4; 10,11 -> 5,6; 16,19 -> 7,8;
    22,23 -> 9,10.  Pair 0: input 1 rises, inputs 2,3,6,7 stay 1.  Then
    10 = T0, 11 = S0, 16 = S1, 19 = S1, 22 = T1, 23 = S0; the path 1 -> 10 -> 22
    is robust (off-path inputs 3 and 16 are S1) and is the only marked path;
    all three nodes are first detected at pair 0, so c(22)=c(10)=c(1)=1 and the
    total is 1."""
    nl = c17()
    a, b = c17_pairs([0, 1, 1, 1, 1], [1, 1, 1, 1, 1])
    r = oracle_grade(nl, a, 64, 64, robust, b_bits=b)
    assert r["total"] == 1
    exp_counts = np.zeros(11, np.uint64)
    exp_counts[[0, 5, 9]] = 1
    assert np.array_equal(r["gate_counts"], exp_counts)
    none = 0xFFFFFFFF
    exp_t0 = np.full(11, none, np.uint32)
    exp_t1 = np.full(11, none, np.uint32)
    exp_t1[0] = 0
    exp_t0[5] = 0
    exp_t1[9] = 0
    assert np.array_equal(r["first_t0"], exp_t0) and np.array_equal(r["first_t1"], exp_t1)


def test_c17_on_path_is_fanin0_only():
    """Input 3 rises with the others at 1.  11 = NAND(3,6) is marked (fanin 0 is
    input 3), but 16 = NAND(2,11) is not: its fanin 0 (input 2) is unmarked
    (markkernel.cuh:254), so nothing reaches a PO and every mark is removed."""
    nl = c17()
    a, b = c17_pairs([1, 1, 0, 1, 1], [1, 1, 1, 1, 1])
    net = orc.OracleNet(nl.arrays())
    codes = orc.sim_chunk(net, a, 64, 0, 64, b_bits=b)
    assert list(codes[:, 0]) == [pg.S1, pg.S1, pg.T1, pg.S1, pg.S1, pg.T0, pg.T0, pg.T1, pg.T1, pg.S1, pg.T0]
    marks, fwd = orc.mark_chunk(net, codes, True)
    assert list(fwd[:, 0]) == [0, 0, 1, 0, 0, 0, 1, 0, 1, 0, 0]
    assert marks.sum() == 0
    assert oracle_grade(nl, a, 64, 64, True, b_bits=b)["total"] == 0


def _compare_with_brute(nl, bits, P, robust, chunk, b_bits=None):
    ref = brute_grade(nl, bits, P, robust, b_bits=b_bits)
    got = oracle_grade(nl, bits, P, chunk, robust, b_bits=b_bits)
    assert got["total"] == ref["total"]
    assert list(got["gate_counts"]) == ref["gate_counts"]
    assert list(got["first_t0"]) == ref["first_t0"]
    assert list(got["first_t1"]) == ref["first_t1"]
    # per-pattern intermediates on the first chunk
    net = orc.OracleNet(nl.arrays())
    w = min(chunk, P)
    codes = orc.sim_chunk(net, bits, P, 0, w, b_bits=b_bits)
    marks, _ = orc.mark_chunk(net, codes, robust)
    for p in range(w):
        fr = ref["frames"][p]
        assert [frames_to_code(a, b) for a, b in fr] == list(codes[:, p])
        assert ref["marks"][p] == list(marks[:, p])
    return ref["total"]


@pytest.mark.parametrize("seed", range(12))
def test_oracle_matches_brute_force(seed):
    nl = pg.Netlist.generate(small_config(seed, wide=seed % 2 == 0, alpha=2.2 if seed % 3 == 0 else 0.0))
    P = 128
    bits = biased_vectors(seed, nl.info()["num_pi"], P, p_one=0.5, p_hold=0.5)
    for robust in (True, False):
        _compare_with_brute(nl, bits, P, robust, chunk=64 if seed % 2 else 128)


@pytest.mark.parametrize("robust", [True, False])
def test_edge_netlist_matches_brute_force(robust):
    """Every gate type with one and several fanins, 12-input gates, POs with
    fanouts and an unused PI."""
    P = 256
    bits = biased_vectors(9, 8, P, p_hold=0.5)
    assert _compare_with_brute(edge_netlist(), bits, P, robust, chunk=128) > 0


def test_oracle_brute_force_has_coverage():
    """The random cases above must exercise non-zero coverage."""
    totals = []
    for seed in range(12):
        nl = pg.Netlist.generate(small_config(seed, wide=seed % 2 == 0))
        bits = biased_vectors(seed, nl.info()["num_pi"], 128, p_hold=0.5)
        totals.append(oracle_grade(nl, bits, 128, 128, True)["total"])
    assert sum(t > 0 for t in totals) >= 9, totals


def test_explicit_pairs_equal_sequence_mode():
    nl = pg.Netlist.generate(small_config(7))
    P = 192
    bits = biased_vectors(3, nl.info()["num_pi"], P, p_hold=0.4)
    vals = np.roll(np.asarray(__import__("helpers").unpack(bits, P)), -1, axis=0)
    seq = oracle_grade(nl, bits, P, 64, True)
    exp = oracle_grade(nl, bits, P, 64, True, b_bits=pack(vals))
    for k in ("total", "flags"):
        assert seq[k] == exp[k]
    for k in ("gate_counts", "first_t0", "first_t1", "chunk_totals"):
        assert np.array_equal(seq[k], exp[k])


def test_chunk_size_invariance_and_constant_vectors():
    nl = pg.Netlist.generate(small_config(11, num_gates=60, levels=8))
    P = 256
    bits = biased_vectors(11, nl.info()["num_pi"], P, p_hold=0.5)
    outs = [oracle_grade(nl, bits, P, c, True) for c in (64, 128, 192, 256)]
    for o in outs[1:]:
        assert o["total"] == outs[0]["total"]
        assert np.array_equal(o["gate_counts"], outs[0]["gate_counts"])
        assert np.array_equal(o["first_t0"], outs[0]["first_t0"])
    const = np.zeros_like(bits)
    assert oracle_grade(nl, const, P, 64, True)["total"] == 0


def test_mark_invariants():
    nl = pg.Netlist.generate(small_config(5, num_gates=80, levels=9))
    arrs = nl.arrays()
    net = orc.OracleNet(arrs)
    P = 128
    bits = biased_vectors(5, arrs["num_pi"], P, p_hold=0.6)
    codes = orc.sim_chunk(net, bits, P, 0, P)
    marks, fwd = orc.mark_chunk(net, codes, True)
    assert np.all(marks <= fwd)
    assert np.all(marks[codes < 2] == 0)  # marks only on transitions
    for g in range(arrs["n"]):
        if arrs["po"][g]:
            continue
        fos = arrs["fout"][arrs["fout_off"][g]:arrs["fout_off"][g + 1]]
        has = marks[fos].max(axis=0) if len(fos) else np.zeros(P, np.uint8)
        assert np.all(marks[g] <= has)


def test_planted_mistakes_are_caught():
    """Each plausible slip in the cover DP changes at least one brute-force
    comparison: they are planted in a Python copy of the DP run on the
    oracle's marks."""
    P = 128
    for seed in range(20):  # first circuit whose coverage involves a PO-internal path
        nl = pg.Netlist.generate(small_config(seed, num_gates=40, levels=6))
        arrs = nl.arrays()
        bits = biased_vectors(seed, arrs["num_pi"], P, p_hold=0.6)
        ref = brute_grade(nl, bits, P, True)
        if ref["total"] >= 4:
            break
    assert ref["total"] >= 4
    net = orc.OracleNet(arrs)
    codes = orc.sim_chunk(net, bits, P, 0, P)
    marks, _ = orc.mark_chunk(net, codes, True)
    f0 = np.full(arrs["n"], 0xFFFFFFFF, np.uint32)
    f1 = f0.copy()
    orc.merge_chunk(net, codes, marks, 0, f0, f1)

    def dp(variant):
        n = arrs["n"]
        counts = [0] * n
        total = 0
        for p in range(P):
            c = [0] * n
            h = [0] * n
            for g in reversed(range(n)):
                m = int(marks[g, p])
                fos = arrs["fout"][arrs["fout_off"][g]:arrs["fout_off"][g + 1]]
                C = sum(c[f] for f in fos)
                H = (m if arrs["po"][g] else 0) + sum(h[f] for f in fos)
                new = f0[g] == p or f1[g] == p
                if variant == "new_ignores_T1":
                    new = f0[g] == p
                if variant == "no_history":
                    c[g] = m * C
                else:
                    c[g] = m * (C + (H if new else 0))
                h[g] = m * (0 if new else H)
                if variant == "keep_history_when_new":
                    h[g] = m * H
                if arrs["type"][g] == pg.INPT:
                    h[g] = 0
                counts[g] += c[g]
            total += sum(c[g] for g in range(arrs["num_pi"]))
        return total, counts

    assert dp("none") == (ref["total"], ref["gate_counts"])
    for v in ("new_ignores_T1", "no_history", "keep_history_when_new"):
        assert dp(v) != (ref["total"], ref["gate_counts"]), v


@pytest.mark.parametrize("seed", range(6))
def test_pi_cover_matches_brute_force(seed):
    """c(g, p) of every PI per pattern (the values of the reference's
    gpucover.log) from the oracle's cover, chunk by chunk, equals the
    brute-force count of marked PI->PO paths through a new node."""
    from helpers import brute_grade
    nl = pg.Netlist.generate(small_config(seed, wide=seed % 2 == 0))
    arrs = nl.arrays()
    I = arrs["num_pi"]
    P, W = 128, 64
    bits = biased_vectors(seed, I, P, p_hold=0.5)
    for robust in (True, False):
        ref = brute_grade(nl, bits, P, robust)
        net = orc.OracleNet(arrs)
        f0 = np.full(net.n, 0xFFFFFFFF, np.uint32)
        f1 = f0.copy()
        gc = np.zeros(net.n, np.uint64)
        for start in range(0, P, W):
            codes = orc.sim_chunk(net, bits, P, start, W)
            marks, _ = orc.mark_chunk(net, codes, robust)
            orc.merge_chunk(net, codes, marks, start, f0, f1)
            newm = orc.new_from_firsts(net.n, W, start, f0, f1)
            tot, _, pic = orc.cover_new_chunk(net, marks, newm, gc, I)
            for p in range(W):
                want = [ref["node_counts"][start + p][g] for g in range(I)]
                assert pic[:, p].tolist() == want, (robust, start + p)
                assert int(pic[:, p].sum()) == ref["per_pattern"][start + p]
        assert gc.tolist() == ref["gate_counts"]
