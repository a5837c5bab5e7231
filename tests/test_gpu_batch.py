"""Batched grading: one captured graph over many circuits equals grading each
circuit alone (the oracle for small circuits, pdf_grade for larger ones, and
the bit-sliced CPU grader for every circuit of the 256-circuit workload)."""
from __future__ import annotations

import numpy as np
import pytest

import pdfgrade as pg
from oracle import oracle as orc
from helpers import bitsliced_grade, small_config

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _wide_cfg(seed, gates, levels, num_pi=64, num_po=32):
    return pg.GenConfig(seed=seed, num_pi=num_pi, num_po=num_po, num_gates=gates, levels=levels,
                        fanin_p=(0.1, 0.6, 0.2, 0.1), prev_level_frac=0.7, fanout_alpha=2.2, fanout_cap=64,
                        path_limit_log2=0)


@pytest.mark.parametrize("robust", [True, False])
def test_small_batch_matches_oracle(robust):
    """Narrow tiles (fewer than 64 words), different depths, ragged P."""
    _say("small batch", robust)
    specs = [(small_config(301, num_pi=10, num_gates=60, levels=6), 256),
             (small_config(302, num_pi=16, num_gates=200, levels=12, alpha=2.2), 1000),
             (small_config(303, num_pi=8, num_gates=30, levels=4), 64),
             (small_config(304, num_pi=24, num_gates=400, levels=15, alpha=2.2), 2048),
             (small_config(305, num_pi=12, num_gates=120, levels=9), 777)]
    nets = [pg.Netlist.generate(c) for c, _ in specs]
    vecs = [pg.gen_vectors(400 + i, c.num_pi, p) for i, (c, p) in enumerate(specs)]
    b = pg.Batch(nets, vecs, [p for _, p in specs])
    wants = [orc.grade(orc.OracleNet(nl.arrays()), v, p, p, robust) for nl, v, (_, p) in zip(nets, vecs, specs)]
    for compact in (1, 0):  # reverse pass over the marked-item list, then over every item
        b.set_tuning(pg.TUNE_REV_COMPACT, compact)
        tot, ms = b.grade(robust)
        for i, want in enumerate(wants):
            assert tot[i] == want["total"], (compact, i)
            m = b.member(i)
            f0, f1 = m.first()
            np.testing.assert_array_equal(f0, want["first_t0"])
            np.testing.assert_array_equal(f1, want["first_t1"])
            np.testing.assert_array_equal(m.gate_counts(), want["gate_counts"])
            assert m.total() == want["total"]
        assert any(t > 0 for t in tot)
    # lists built by the forward launches need whole-wave tiles (>= 64 words)
    with pytest.raises(pg.PdfError):
        b.set_tuning(pg.TUNE_REV_COMPACT, 2)
    b.close()


def _plain(nl, v, p, robust):
    g = pg.Grader(nl)
    g.set_vectors(v, p)
    g.set_chunk(p)
    t = g.grade(robust)
    out = (t, *g.first(), g.gate_counts())
    g.close()
    return out


def test_wide_batch_matches_plain_grading_and_tuning():
    """Circuits of 2k-30k gates with 4k-16k pairs (whole-wave tiles): equal to
    per-circuit pdf_grade under every tuning and on graph replays."""
    _say("wide batch")
    rng = np.random.default_rng(7)
    specs = []
    for i in range(10):
        gates = int(np.exp(rng.uniform(np.log(2000), np.log(30000))))
        levels = int(round(8 * np.log2(gates)))
        specs.append((_wide_cfg(3100 + i, gates, levels), int(rng.integers(64, 257)) * 64))
    nets = [pg.Netlist.generate(c) for c, _ in specs]
    vecs = [pg.gen_vectors(13100 + i, c.num_pi, p) for i, (c, p) in enumerate(specs)]
    want = [_plain(nl, v, p, True) for nl, v, (_, p) in zip(nets, vecs, specs)]
    b = pg.Batch(nets, vecs, [p for _, p in specs])
    for order in (0, 1):
        for waves in (64, 1024):
            for graph in (0, 1):
                for compact, wpl in ((0, 1), (1, 1), (2, 1), (0, 2), (1, 2), (2, 2)):
                    b.set_tuning(pg.TUNE_SEGMENT_ORDER, order)
                    b.set_tuning(pg.TUNE_COVER_WAVES, waves)
                    b.set_tuning(pg.TUNE_USE_GRAPH, graph)
                    b.set_tuning(pg.TUNE_REV_COMPACT, compact)
                    b.set_tuning(pg.TUNE_WORDS_PER_LANE, wpl)
                    ctx = f"order {order} waves {waves} graph {graph} compact {compact} wpl {wpl}"
                    for rep in range(2):
                        tot, _ = b.grade(True)
                        np.testing.assert_array_equal(tot, np.array([w[0] for w in want], np.uint64),
                                                      err_msg=f"{ctx} rep {rep}")
                    for i, w in enumerate(want):
                        m = b.member(i)
                        f0, f1 = m.first()
                        np.testing.assert_array_equal(f0, w[1], err_msg=ctx)
                        np.testing.assert_array_equal(f1, w[2], err_msg=ctx)
                        np.testing.assert_array_equal(m.gate_counts(), w[3], err_msg=ctx)
    best = b.autotune(True, reps=2)
    _say("autotune", best)
    assert best["cover_waves"] in (256, 512, 1024, 2048) and best["rev_compact"] in (0, 1, 2) and best["ms"] > 0
    assert best["words_per_lane"] in (1, 2)
    tot, _ = b.grade(True)
    np.testing.assert_array_equal(tot, np.array([w[0] for w in want], np.uint64))
    b.close()


@pytest.mark.parametrize("robust", [True, False])
def test_wide_batch_ragged_pairs_two_words_per_lane(robust):
    """Pair counts that are not multiples of 64 and word counts that are not
    multiples of 128 (the second half of a 128-word tile partly or wholly
    past the end): every reverse and forward mode equals pdf_grade."""
    _say("ragged wide batch", robust)
    specs = [(_wide_cfg(3301, 3000, 92), 4305), (_wide_cfg(3302, 6000, 100), 8191),
             (_wide_cfg(3303, 2500, 90), 4096 + 64 * 70 + 5), (_wide_cfg(3304, 9000, 105), 12345)]
    nets = [pg.Netlist.generate(c) for c, _ in specs]
    vecs = [pg.gen_vectors(13300 + i, c.num_pi, p) for i, (c, p) in enumerate(specs)]
    want = [_plain(nl, v, p, robust) for nl, v, (_, p) in zip(nets, vecs, specs)]
    assert any(w[0] > 0 for w in want)
    b = pg.Batch(nets, vecs, [p for _, p in specs])
    for wpl in (2, 1):
        for compact in (2, 1, 0):
            b.set_tuning(pg.TUNE_WORDS_PER_LANE, wpl)
            b.set_tuning(pg.TUNE_REV_COMPACT, compact)
            tot, _ = b.grade(robust)
            ctx = f"wpl {wpl} compact {compact}"
            np.testing.assert_array_equal(tot, np.array([w[0] for w in want], np.uint64), err_msg=ctx)
            for i, w in enumerate(want):
                m = b.member(i)
                f0, f1 = m.first()
                np.testing.assert_array_equal(f0, w[1], err_msg=ctx)
                np.testing.assert_array_equal(f1, w[2], err_msg=ctx)
                np.testing.assert_array_equal(m.gate_counts(), w[3], err_msg=ctx)
    with pytest.raises(pg.PdfError):
        b.set_tuning(pg.TUNE_WORDS_PER_LANE, 3)
    with pytest.raises(pg.PdfError):
        b.set_tuning(pg.TUNE_REV_COMPACT, 3)
    b.close()


def test_batch_non_robust_mixed_widths():
    _say("mixed widths non-robust")
    specs = [(_wide_cfg(3201, 5000, 98), 8192), (small_config(3202, num_pi=20, num_gates=300, levels=10), 512),
             (_wide_cfg(3203, 12000, 108), 4096)]
    nets = [pg.Netlist.generate(c) for c, _ in specs]
    vecs = [pg.gen_vectors(13200 + i, c.num_pi, p) for i, (c, p) in enumerate(specs)]
    want = [_plain(nl, v, p, False) for nl, v, (_, p) in zip(nets, vecs, specs)]
    b = pg.Batch(nets, vecs, [p for _, p in specs])
    tot, _ = b.grade(False)
    np.testing.assert_array_equal(tot, np.array([w[0] for w in want], np.uint64))
    for i, w in enumerate(want):
        np.testing.assert_array_equal(b.member(i).gate_counts(), w[3])
    b.close()


def test_batch_256_circuits_against_bitsliced():
    """The whole 256-circuit workload (1k-200k gates, 4k-64k pairs each), one
    batched grade, every circuit's total, first-detection arrays and per-node
    counts bit-exact against the bit-sliced CPU grader."""
    from pdfgrade.configs import BATCH_256
    _say("batch 256 vs bit-sliced")
    members = BATCH_256.members()
    nets = [pg.Netlist.generate(c) for c, _, _ in members]
    vecs = [pg.gen_vectors(v, c.num_pi, p) for c, v, p in members]
    pairs = [p for _, _, p in members]
    b = pg.Batch(nets, vecs, pairs)
    tot, _ = b.grade(True)
    checked = []
    for i, (nl, v, p) in enumerate(zip(nets, vecs, pairs)):
        ref = bitsliced_grade(nl, v, p, 16384, True)
        m = b.member(i)
        f0, f1 = m.first()
        assert int(tot[i]) == ref["total"], i
        np.testing.assert_array_equal(f0, ref["first_t0"], err_msg=str(i))
        np.testing.assert_array_equal(f1, ref["first_t1"], err_msg=str(i))
        np.testing.assert_array_equal(m.gate_counts(), ref["gate_counts"], err_msg=str(i))
        checked.append((int(tot[i]), f0, f1, m.gate_counts()))
        if i % 32 == 31:
            _say("checked", i + 1)
    assert int(tot.sum()) > 0
    # the default reverse pass runs over lists the forward launches build;
    # over the list a scan launch builds: the same bit-exact results
    b.set_tuning(pg.TUNE_REV_COMPACT, 1)
    _say("batch 256, scan-built reverse lists")
    tot, _ = b.grade(True)
    for i, (t_i, f0, f1, counts) in enumerate(checked):
        m = b.member(i)
        g0, g1 = m.first()
        assert int(tot[i]) == t_i, i
        np.testing.assert_array_equal(g0, f0, err_msg=str(i))
        np.testing.assert_array_equal(g1, f1, err_msg=str(i))
        np.testing.assert_array_equal(m.gate_counts(), counts, err_msg=str(i))
    b.close()
