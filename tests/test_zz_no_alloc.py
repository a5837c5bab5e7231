"""pdf_grade_noalloc: the fused grade with every chunk buffer in the library's
static per-GPU scratch array, checked against pdf_grade on the same grader
(total, first-detection pairs, per-gate counts), on two calls in a row with
the second on a smaller input, so the second call runs over the first
call's leftovers in the shared array."""
import sys

import numpy as np
import pytest
import torch

import pdfgrade as pg

from helpers import biased_vectors, small_config

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _state(g):
    f0, f1 = g.first()
    return g.total(), f0, f1, g.gate_counts()


def _same(a, b, ctx):
    assert a[0] == b[0], (ctx, a[0], b[0])
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y), ctx


def _cases():
    # 20,000 gates x 65,536 pairs does not fit the scratch in one chunk, so
    # the first call grades several scratch-sized chunks; the second input is
    # smaller in every dimension (one chunk)
    big = pg.GenConfig(seed=21, num_pi=128, num_po=64, num_gates=20_000, levels=40, fanin_p=(0.1, 0.6, 0.2, 0.1),
                       prev_level_frac=0.7, fanout_alpha=2.2, fanout_cap=64)
    small = small_config(22, num_pi=12, num_po=16, num_gates=600, levels=12, wide=True)
    return [("big", big, 65_536, 0.6), ("small", small, 2_048, 0.7)]


def test_noalloc_grade_equals_grade_twice():
    total = torch.zeros(1, dtype=torch.int64, device="cuda")
    err = torch.ones(1, dtype=torch.int32, device="cuda")
    s = torch.cuda.Stream()
    for name, cfg, P, hold in _cases():
        nl = pg.Netlist.generate(cfg)
        I = nl.info()["num_pi"]
        bits = biased_vectors(cfg.seed, I, P, p_hold=hold)
        with pg.Grader(nl) as g:
            g.set_vectors(bits, P)
            for robust in (True, False):
                say(f"noalloc {name}: n={nl.info()['n']} pairs={P} robust={robust}")
                ref_total = g.grade(robust)
                ref = _state(g)
                total.fill_(-1)
                err.fill_(1)
                torch.cuda.synchronize()
                g.grade_noalloc(robust, s.cuda_stream, total.data_ptr(), err.data_ptr())
                s.synchronize()
                assert int(err.item()) == 0, name
                assert int(total.item()) == ref_total, (name, robust, int(total.item()), ref_total)
                _same(_state(g), ref, (name, robust))


def test_noalloc_grade_with_walk_and_compaction_knobs():
    """The grader's level-walk and compacted-reverse knobs set: the scratch
    holds neither the walk's tile-major plane nor the item lists, so the
    no-alloc grade runs per-level launches with dense reverse launches; it
    must equal the grader's own grades, and the grader's own grades after it
    (lists refilled by their forward pass) must too."""
    cfg = pg.GenConfig(seed=25, num_pi=64, num_po=32, num_gates=5_000, levels=30, fanin_p=(0.1, 0.6, 0.2, 0.1),
                       prev_level_frac=0.7, fanout_alpha=2.2, fanout_cap=64)
    nl = pg.Netlist.generate(cfg)
    P = 65_536
    bits = biased_vectors(25, nl.info()["num_pi"], P, p_hold=0.6)
    total = torch.zeros(1, dtype=torch.int64, device="cuda")
    err = torch.ones(1, dtype=torch.int32, device="cuda")
    s = torch.cuda.Stream()
    with pg.Grader(nl) as g:
        g.set_vectors(bits, P)
        g.set_chunk(P)
        for robust in (True, False):
            g.set_tuning(pg.TUNE_WALK, 0)
            g.set_tuning(pg.TUNE_REV_COMPACT, 0)
            ref_total = g.grade(robust)
            ref = _state(g)
            assert ref_total > 0
            for walk, compact in ((0, 1), (4, 0), (4, 1)):
                say(f"noalloc knobs robust={robust} walk={walk} compact={compact}")
                g.set_tuning(pg.TUNE_WALK, walk)
                g.set_tuning(pg.TUNE_REV_COMPACT, compact)
                assert g.grade(robust) == ref_total
                _same(_state(g), ref, (robust, walk, compact, "before"))
                total.fill_(-1)
                err.fill_(1)
                torch.cuda.synchronize()
                g.grade_noalloc(robust, s.cuda_stream, total.data_ptr(), err.data_ptr())
                s.synchronize()
                assert int(err.item()) == 0
                assert int(total.item()) == ref_total, (robust, walk, compact, int(total.item()), ref_total)
                _same(_state(g), ref, (robust, walk, compact, "noalloc"))
                assert g.grade(robust) == ref_total
                _same(_state(g), ref, (robust, walk, compact, "after"))


def test_noalloc_grade_explicit_pairs():
    cfg = small_config(23, num_pi=10, num_po=12, num_gates=400, levels=10, wide=True)
    nl = pg.Netlist.generate(cfg)
    I = nl.info()["num_pi"]
    P = 3_000  # not a multiple of 64
    a = biased_vectors(5, I, P)
    b = biased_vectors(6, I, P)
    total = torch.zeros(1, dtype=torch.int64, device="cuda")
    s = torch.cuda.Stream()
    with pg.Grader(nl) as g:
        g.set_pairs(a, b, P)
        for robust in (True, False):
            say(f"noalloc explicit robust={robust}")
            ref_total = g.grade(robust)
            ref = _state(g)
            g.grade_noalloc(robust, s.cuda_stream, total.data_ptr())
            s.synchronize()
            assert int(total.item()) == ref_total
            _same(_state(g), ref, robust)


def test_noalloc_grade_refuses_oversized_netlists():
    say("noalloc oversized")
    cfg = pg.GenConfig(seed=24, num_pi=4096, num_po=2048, num_gates=700_000, levels=60,
                       fanin_p=(0.1, 0.6, 0.2, 0.1), prev_level_frac=0.7, fanout_alpha=2.2, fanout_cap=64,
                       path_limit_log2=0)
    nl = pg.Netlist.generate(cfg)
    P = 128
    bits = pg.gen_vectors(7, nl.info()["num_pi"], P)
    total = torch.zeros(1, dtype=torch.int64, device="cuda")
    s = torch.cuda.Stream()
    with pg.Grader(nl) as g:
        g.set_vectors(bits, P)
        with pytest.raises(pg.PdfError) as ei:
            g.grade_noalloc(True, s.cuda_stream, total.data_ptr())
        assert ei.value.code == pg.ERR_UNSUPPORTED
        assert "static scratch" in str(ei.value)
