"""pdf_grade_async: the fused grade queued on a caller's HIP stream (here a
torch stream) and captured into a HIP graph, and the four stages queued the
same way (pdf_sim_async ... pdf_cover_async), checked against the grader's
own synchronous grade; graph replays are checked one by one, with eager work
and synchronizations between them."""
import sys

import numpy as np
import pytest
import torch

import pdfgrade as pg
from pdfgrade.configs import GATES_50K

from helpers import biased_vectors, small_config

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _state(g):
    f0, f1 = g.first()
    return g.total(), f0, f1, g.gate_counts()


def _same(a, b, ctx):
    assert a[0] == b[0], ctx
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y), ctx


def test_async_grade_on_a_torch_stream():
    say("async on stream")
    nl = pg.Netlist.generate(GATES_50K.netlist)
    P = 16384
    bits = pg.gen_vectors(GATES_50K.vector_seed, nl.info()["num_pi"], P)
    total = torch.zeros(1, dtype=torch.int64, device="cuda")
    err = torch.ones(1, dtype=torch.int32, device="cuda")
    s = torch.cuda.Stream()
    with pg.Grader(nl) as g:
        g.set_vectors(bits, P)
        g.set_tuning(pg.TUNE_LANES, 2)
        g.set_chunk(4096)
        for robust in (True, False):
            ref_total = g.grade(robust)
            ref = _state(g)
            g.grade_async(robust, 1, s.cuda_stream, total.data_ptr(), err.data_ptr())
            s.synchronize()
            assert int(total.item()) == ref_total and int(err.item()) == 0
            _same(_state(g), ref, f"robust {robust}")


def test_async_grade_captured_in_a_graph():
    say("async captured")
    nl = pg.Netlist.generate(small_config(3, num_gates=400, levels=10, wide=True))
    I = nl.info()["num_pi"]
    P = 2048
    bits = biased_vectors(3, I, P, p_hold=0.6)
    total = torch.zeros(1, dtype=torch.int64, device="cuda")
    s = torch.cuda.Stream()
    with pg.Grader(nl) as g:
        g.set_vectors(bits, P)
        g.set_chunk(512)
        for K in (1, 2):
            ref_total = g.grade_segs(False, K)  # also sets up the segment list for K = 2
            ref = _state(g)
            seg_ref = g.segment_first() if K > 1 else None
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph, stream=s, capture_error_mode="thread_local"):
                g.grade_async(False, K, s.cuda_stream, total.data_ptr())
            for rep in range(3):
                total.zero_()
                g.grade(True)                      # eager work between replays
                torch.cuda.synchronize()
                graph.replay()
                torch.cuda.synchronize()
                assert int(total.item()) == ref_total, (K, rep)
                _same(_state(g), ref, (K, rep))
                if K > 1:
                    s0, s1 = g.segment_first()
                    assert np.array_equal(s0, seg_ref[0]) and np.array_equal(s1, seg_ref[1]), (K, rep)
            del graph


def test_async_grade_errors():
    nl = pg.Netlist.generate(small_config(4))
    P = 256
    bits = biased_vectors(4, nl.info()["num_pi"], P)
    total = torch.zeros(1, dtype=torch.int64, device="cuda")
    s = torch.cuda.Stream()
    with pg.Grader(nl) as g:
        g.set_vectors(bits, P)
        g.set_chunk(0)
        with pytest.raises(pg.PdfError):
            g.grade_async(True, 3, s.cuda_stream, total.data_ptr())  # no segment list of length 3 yet
        with pytest.raises(pg.PdfError):
            g.grade_async(True, 9, s.cuda_stream, total.data_ptr())


def dense_circuit(seed):
    """Mostly NOT/BUFF gates in 8 levels: 50-100 paths over a few thousand
    held pairs."""
    return pg.GenConfig(seed=seed, num_pi=24, num_po=60, num_gates=600, levels=8, fanin_p=(0.7, 0.25, 0.05, 0.0),
                        prev_level_frac=0.7, fanout_alpha=2.2, fanout_cap=16)


def _stage_chunks(g, robust, K, s, totals, gates=None, err=None):
    """Queue reset + the four stages of every chunk on stream s; chunk c's
    total lands in totals[c]."""
    W = g.chunk_width(0)  # every chunk but the last is this wide
    g.reset_async(s.cuda_stream)
    for c in range(g.num_chunks()):
        g.sim_async(c, c * W, s.cuda_stream)
        g.mark_async(c, c * W, robust, s.cuda_stream)
        g.merge_async(c, c * W, K, s.cuda_stream)
        g.cover_async(c, c * W, K, s.cuda_stream, totals[c:].data_ptr(),
                      gates.data_ptr() if gates is not None else 0, err.data_ptr() if err is not None else 0)


def test_async_stages_on_a_torch_stream():
    """The stages queued per chunk on a torch stream equal the synchronous
    grade: per-chunk totals (summed), first detections and per-gate counts
    (through per_gate_accum), in both modes and with segments of length 2."""
    say("async stages")
    nl = pg.Netlist.generate(dense_circuit(302))
    I, n = nl.info()["num_pi"], nl.info()["n"]
    P = 5000  # ragged: not a multiple of 64
    bits = biased_vectors(302, I, P, p_hold=0.8)
    s = torch.cuda.Stream()
    with pg.Grader(nl) as g:
        g.set_vectors(bits, P)
        g.set_chunk(1024)
        nch = g.num_chunks()
        for K in (1, 2):
            for robust in (True, False):
                ref_total = g.grade_segs(robust, K)
                assert ref_total > 0
                ref = _state(g)
                totals = torch.full((nch,), -1, dtype=torch.int64, device="cuda")
                gates = torch.zeros(n, dtype=torch.int64, device="cuda")
                err = torch.ones(1, dtype=torch.int32, device="cuda")
                _stage_chunks(g, robust, K, s, totals, gates, err)
                s.synchronize()
                assert int(totals.sum().item()) == ref_total, (K, robust)
                assert int(err.item()) == 0
                f0, f1 = g.first()
                assert np.array_equal(f0, ref[1]) and np.array_equal(f1, ref[2]), (K, robust)
                assert np.array_equal(gates.cpu().numpy().view(np.uint64), ref[3]), (K, robust)


def test_async_stages_captured_in_a_graph():
    """One vector set's reset and stages captured into a HIP graph; replays
    (with eager grading in between) give the synchronous grade's results."""
    say("async stages captured")
    nl = pg.Netlist.generate(dense_circuit(301))
    I, n = nl.info()["num_pi"], nl.info()["n"]
    P = 3072
    bits = biased_vectors(301, I, P, p_hold=0.8)
    s = torch.cuda.Stream()
    with pg.Grader(nl) as g:
        g.set_vectors(bits, P)
        g.set_chunk(1024)
        ref_total = g.grade(False)
        assert ref_total > 0
        ref = _state(g)
        totals = torch.zeros(g.num_chunks(), dtype=torch.int64, device="cuda")
        gates = torch.zeros(n, dtype=torch.int64, device="cuda")
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=s, capture_error_mode="thread_local"):
            _stage_chunks(g, False, 1, s, totals, gates)
        for rep in range(3):
            totals.zero_()
            gates.zero_()
            g.grade(True)  # eager work between replays
            torch.cuda.synchronize()
            graph.replay()
            torch.cuda.synchronize()
            assert int(totals.sum().item()) == ref_total, rep
            f0, f1 = g.first()
            assert np.array_equal(f0, ref[1]) and np.array_equal(f1, ref[2]), rep
            assert np.array_equal(gates.cpu().numpy().view(np.uint64), ref[3]), rep
        del graph


def test_async_stage_order_and_errors():
    nl = pg.Netlist.generate(small_config(10))
    P = 512
    bits = biased_vectors(10, nl.info()["num_pi"], P)
    t = torch.zeros(1, dtype=torch.int64, device="cuda")
    s = torch.cuda.Stream()
    with pg.Grader(nl) as g:
        g.set_vectors(bits, P)
        g.set_chunk(256)
        with pytest.raises(pg.PdfError):
            g.mark_async(0, 0, True, s.cuda_stream)  # not simulated
        g.sim_async(0, 0, s.cuda_stream)
        with pytest.raises(pg.PdfError):
            g.merge_async(0, 0, 1, s.cuda_stream)  # not marked
        with pytest.raises(pg.PdfError):
            g.sim_async(0, 64, s.cuda_stream)  # start != chunk * chunk_pairs
        g.mark_async(0, 0, True, s.cuda_stream)
        with pytest.raises(pg.PdfError):
            g.merge_async(0, 0, 3, s.cuda_stream)  # no segment list of length 3
        g.merge_async(0, 0, 1, s.cuda_stream)
        with pytest.raises(pg.PdfError):
            g.cover_async(0, 0, 2, s.cuda_stream, t.data_ptr())  # segment length differs
        g.cover_async(0, 0, 1, s.cuda_stream, t.data_ptr())
        s.synchronize()
        g.set_debug(pg.DEBUG_FORWARD_MARKS)
        with pytest.raises(pg.PdfError):
            g.sim_async(1, 256, s.cuda_stream)  # debug dumps need the synchronous calls
