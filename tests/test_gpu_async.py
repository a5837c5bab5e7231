"""pdf_grade_async: the fused grade queued on a caller's HIP stream (here a
torch stream) and captured into a HIP graph, checked against the grader's own
synchronous grade; graph replays are checked one by one, with eager work and
synchronizations between them."""
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
