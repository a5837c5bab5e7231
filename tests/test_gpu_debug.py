"""The stage API's debug copies (pdf_grader_set_debug: forward marks before
removal, c of every PI per pair) and fcount's --verbose logs in the
reference's formats (gpusim-p2.log, gpumark-full.log, gpumark.log,
gpumerge.log, gpucover.log; simkernel.cuh:250-295, markkernel.cuh:323-395,
mergekernel.cuh:25-47, coverkernel.cuh:44-66), against the oracle chunk by
chunk."""
import os
import subprocess
import sys

import numpy as np
import pytest

import pdfgrade as pg
from oracle import oracle as orc

from helpers import biased_vectors, small_config

pytestmark = pytest.mark.gpu

FCOUNT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpu-path-delay-coverage_amd",
                      "fcount")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_chunks(nl, bits, P, W, robust):
    """The chunk loop in the oracle, keeping every chunk's intermediates."""
    arrs = nl.arrays()
    net = orc.OracleNet(arrs)
    f0 = np.full(net.n, 0xFFFFFFFF, np.uint32)
    f1 = f0.copy()
    gc = np.zeros(net.n, np.uint64)
    out = []
    for start in range(0, P, W):
        w = min(W, P - start)
        codes = orc.sim_chunk(net, bits, P, start, w)
        marks, fwd = orc.mark_chunk(net, codes, robust)
        orc.merge_chunk(net, codes, marks, start, f0, f1)
        newm = orc.new_from_firsts(net.n, w, start, f0, f1)
        _, _, pic = orc.cover_new_chunk(net, marks, newm, gc, arrs["num_pi"])
        out.append(dict(start=start, width=w, codes=codes, fwd=fwd, marks=marks, pic=pic))
    return out, f0, f1, gc


def _circuit():
    return pg.Netlist.generate(small_config(1, num_gates=60, levels=8))


def test_debug_copies_match_oracle():
    nl = _circuit()
    I = nl.info()["num_pi"]
    P, W = 320, 128  # chunks of 128, 128 and 64 pairs
    bits = biased_vectors(1, I, P, p_hold=0.5)
    for robust in (True, False):
        say("debug copies", robust)
        ref, f0, f1, gc = oracle_chunks(nl, bits, P, W, robust)
        assert any(int(r["pic"].sum()) > 0 for r in ref)
        with pg.Grader(nl) as g:
            g.set_vectors(bits, P)
            g.set_chunk(W)
            g.set_debug(pg.DEBUG_FORWARD_MARKS | pg.DEBUG_PI_COVER)
            for c, r in enumerate(ref):
                g.sim(c, r["start"])
                g.mark(c, r["start"], robust)
                assert np.array_equal(pg.plane_to_bytes(g.forward_mark_plane(r["width"]), r["width"]), r["fwd"])
                assert np.array_equal(pg.plane_to_bytes(g.mark_plane(r["width"]), r["width"]), r["marks"])
                g.merge(c, r["start"])
                g.cover(c, r["start"])
                assert np.array_equal(g.pi_cover(I, r["width"]), r["pic"]), c
            assert np.array_equal(g.gate_counts(), gc)
            assert np.array_equal(g.first()[0], f0) and np.array_equal(g.first()[1], f1)
    with pg.Grader(nl) as g:  # without the flags the getters refuse
        g.set_vectors(bits, P)
        g.set_chunk(W)
        g.sim(0, 0)
        g.mark(0, 0, True)
        with pytest.raises(pg.PdfError):
            g.forward_mark_plane(W)
        with pytest.raises(pg.PdfError):
            g.set_debug(8)


def _header(label, count):
    return label + "\t" + "".join(f"{i:>4} " for i in range(count)) + "\n"


def expected_logs(nl, bits, P, W, robust):
    ref, f0, f1, _ = oracle_chunks(nl, bits, P, W, robust)
    n = nl.info()["n"]
    I = nl.info()["num_pi"]
    names = ["S0", "S1", "T0", "T1"]
    sim = _header("Gate:     ", n)
    full = _header("Gate:     ", n)
    mark = _header("Gate:     ", n)
    cover = _header("Gate:   ", I)
    for r in ref:
        for p in range(r["width"]):
            row = f"Vector {r['start'] + p}:\t"
            sim += row + "".join(f"{names[int(r['codes'][i, p])] + ' ':>5}" for i in range(n)) + "\n"
            full += row + "".join(f"{'Y' if r['fwd'][i, p] else 'N':>4} " for i in range(n)) + "\n"
            mark += row + "".join(f"{'Y' if r['marks'][i, p] else 'N':>4} " for i in range(n)) + "\n"
            cover += row + "".join(f"{int(r['pic'][i, p]):>4} " for i in range(I)) + "\n"
    merge = "".join(f"Segment {i}({i}): {np.int32(np.uint32(f0[i]).view(np.int32))},"
                    f"{np.int32(np.uint32(f1[i]).view(np.int32))} \n" for i in range(n))
    return {"gpusim-p2.log": sim, "gpumark-full.log": full, "gpumark.log": mark, "gpucover.log": cover,
            "gpumerge.log": merge}


def test_fcount_verbose_logs_match_oracle(tmp_path):
    """fcount --verbose --sim --mark --merge on a 3-chunk run (--num_patterns
    129 gives 128-pair chunks): every log equals the oracle's text."""
    nl = _circuit()
    I = nl.info()["num_pi"]
    V = 320
    bits = biased_vectors(1, I, V, p_hold=0.5)
    ckt = tmp_path / "ckt.level"
    nl.write_level(str(ckt))
    vec = tmp_path / "v.vec"
    pg.write_vectors(str(vec), bits, V)
    for robust in (True, False):
        say("fcount verbose", robust)
        args = [FCOUNT, "--bench", str(ckt), "--num_patterns", "129", "--verbose", "--sim", "--mark", "--merge",
                str(vec)] + (["--robust"] if robust else [])
        r = subprocess.run(args, capture_output=True, text=True, timeout=120, cwd=tmp_path)
        assert r.returncode == 0, r.stderr
        want = expected_logs(pg.Netlist.read_level(str(ckt)), bits, V, 128, robust)
        for name, text in want.items():
            got = (tmp_path / name).read_text()
            assert got == text, (robust, name)
