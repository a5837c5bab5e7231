"""REF mode on the GPU: the reference's 4,224-pattern merge window
(mergekernel.cuh:125-130) against the oracle with the same window, the
flags for values past the reference's 16-bit and uint32 ranges against the
oracle's, the segment-stride count, and fcount reproducing the reference's
default run (one chunk, windowed merge)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import pdfgrade as pg
from oracle import oracle as orc

from helpers import biased_vectors

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WINDOW = pg.REF_MERGE_PATTERNS


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def circuit():
    return pg.Netlist.generate(pg.GenConfig(seed=61, num_pi=32, num_po=48, num_gates=2000, levels=16,
                                            fanin_p=(0.3, 0.5, 0.15, 0.05), prev_level_frac=0.7,
                                            fanout_alpha=2.2, fanout_cap=32))


def late_vectors(seed, num_pi, P, quiet=5000):
    """Hold-biased vectors whose first `quiet` vectors are all equal: no pair
    before quiet - 1 has a transition, so the first chunk's detections lie
    past the 4,224-pattern window."""
    from helpers import pack, unpack
    vals = unpack(biased_vectors(seed, num_pi, P, p_hold=0.93), P)
    vals[:quiet] = vals[0]
    return pack(vals)


def state(g, total):
    f0, f1 = g.first()
    return dict(total=total, first_t0=f0, first_t1=f1, gate_counts=g.gate_counts())


def assert_same(got, ref, ctx):
    assert got["total"] == ref["total"], (ctx, got["total"], ref["total"])
    assert np.array_equal(got["first_t0"], ref["first_t0"]), ctx
    assert np.array_equal(got["first_t1"], ref["first_t1"]), ctx
    assert np.array_equal(got["gate_counts"], ref["gate_counts"]), ctx


@pytest.mark.parametrize("robust", [True, False])
def test_merge_window_16k_vectors_chunk_8192(robust):
    nl = circuit()
    P, chunk = 16384, 8192
    bits = late_vectors(62, nl.info()["num_pi"], P)
    ref = orc.grade(orc.OracleNet(nl.arrays()), bits, P, chunk, robust, merge_window=WINDOW)
    full = orc.grade(orc.OracleNet(nl.arrays()), bits, P, chunk, robust)
    # the window matters here: some first detections lie past it
    assert not np.array_equal(ref["first_t0"], full["first_t0"]) or ref["total"] != full["total"]
    for lanes, walk in ((1, 0), (2, 0), (1, 4)):
        say("merge window", robust, lanes, walk)
        with pg.Grader(nl) as g:
            g.set_vectors(bits, P)
            g.set_chunk(chunk)
            g.set_tuning(pg.TUNE_LANES, lanes)
            g.set_tuning(pg.TUNE_WALK, walk)
            g.set_merge_window(WINDOW)
            assert_same(state(g, g.grade(robust)), ref, ("fused", lanes, walk))
            flags, _ = g.ref_flags()
            assert not flags & pg.REF_MERGE_WINDOW
            total, per_chunk = g.grade_stages(robust)
            assert_same(state(g, total), ref, ("stages", lanes, walk))
            assert list(per_chunk) == [int(x) for x in ref["chunk_totals"]]
            g.set_merge_window(0)
            assert_same(state(g, g.grade(robust)), full, ("no window", lanes, walk))
            flags, _ = g.ref_flags()
            assert flags & pg.REF_MERGE_WINDOW


def test_merge_window_segments():
    nl = circuit()
    P, chunk = 12288, 6144
    bits = late_vectors(63, nl.info()["num_pi"], P)
    for K in (2, 3):
        say("merge window segments", K)
        ref = orc.grade_segs(orc.OracleNet(nl.arrays()), bits, P, chunk, False, K, merge_window=WINDOW)
        with pg.Grader(nl) as g:
            g.set_vectors(bits, P)
            g.set_chunk(chunk)
            g.set_merge_window(WINDOW)
            total = g.grade_segs(False, K)
            assert total == ref["total"]
            assert np.array_equal(g.gate_counts(), ref["gate_counts"])
            s0, s1 = g.segment_first()
            assert np.array_equal(s0, ref["seg_t0"]) and np.array_equal(s1, ref["seg_t1"])
            flags, unmerged = g.ref_flags()
            assert unmerged == pg.lib().pdf_ref_unmerged_segments(len(s0))


def fan_tree(width: int):
    types, fanins, po = [pg.INPT], [[]], [0]
    mids = []
    for _ in range(width):
        types.append(pg.BUFF); fanins.append([0]); po.append(0)
        mids.append(len(types) - 1)
    for m in mids:
        for _ in range(width):
            types.append(pg.BUFF); fanins.append([m]); po.append(1)
    return pg.Netlist.from_arrays(1, types, po, fanins)


def test_u16_flag_on_fan_trees():
    """width^2 marked paths from the PI (every node new at its first
    toggles): past 65,535 the reference's 16-bit fields overflow."""
    from helpers import pack
    P = 64
    bits = pack(np.array([[[0, 1, 0, 0][i % 4]] for i in range(P)], np.uint8))
    for width, over in ((200, False), (260, True)):
        say("u16 flag", width)
        nl = fan_tree(width)
        ref = orc.grade(orc.OracleNet(nl.arrays()), bits, P, P, True)
        with pg.Grader(nl) as g:
            g.set_vectors(bits, P)
            total = g.grade(True)
            flags, unmerged = g.ref_flags()
        assert total == ref["total"] == 2 * width * width
        assert bool(flags & pg.REF_EXCEEDS_U16) == bool(ref["flags"] & orc.FLAG_EXCEEDS_U16_DP) == over
        assert not flags & pg.REF_EXCEEDS_U32 and not ref["flags"] & orc.FLAG_EXCEEDS_U32
        assert not flags & pg.REF_SEGMENT_STRIDE and unmerged == 0


def test_fcount_reproduces_reference_default_run(tmp_path):
    """fcount with --ref_window, the chunk left to free memory (the whole
    vector set in one chunk, as the reference sizes it), prints the oracle's
    total with the reference's window; without it the full-chunk total."""
    nl = circuit()
    lvl = tmp_path / "ckt.level"
    nl.write_level(str(lvl))
    V = 16384
    bits = late_vectors(64, nl.info()["num_pi"], V)
    vec = tmp_path / "ckt.vec"
    pg.write_vectors(str(vec), bits, V)
    exe = os.path.join(ROOT, "gpu-path-delay-coverage_amd", "fcount")
    net = orc.OracleNet(pg.Netlist.read_level(str(lvl)).arrays())
    for args, window in ((["--ref_window"], WINDOW), ([], 0)):
        for robust in (True, False):
            say("fcount", args, robust)
            cmd = [exe, "--bench", str(lvl), str(vec)] + args + (["--robust"] if robust else [])
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
            assert r.returncode == 0, r.stderr
            line = [l for l in r.stdout.splitlines() if l.startswith(str(vec))][-1]
            want = orc.grade(net, bits, V, V, robust, merge_window=window)["total"]
            assert int(line.rsplit(";", 1)[1]) == want, (line, want)
            assert ("4224-pattern merge window" in r.stderr) == (window == 0)
