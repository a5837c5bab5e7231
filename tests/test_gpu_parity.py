"""Bit-exact parity of the HIP grading path (through the C ABI) against the CPU
oracle, on the same seeded inputs.  Integer work: every output must be equal."""
import os
import sys

import numpy as np
import pytest

import pdfgrade as pg
from pdfgrade.configs import GATES_50K, TINY
from oracle import oracle as orc

from helpers import (biased_vectors, bitsliced_grade, edge_netlist, oracle_grade, pack, small_config,
                     unpack)

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_grade(nl, bits, P, chunk, robust, b_bits=None, stages=False):
    with pg.Grader(nl) as g:
        if b_bits is None:
            g.set_vectors(bits, P)
        else:
            g.set_pairs(bits, b_bits, P)
        g.set_chunk(chunk)
        if stages:
            total, per_chunk = g.grade_stages(robust)
        else:
            total, per_chunk = g.grade(robust), None
        f0, f1 = g.first()
        return dict(total=total, per_chunk=per_chunk, first_t0=f0, first_t1=f1, gate_counts=g.gate_counts())


def assert_same(got, ref, ctx=""):
    assert got["total"] == ref["total"], ctx
    assert np.array_equal(got["first_t0"], ref["first_t0"]), ctx
    assert np.array_equal(got["first_t1"], ref["first_t1"]), ctx
    assert np.array_equal(got["gate_counts"], ref["gate_counts"]), ctx
    if got.get("per_chunk") is not None:
        assert np.array_equal(np.array(got["per_chunk"], np.uint64), ref["chunk_totals"]), ctx


@pytest.mark.parametrize("robust", [True, False])
@pytest.mark.parametrize("chunk", [64, 256, 1024])
def test_tiny_workload(robust, chunk):
    say("tiny", robust, chunk)
    nl = pg.Netlist.generate(TINY.netlist)
    bits = pg.gen_vectors(TINY.vector_seed, nl.info()["num_pi"], TINY.pairs)
    ref = oracle_grade(nl, bits, TINY.pairs, chunk, robust)
    assert_same(gpu_grade(nl, bits, TINY.pairs, chunk, robust), ref, "fused")
    assert_same(gpu_grade(nl, bits, TINY.pairs, chunk, robust, stages=True), ref, "stages")


def check_intermediates(nl, bits, P, chunk, robust, b_bits=None):
    """Sim codes and final marks of every chunk against the oracle."""
    net = orc.OracleNet(nl.arrays())
    with pg.Grader(nl) as g:
        if b_bits is None:
            g.set_vectors(bits, P)
        else:
            g.set_pairs(bits, b_bits, P)
        g.set_chunk(chunk)
        f0 = np.full(net.n, 0xFFFFFFFF, np.uint32)
        f1 = f0.copy()
        gc = np.zeros(net.n, np.uint64)
        start = 0
        for c in range(g.num_chunks()):
            w = g.chunk_width(c)
            g.sim(c, start)
            a, b = g.sim_planes(w)
            codes = orc.sim_chunk(net, bits, P, start, w, b_bits=b_bits)
            assert np.array_equal(pg.planes_to_codes(a, b, w), codes), f"sim chunk {c}"
            # lanes past the last pair never transition
            assert (pg.planes_to_codes(a, b)[:, w:] < 2).all()
            g.mark(c, start, robust)
            marks, _ = orc.mark_chunk(net, codes, robust)
            assert np.array_equal(pg.plane_to_bytes(g.mark_plane(w), w), marks), f"mark chunk {c}"
            g.merge(c, start)
            orc.merge_chunk(net, codes, marks, start, f0, f1)
            gf0, gf1 = g.first()
            assert np.array_equal(gf0, f0) and np.array_equal(gf1, f1), f"merge chunk {c}"
            cov, _ = g.cover(c, start)
            ref_cov, _ = orc.cover_chunk(net, marks, start, f0, f1, gc)
            assert cov == ref_cov, f"cover chunk {c}"
            assert np.array_equal(g.gate_counts(), gc), f"gate counts chunk {c}"
            start += w


@pytest.mark.parametrize("robust", [True, False])
def test_tiny_intermediates(robust):
    say("tiny intermediates", robust)
    nl = pg.Netlist.generate(TINY.netlist)
    bits = pg.gen_vectors(TINY.vector_seed, nl.info()["num_pi"], TINY.pairs)
    check_intermediates(nl, bits, TINY.pairs, 256, robust)


@pytest.mark.parametrize("seed", range(16))
def test_random_small_circuits(seed):
    say("random", seed)
    cfg = small_config(seed, num_pi=6 + seed % 5, num_gates=30 + 7 * seed, levels=4 + seed % 6,
                       wide=seed % 2 == 0, alpha=2.2 if seed % 3 == 0 else 0.0)
    nl = pg.Netlist.generate(cfg)
    P = 64 * (2 + seed % 4)
    bits = biased_vectors(seed, nl.info()["num_pi"], P, p_hold=0.3 + 0.1 * (seed % 5))
    for robust in (True, False):
        chunk = 64 * (1 + seed % 3)
        ref = oracle_grade(nl, bits, P, chunk, robust)
        assert_same(gpu_grade(nl, bits, P, chunk, robust), ref, f"seed {seed} robust {robust}")
    check_intermediates(nl, bits, P, 64, True)


def test_explicit_pairs():
    say("explicit pairs")
    nl = pg.Netlist.generate(small_config(21, num_gates=60, levels=7))
    P = 256
    I = nl.info()["num_pi"]
    a = biased_vectors(1, I, P, p_hold=0.0)
    b = pack(np.where(np.random.default_rng(2).random((P, I)) < 0.3, 1 - unpack(a, P), unpack(a, P)))
    for robust in (True, False):
        ref = oracle_grade(nl, a, P, 128, robust, b_bits=b)
        assert_same(gpu_grade(nl, a, P, 128, robust, b_bits=b), ref, "explicit")
    check_intermediates(nl, a, P, 128, False, b_bits=b)


@pytest.mark.parametrize("robust", [True, False])
def test_edge_netlist(robust):
    say("edge netlist", robust)
    nl = edge_netlist()
    P = 512
    bits = biased_vectors(9, 8, P, p_hold=0.5)
    ref = oracle_grade(nl, bits, P, 128, robust)
    assert_same(gpu_grade(nl, bits, P, 128, robust), ref, "edge")
    check_intermediates(nl, bits, P, 192, robust)


@pytest.mark.parametrize("P,chunk", [(1000, 256), (130, 64), (77, 0)])
def test_pair_count_not_multiple_of_64(P, chunk):
    """The last pair wraps to vector 0 (simkernel.cuh:116-120); lanes past the
    last pair never transition."""
    say("ragged", P, chunk)
    nl = pg.Netlist.generate(small_config(33, num_gates=70, levels=6))
    bits = biased_vectors(33, nl.info()["num_pi"], P, p_hold=0.5)
    for robust in (True, False):
        ref = oracle_grade(nl, bits, P, chunk, robust)
        assert_same(gpu_grade(nl, bits, P, chunk, robust), ref, f"P={P}")
        assert_same(gpu_grade(nl, bits, P, chunk, robust, stages=True), ref, f"P={P} stages")
    check_intermediates(nl, bits, P, chunk or P, True)


def test_constant_vectors_give_zero():
    say("constant")
    nl = pg.Netlist.generate(TINY.netlist)
    bits = np.zeros((nl.info()["num_pi"], 16), np.uint64)
    r = gpu_grade(nl, bits, 1024, 512, True)
    assert r["total"] == 0 and not r["gate_counts"].any()
    assert np.all(r["first_t0"] == pg.NONE) and np.all(r["first_t1"] == pg.NONE)


def test_50k_gates_sample_against_oracle():
    """The 50k-gate netlist on its first 4,096 pairs (as a 4,096-vector set)."""
    say("50k sample")
    nl = pg.Netlist.generate(GATES_50K.netlist)
    P = 4096
    bits = pg.gen_vectors(GATES_50K.vector_seed, nl.info()["num_pi"], P)
    ref = oracle_grade(nl, bits, P, 1024, True)
    assert_same(gpu_grade(nl, bits, P, 1024, True), ref, "50k fused")
    st = gpu_grade(nl, bits, P, 4096, True, stages=True)
    assert st["per_chunk"] == [ref["total"]]  # one 4,096-pair chunk
    assert_same({**st, "per_chunk": None}, ref, "50k stages")


def test_50k_gates_full_size_properties():
    """Full 65,536 pairs: chunk invariance, stage path == fused path, and the
    planes of an interior chunk equal the oracle's for that chunk."""
    say("50k full")
    nl = pg.Netlist.generate(GATES_50K.netlist)
    P = GATES_50K.pairs
    bits = pg.gen_vectors(GATES_50K.vector_seed, nl.info()["num_pi"], P)
    full = gpu_grade(nl, bits, P, P, True)
    for chunk in (8192, 4096 * 3):
        r = gpu_grade(nl, bits, P, chunk, True)
        r["per_chunk"] = None
        assert_same(r, full, f"chunk {chunk}")
    st = gpu_grade(nl, bits, P, 16384, True, stages=True)
    assert sum(st["per_chunk"]) == full["total"]
    assert_same({**st, "per_chunk": None}, full, "stages")
    # every first-detection pattern is a valid pair index, and counts exist only
    # on nodes that were detected or sit above a detected node
    det = (full["first_t0"] != pg.NONE) | (full["first_t1"] != pg.NONE)
    assert np.all(full["first_t0"][det & (full["first_t0"] != pg.NONE)] < P)
    assert full["total"] <= int(full["gate_counts"].sum())
    # interior chunk planes vs the oracle (planes carry no cross-chunk state)
    net = orc.OracleNet(nl.arrays())
    with pg.Grader(nl) as g:
        g.set_vectors(bits, P)
        g.set_chunk(2048)
        c = 17
        start = c * 2048
        g.sim(c, start)
        a, b = g.sim_planes(2048)
        codes = orc.sim_chunk(net, bits, P, start, 2048)
        assert np.array_equal(pg.planes_to_codes(a, b), codes)
        g.mark(c, start, True)
        marks, _ = orc.mark_chunk(net, codes, True)
        assert np.array_equal(pg.plane_to_bytes(g.mark_plane(2048)), marks)
    # the last chunk wraps to vector 0 (simkernel.cuh:116-120)
    with pg.Grader(nl) as g:
        g.set_vectors(bits, P)
        g.set_chunk(1024)
        c = 63
        g.sim(c, c * 1024)
        a, b = g.sim_planes(1024)
        codes = orc.sim_chunk(net, bits, P, c * 1024, 1024)
        assert np.array_equal(pg.planes_to_codes(a, b), codes)


def test_api_errors():
    say("api errors")
    nl = pg.Netlist.generate(TINY.netlist)
    bits = pg.gen_vectors(1, nl.info()["num_pi"], 1024)
    with pg.Grader(nl) as g:
        with pytest.raises(pg.PdfError) as e:
            g.set_vectors(bits, 0)
        assert e.value.code == pg.ERR_ARG
        g.set_vectors(bits, 1024)
        with pytest.raises(pg.PdfError) as e:
            g.set_chunk(100)
        assert e.value.code == pg.ERR_ARG
        g.set_chunk(256)
        with pytest.raises(pg.PdfError) as e:
            g.mark(0, 0, True)
        assert e.value.code == pg.ERR_STATE
        g.sim(0, 0)
        with pytest.raises(pg.PdfError) as e:
            g.sim(1, 0)  # start must be chunk * chunk_pairs
        assert e.value.code == pg.ERR_ARG
        g.mark(0, 0, True)
        with pytest.raises(pg.PdfError) as e:
            g.merge(0, 0, segs=9)  # segment lengths are 1..8 (main.cu:97-103)
        assert e.value.code == pg.ERR_ARG
        with pytest.raises(pg.PdfError) as e:
            g.sim(4, 1024)
        assert e.value.code == pg.ERR_ARG


@pytest.mark.parametrize("robust", [True, False])
def test_two_words_per_lane_path(robust):
    """A 65,536-pair chunk (1,024 words) takes the two-words-per-lane forward
    kernel; a 1,500-gate netlist keeps the oracle fast."""
    say("two words per lane", robust)
    cfg = pg.GenConfig(seed=77, num_pi=64, num_po=64, num_gates=1500, levels=20, fanin_p=(0.1, 0.6, 0.2, 0.1),
                       prev_level_frac=0.7, fanout_alpha=2.2, fanout_cap=64)
    nl = pg.Netlist.generate(cfg)
    P = 65536
    bits = biased_vectors(77, nl.info()["num_pi"], P, p_hold=0.6)
    ref = oracle_grade(nl, bits, P, P, robust)
    assert ref["total"] > 0
    assert_same(gpu_grade(nl, bits, P, P, robust), ref, "nw2")
    assert_same(gpu_grade(nl, bits, P, 8192, robust), ref, "nw1 chunks")


@pytest.mark.parametrize("robust", [True, False])
def test_reverse_tiles_per_wave(robust):
    """The reverse pass with 1, 2, 4 and 8 64-word tiles per wave, and the
    automatic choice, on a 1,500-gate netlist: a 65,536-pair chunk (16
    tiles) and a ragged 40,859-pair last chunk (639 words: 9 full tiles and
    a partial one, so the last tile group is short), fused and stage by
    stage, against the oracle."""
    say("reverse tiles", robust)
    cfg = pg.GenConfig(seed=77, num_pi=64, num_po=64, num_gates=1500, levels=20, fanin_p=(0.1, 0.6, 0.2, 0.1),
                       prev_level_frac=0.7, fanout_alpha=2.2, fanout_cap=64)
    nl = pg.Netlist.generate(cfg)
    P = 65536 + 40896 - 37
    bits = biased_vectors(77, nl.info()["num_pi"], P, p_hold=0.6)
    ref = oracle_grade(nl, bits, P, 65536, robust)
    assert ref["total"] > 0
    for rt in (0, 1, 2, 4, 8):
        with pg.Grader(nl) as g:
            g.set_vectors(bits, P)
            g.set_chunk(65536)
            g.set_tuning(pg.TUNE_REV_TILES, rt)
            for stages in (False, True):
                if stages:
                    total, per_chunk = g.grade_stages(robust)
                else:
                    total, per_chunk = g.grade(robust), None
                f0, f1 = g.first()
                got = dict(total=total, per_chunk=per_chunk, first_t0=f0, first_t1=f1, gate_counts=g.gate_counts())
                assert_same(got, ref, f"rt {rt} stages {stages}")
        with pg.Grader(nl) as g, pytest.raises(pg.PdfError):
            g.set_tuning(pg.TUNE_REV_TILES, 3)


def test_graph_replays_match_eager():
    """pdf_grade replays a captured HIP graph; every replay, with eager stage
    calls and synchronizations in between, equals the eager launch sequence."""
    say("graph replays")
    nl = pg.Netlist.generate(small_config(41, num_gates=300, levels=12))
    P = 4096
    bits = biased_vectors(41, nl.info()["num_pi"], P, p_hold=0.6)
    with pg.Grader(nl) as g:
        g.set_vectors(bits, P)
        g.set_chunk(1024)
        g.set_tuning(pg.TUNE_USE_GRAPH, 0)
        eager = [g.grade(True), g.gate_counts().copy(), *g.first()]
        g.set_tuning(pg.TUNE_USE_GRAPH, 1)
        for rep in range(4):
            assert g.grade(True) == eager[0], rep
            assert np.array_equal(g.gate_counts(), eager[1])
            f0, f1 = g.first()
            assert np.array_equal(f0, eager[2]) and np.array_equal(f1, eager[3])
            # eager work between replays
            total, _ = g.grade_stages(True)
            assert total == eager[0]
        assert g.grade(False) == oracle_grade(nl, bits, P, 1024, False)["total"]
        for nw in (1, 2):
            g.set_tuning(pg.TUNE_WORDS_PER_LANE, nw)
            assert g.grade(True) == eager[0]


def test_grader_autotune_candidates():
    """pdf_grader_autotune on 262,144 pairs of a 1,500-gate netlist: the
    chosen configuration is one of its candidates (1-4 lanes; chunk = pairs /
    lanes, or a power of two from 32,768 pairs below it), and the grader left
    in it grades bit-exact against the bit-sliced CPU grader, twice."""
    say("grader autotune")
    cfg = pg.GenConfig(seed=77, num_pi=64, num_po=64, num_gates=1500, levels=20, fanin_p=(0.1, 0.6, 0.2, 0.1),
                       prev_level_frac=0.7, fanout_alpha=2.2, fanout_cap=64)
    nl = pg.Netlist.generate(cfg)
    P = 262144
    bits = biased_vectors(77, nl.info()["num_pi"], P, p_hold=0.6)
    ref = bitsliced_grade(nl, bits, P, 65536, True)
    assert ref["total"] > 0
    with pg.Grader(nl) as g:
        g.set_vectors(bits, P)
        g.set_chunk(8192)
        best = g.autotune(True, reps=2)
        say("autotune", best)
        lanes, chunk = best["lanes"], best["chunk_pairs"]
        even = P if lanes == 1 else (-(-P // lanes) + 63) // 64 * 64
        assert 1 <= lanes <= 4
        assert chunk == even or (32768 <= chunk < even and chunk & (chunk - 1) == 0), best
        assert best["use_graph"] in (0, 1) and best["words_per_lane"] in (1, 2) and best["ms"] > 0
        assert best["walk"] in (0, 2, 4) and not (best["walk"] and best["words_per_lane"] == 2), best
        assert g.num_chunks() == -(-P // chunk)
        for rep in range(2):
            got = dict(total=g.grade(True), gate_counts=g.gate_counts())
            got["first_t0"], got["first_t1"] = g.first()
            assert_same(got, ref, f"rep {rep}")


def test_2m_gates_properties():
    """The 2,000,000-gate / 200-level netlist on 65,536 of its pairs: chunk
    invariance across 32,768- and 16,384-pair blocks, and the planes of an
    interior 64-pair chunk and of the wrapping last chunk equal the oracle's."""
    from pdfgrade.configs import GATES_2M
    say("2m gates")
    nl = pg.Netlist.generate(GATES_2M.netlist)
    P = 65536
    bits = pg.gen_vectors(GATES_2M.vector_seed, nl.info()["num_pi"], P)
    a = gpu_grade(nl, bits, P, 32768, True)
    b = gpu_grade(nl, bits, P, 16384, True)
    b["per_chunk"] = None
    assert_same(b, a, "2m chunking")
    net = orc.OracleNet(nl.arrays())
    with pg.Grader(nl) as g:
        g.set_vectors(bits, P)
        g.set_chunk(64)
        for c in (300, P // 64 - 1):
            g.sim(c, c * 64)
            pa, pb = g.sim_planes(64)
            codes = orc.sim_chunk(net, bits, P, c * 64, 64)
            assert np.array_equal(pg.planes_to_codes(pa, pb), codes), c
            g.mark(c, c * 64, True)
            marks, _ = orc.mark_chunk(net, codes, True)
            assert np.array_equal(pg.plane_to_bytes(g.mark_plane(64)), marks), c


def test_fcount_cli_matches_oracle(tmp_path):
    """The fcount CLI (main.cu's flags and result line) on c17 and a vector file."""
    import subprocess
    from helpers import C17_BENCH
    bench = tmp_path / "c17.bench"
    bench.write_text(C17_BENCH)
    nl = pg.Netlist.read_bench(str(bench))
    V = 256
    bits = pg.gen_vectors(17, 5, V)
    vec = tmp_path / "c17.vec"
    pg.write_vectors(str(vec), bits, V)
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpu-path-delay-coverage_amd",
                       "fcount")
    for robust in (True, False):
        args = [exe, "--bench", str(bench), str(vec)] + (["--robust"] if robust else [])
        print("fcount", robust, flush=True)
        r = subprocess.run(args, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        line = [l for l in r.stdout.splitlines() if l.startswith(str(vec))][-1]
        total = int(line.rsplit(";", 1)[1])
        want = orc.grade(orc.OracleNet(nl.arrays()), bits, V, V, robust)["total"]
        assert total == want, (line, want)


@pytest.mark.parametrize("robust", [True, False])
def test_50k_gates_full_size_against_bitsliced(robust):
    """All 65,536 pairs of the 50k-gate workload with its recorded chunk,
    lane and walk tuning, bit-exact against the bit-sliced CPU grader (an independent
    second CPU implementation, itself checked against the oracle)."""
    say("50k full vs bit-sliced", robust)
    nl = pg.Netlist.generate(GATES_50K.netlist)
    P = GATES_50K.pairs
    bits = pg.gen_vectors(GATES_50K.vector_seed, nl.info()["num_pi"], P)
    ref = bitsliced_grade(nl, bits, P, GATES_50K.chunk, robust)
    with pg.Grader(nl) as g:
        g.set_vectors(bits, P)
        g.set_tuning(pg.TUNE_LANES, GATES_50K.lanes)
        g.set_tuning(pg.TUNE_WALK, GATES_50K.walk)
        g.set_chunk(GATES_50K.chunk)
        for rep in range(2):
            got = dict(total=g.grade(robust), gate_counts=g.gate_counts())
            got["first_t0"], got["first_t1"] = g.first()
            assert_same(got, ref, f"rep {rep}")
    if not robust:
        assert ref["total"] > 0


def test_2m_gates_recorded_tuning():
    """The 2,000,000-gate netlist on 262,144 pairs with its recorded tuning
    (2 lanes x 131,072-pair chunks x 2 words per lane, reverse tiles per wave
    chosen per level: 8 on the large levels) against 1 lane x 32,768-pair
    chunks x 1 word with one tile per reverse wave: same total, first
    detections and counts."""
    from pdfgrade.configs import GATES_2M
    say("2m gates recorded tuning")
    assert (GATES_2M.lanes, GATES_2M.chunk, GATES_2M.words_per_lane) == (2, 131072, 2)
    nl = pg.Netlist.generate(GATES_2M.netlist)
    P = 262144
    bits = pg.gen_vectors(GATES_2M.vector_seed, nl.info()["num_pi"], P)
    with pg.Grader(nl) as g:
        g.set_vectors(bits, P)
        g.set_chunk(32768)
        g.set_tuning(pg.TUNE_REV_TILES, 1)
        ref = dict(total=g.grade(True), gate_counts=g.gate_counts())
        ref["first_t0"], ref["first_t1"] = g.first()
    with pg.Grader(nl) as g:
        g.set_vectors(bits, P)
        g.set_chunk(GATES_2M.chunk)
        g.set_tuning(pg.TUNE_LANES, GATES_2M.lanes)
        g.set_tuning(pg.TUNE_WORDS_PER_LANE, GATES_2M.words_per_lane)
        for rep in range(2):
            total = g.grade(True)
            f0, f1 = g.first()
            assert_same(dict(total=total, first_t0=f0, first_t1=f1, gate_counts=g.gate_counts()), ref, rep)


def test_2m_gates_prefix_against_bitsliced():
    """The 2,000,000-gate / 200-level netlist on 4,096 pairs (as a vector set
    of its own), bit-exact against the bit-sliced CPU grader, both modes."""
    from pdfgrade.configs import GATES_2M
    say("2m gates vs bit-sliced")
    nl = pg.Netlist.generate(GATES_2M.netlist)
    P = 4096
    bits = pg.gen_vectors(GATES_2M.vector_seed, nl.info()["num_pi"], P)
    for robust in (True, False):
        ref = bitsliced_grade(nl, bits, P, 2048, robust)
        assert_same(gpu_grade(nl, bits, P, 2048, robust), ref, f"robust {robust}")
