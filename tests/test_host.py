"""Host-side checks that need no GPU: the C ABI library loads and exports every
symbol of include/pdfgrade.h, and the netlist/vector host code behaves."""
import os
import re
import sys
import tempfile

import numpy as np
import pytest

import pdfgrade as pg
from pdfgrade.configs import GATES_50K, TINY

from helpers import C17_BENCH, small_config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "pdfgrade.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pdf_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) >= 25
    lib = pg.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes signatures cover exactly the declared functions
    assert sorted(pg.SIGNATURES) == names
    assert lib.pdf_abi_version() == 1


def _arr(nl):
    return nl.arrays()


def test_generator_is_deterministic_and_seeded():
    a = _arr(pg.Netlist.generate(TINY.netlist))
    b = _arr(pg.Netlist.generate(TINY.netlist))
    for k in ("type", "po", "fin_off", "fin"):
        assert np.array_equal(a[k], b[k])
    c = _arr(pg.Netlist.generate(pg.GenConfig(**{**TINY.netlist.__dict__, "seed": 99})))
    assert not np.array_equal(a["fin"][:50], c["fin"][:50])


def check_structure(arrs, cfg):
    n, I = arrs["n"], arrs["num_pi"]
    assert n == cfg.num_pi + cfg.num_gates and I == cfg.num_pi
    assert arrs["L"] == cfg.levels
    assert int(arrs["po"].sum()) == cfg.num_po
    lv = arrs["level"]
    assert np.all(np.diff(lv.astype(np.int64)) >= 0)
    assert np.all(arrs["type"][:I] == pg.INPT) and np.all(arrs["type"][I:] != pg.INPT)
    fo, fi = arrs["fin_off"], arrs["fin"]
    nfi = np.diff(fo)
    assert nfi.max() <= 4 and nfi[I:].min() >= 1
    for g in range(I, n):
        ins = fi[fo[g]:fo[g + 1]]
        assert len(set(ins.tolist())) == len(ins)
        assert lv[ins].max() == lv[g] - 1  # at least one fanin on the previous level
        if nfi[g] == 1:
            assert arrs["type"][g] in (pg.NOT, pg.BUFF)
        else:
            assert arrs["type"][g] in (pg.AND, pg.NAND, pg.OR, pg.NOR, pg.XOR, pg.XNOR)
    fanout = np.diff(arrs["fout_off"])
    if cfg.fanout_cap:
        assert fanout.max() <= cfg.fanout_cap
    return nfi[I:], fanout


def test_generator_tiny_structure():
    arrs = _arr(pg.Netlist.generate(TINY.netlist))
    nfi, _ = check_structure(arrs, TINY.netlist)
    assert 0.05 < np.mean(nfi == 1) < 0.2
    lv = arrs["level"]
    fo, fi = arrs["fin_off"], arrs["fin"]
    prev = sum(int(lv[f] == lv[g] - 1) for g in range(arrs["num_pi"], arrs["n"]) for f in fi[fo[g]:fo[g + 1]])
    assert 0.6 < prev / len(fi) < 0.9


def test_generator_50k_structure_and_path_bound():
    nl = pg.Netlist.generate(GATES_50K.netlist)
    inf = nl.info()
    assert inf["structural_paths"] < 2 ** 63
    arrs = nl.arrays()
    nfi, fanout = check_structure(arrs, GATES_50K.netlist)
    hist = np.bincount(nfi, minlength=5)[1:5] / len(nfi)
    assert np.allclose(hist, (0.1, 0.6, 0.2, 0.1), atol=0.02)
    # heavy-tailed fanout: some nodes far above the mean
    assert fanout.max() >= 20 * max(1.0, fanout.mean()) / 4


def test_vectors_packed_and_deterministic():
    v = pg.gen_vectors(5, 7, 200)
    assert v.shape == (7, 4)
    assert np.array_equal(v, pg.gen_vectors(5, 7, 200))
    assert np.all(v[:, 3] >> np.uint64(8) == 0)  # bits past vector 199 are zero
    ones = sum(bin(int(x)).count("1") for x in v[:, :3].ravel())
    assert 0.4 < ones / (7 * 192) < 0.6


def write_tmp(text, suffix):
    f = tempfile.NamedTemporaryFile("w", suffix=suffix, delete=False)
    f.write(text)
    f.close()
    return f.name


def test_bench_reader_c17():
    path = write_tmp(C17_BENCH, ".bench")
    try:
        nl = pg.Netlist.read_bench(path)
    finally:
        os.unlink(path)
    a = nl.arrays()
    assert a["n"] == 11 and a["num_pi"] == 5 and a["L"] == 3
    assert list(a["type"][5:]) == [pg.NAND] * 6
    assert list(a["po"]) == [0] * 9 + [1, 1]
    fo, fi = a["fin_off"], a["fin"]
    assert [list(fi[fo[g]:fo[g + 1]]) for g in range(5, 11)] == [[0, 2], [2, 3], [1, 6], [6, 4], [5, 7], [7, 8]]
    assert nl.info()["structural_paths"] == 11


@pytest.mark.parametrize("text,code", [
    ("INPUT(a)\nOUTPUT(c)\nb = AND(a, c)\nc = NOT(b)\n", pg.ERR_NETLIST),   # cycle
    ("INPUT(a)\nOUTPUT(b)\nb = AND(a, z)\n", pg.ERR_NETLIST),              # undriven
    ("INPUT(a)\nOUTPUT(b)\nb = FOO(a)\n", pg.ERR_IO),                      # unknown gate
    ("INPUT(a)\nOUTPUT(q)\nb = NOT(a)\n", pg.ERR_NETLIST),                 # undefined output
])
def test_bench_reader_errors(text, code):
    path = write_tmp(text, ".bench")
    try:
        with pytest.raises(pg.PdfError) as e:
            pg.Netlist.read_bench(path)
        assert e.value.code == code
    finally:
        os.unlink(path)


def test_level_round_trip(tmp_path):
    nl = pg.Netlist.generate(TINY.netlist)
    p = str(tmp_path / "tiny.level")
    nl.write_level(p)
    back = pg.Netlist.read_level(p).arrays()
    a = nl.arrays()
    for k in ("type", "po", "fin_off", "fin", "fout_off", "fout", "level_off"):
        assert np.array_equal(a[k], back[k])


@pytest.mark.parametrize("types,po,fanins,why", [
    ([pg.AND, pg.INPT], [0, 1], [[1], []], "INPT"),
    ([pg.INPT, pg.AND, pg.AND], [0, 0, 1], [[], [2], [0]], "earlier"),
    ([pg.INPT, pg.AND], [0, 1], [[], []], "no fanins"),
    ([pg.INPT, pg.AND, pg.OR], [0, 1, 1], [[], [0], [0]], ""),  # valid
])
def test_netlist_validation(types, po, fanins, why):
    if not why:
        nl = pg.Netlist.from_arrays(1, types, po, fanins)
        assert nl.info()["levels"] == 1
        return
    with pytest.raises(pg.PdfError) as e:
        pg.Netlist.from_arrays(1, types, po, fanins)
    assert e.value.code == pg.ERR_NETLIST and why in str(e.value)


def test_gen_toggle_pairs_statistics():
    """A equals pdf_gen_vectors' output; the flip fraction per pair follows
    the log-uniform flip probability; bits past the count are zero."""
    I, P = 256, 20000
    a, b = pg.gen_toggle_pairs(7, I, P, 1.0 / 256, 0.5)
    np.testing.assert_array_equal(a, pg.gen_vectors(7, I, P))
    a2, b2 = pg.gen_toggle_pairs(7, I, P, 1.0 / 256, 0.5)
    np.testing.assert_array_equal(b, b2)
    flips = np.unpackbits((a ^ b).view(np.uint8).reshape(I, -1), axis=1, bitorder="little")[:, :P].sum(axis=0)
    # E[q] for q log-uniform on [lo, hi] is (hi - lo) / ln(hi / lo)
    eq = (0.5 - 1 / 256) / np.log(128)
    assert abs(flips.mean() / I - eq) < 0.01
    # P(no flip) = E[(1 - q)^I] ~ E1(1) / ln(hi / lo) (exponential integral)
    from scipy.special import exp1
    assert abs((flips == 0).mean() - (exp1(1.0) - exp1(128.0)) / np.log(128)) < 0.01
    tail = (a ^ b)[:, -1] >> np.uint64(P % 64)
    assert not tail.any()
    # q = 1 flips every PI of every pair
    a1, b1 = pg.gen_toggle_pairs(3, 8, 100, 1.0, 1.0)
    want = np.array([[2**64 - 1, (1 << 36) - 1]] * 8, np.uint64)
    np.testing.assert_array_equal(a1 ^ b1, want)
    with pytest.raises(pg.PdfError):
        pg.gen_toggle_pairs(1, 8, 64, 0.0, 0.5)


def test_vector_file_round_trip(tmp_path):
    I, V = 7, 150
    bits = pg.gen_vectors(11, I, V)
    path = str(tmp_path / "v.vec")
    pg.write_vectors(path, bits, V)
    got, n = pg.read_vectors(path, I)
    assert n == V
    np.testing.assert_array_equal(got, bits)
    # comments, blank lines and spaces are ignored
    text = "# header\n\n0 1 0 1 0 1 1  # first\n1111111\n"
    (tmp_path / "w.vec").write_text(text)
    got, n = pg.read_vectors(str(tmp_path / "w.vec"), I)
    assert n == 2
    assert [int(got[i][0] & 1) for i in range(I)] == [0, 1, 0, 1, 0, 1, 1]
    assert all(int(got[i][0] >> 1) & 1 for i in range(I))


def test_vector_file_errors(tmp_path):
    (tmp_path / "bad.vec").write_text("01x1\n")
    with pytest.raises(pg.PdfError):
        pg.read_vectors(str(tmp_path / "bad.vec"), 4)
    (tmp_path / "wide.vec").write_text("0101\n011\n")
    with pytest.raises(pg.PdfError):
        pg.read_vectors(str(tmp_path / "wide.vec"), 4)
    with pytest.raises(pg.PdfError):
        pg.read_vectors(str(tmp_path / "missing.vec"), 4)
    # more vectors than the caller's capacity is an error, with no write past it
    import ctypes as C
    (tmp_path / "many.vec").write_text("1111\n" * 70)
    buf = np.zeros((4, 2), np.uint64)
    got = C.c_uint64()
    rc = pg.lib().pdf_vectors_read(str(tmp_path / "many.vec").encode(), 4, buf[:, :1].copy().ctypes.data, 64,
                                   C.byref(got))
    assert rc == pg.ERR_ARG


@pytest.mark.parametrize("seed", range(4))
def test_library_segments_equal_oracle(seed):
    """pdf_netlist_segments (the grader's segment list) equals the oracle's
    enumeration entry for entry, K = 1..8."""
    from oracle import oracle as orc
    nl = pg.Netlist.generate(small_config(seed, num_gates=50, levels=7, wide=seed % 2 == 0,
                                          alpha=2.2 if seed % 2 else 0.0))
    net = orc.OracleNet(nl.arrays())
    for K in range(1, 9):
        assert np.array_equal(nl.segments(K), orc.segments(net, K)), K
    with pytest.raises(pg.PdfError):
        nl.segments(9)


def test_hold_vectors_layout_and_hold_rate():
    """pdfgrade.hold_vectors: PI-major packing like gen_vectors (bit v % 64
    of word v // 64 is vector v), vector 0 drawn fresh, and each PI equal to
    its previous value at rate p_hold + (1 - p_hold) / 2."""
    from helpers import pack, unpack

    V, I, p = 20000, 9, 0.8
    bits = pg.hold_vectors(3, I, V, p)
    assert bits.shape == (I, (V + 63) // 64) and bits.dtype == np.uint64
    vals = unpack(bits, V)
    assert np.array_equal(pack(vals), bits)  # no stray bits past vector V - 1
    same = (vals[1:] == vals[:-1]).mean()
    assert abs(same - (p + (1 - p) / 2)) < 0.01
    assert abs(vals.mean() - 0.5) < 0.02
    assert np.array_equal(pg.hold_vectors(3, I, V, p), bits)
    assert not np.array_equal(pg.hold_vectors(4, I, V, p), bits)


def test_bench_walk_byte_model_counts_far_fanins():
    """bench.walk_model_bytes on the 50k-gate netlist: the fanin edges it
    counts as outside the walk's LDS window, and the nodes whose records are
    written for them, equal a direct loop over the gates (fanin 0, or any
    fanin of AND/NAND/OR/NOR, below the node's level end minus the
    2,048-node window of 80-byte slots), and the byte totals follow."""
    import bench

    nl = pg.Netlist.generate(GATES_50K.netlist)
    a = nl.arrays()
    m = bench.walk_model_bytes(nl, 65536, 4)
    cap = m["window_nodes"]
    assert cap == (160 * 1024) // 80 == 2048 and m["record_bytes"] == 48 and m["tiles"] == 256
    far = 0
    kept = set()
    lo = a["level_off"]
    for g in range(a["n"]):
        t = a["type"][g]
        hi = int(lo[a["level"][g] + 1])
        for j, f in enumerate(a["fin"][a["fin_off"][g]:a["fin_off"][g + 1]]):
            if (j == 0 or pg.AND <= t <= pg.NOR) and int(f) < hi - cap:
                far += 1
                kept.add(int(f))
    assert m["far_fanins"] == far > 0
    assert m["kept_records"] == len(kept) < a["n"]
    assert m["record_writes"] == len(kept) * 256 * 48
    assert m["far_fanin_reads"] == far * 256 * 48
    assert m["bit_bytes"] == 2 * a["n"] * 256
    assert m["bytes"] == m["record_writes"] + m["far_fanin_reads"] + m["pi_reads"] + m["bit_bytes"]


def test_bench_tune_option_names_knobs(monkeypatch):
    """bench.py --tune KNOB=VALUE: each knob name is a pdfgrade TUNE_*
    constant in lower case, repeatable; the recorded workloads' tunings use
    values the grader accepts."""
    import bench
    from pdfgrade.configs import WORKLOADS

    monkeypatch.setattr(sys, "argv", ["bench.py", "--tune", "fwd_order=1", "--tune", "rev_compact=0"])
    args = bench.parse()
    assert args.tune == ["fwd_order=1", "rev_compact=0"]
    for kv in args.tune:
        knob, value = kv.split("=", 1)
        assert isinstance(getattr(pg, "TUNE_" + knob.upper()), int) and int(value) in (0, 1)
    for wl in WORKLOADS.values():
        assert wl.walk in (0, 1, 2, 4, 8) and wl.rev_compact in (0, 1) and wl.words_per_lane in (1, 2)
        assert not (wl.walk and wl.rev_compact), wl.name  # the walk has its own reverse pass
