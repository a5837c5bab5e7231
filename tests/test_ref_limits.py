"""The reference's resource limits (CPU): the merge window of the oracle
against a numpy restatement, the segment-stride count of the library against
the oracle's segment-by-segment count and the figure quoted for the 2M-gate
workload, and the 16-bit flags on fan trees whose path counts are known in
closed form."""
import numpy as np
import pytest

import pdfgrade as pg
from oracle import oracle as orc

from helpers import biased_vectors, small_config


@pytest.mark.parametrize("S", [0, 1, 19, 20, 21, 65535, 1_310_699, 1_310_700, 1_310_701, 1_376_235, 1_376_236,
                               2_004_096, 2_621_400, 3_000_000, 4_000_000])
def test_unmerged_segments_library_matches_oracle(S):
    assert pg.lib().pdf_ref_unmerged_segments(S) == orc.ref_unmerged_segments(S)


def test_unmerged_segments_known_answers():
    # one launch reaches 65,535 x 20 segments; the 2M-gate netlist's
    # 2,004,096 nodes leave 693,396 unmerged (count += BLOCK_STEP,
    # mergekernel.cuh:243)
    assert orc.ref_unmerged_segments(1_310_700) == 0
    assert orc.ref_unmerged_segments(2_004_096) == 693_396


def windowed_first(codes, marks, P, chunk, window):
    """Earliest marked T0 / T1 pattern per node, each chunk contributing only
    its first `window` patterns (mergekernel.cuh:125-130), in numpy."""
    n = codes.shape[0]
    f = [np.full(n, 0xFFFFFFFF, np.uint64), np.full(n, 0xFFFFFFFF, np.uint64)]
    rel = np.arange(P) % chunk
    ok = (marks != 0) & (rel < window)[None, :]
    pid = np.arange(P, dtype=np.uint64)[None, :]
    for pol in (0, 1):
        hit = ok & ((codes & 1) == pol)
        cand = np.where(hit, pid, np.uint64(0xFFFFFFFF))
        f[pol] = cand.min(axis=1)
    return f[0].astype(np.uint32), f[1].astype(np.uint32)


@pytest.mark.parametrize("seed", range(4))
def test_oracle_merge_window_against_numpy(seed):
    nl = pg.Netlist.generate(small_config(seed + 40, num_pi=10, num_gates=120, levels=7))
    net = orc.OracleNet(nl.arrays())
    P, chunk, window = 1536, 512, 3 + 29 * seed
    bits = biased_vectors(seed, nl.info()["num_pi"], P, p_hold=0.6)
    codes = orc.sim_chunk(net, bits, P, 0, P)
    for robust in (True, False):
        marks, _ = orc.mark_chunk(net, codes, robust)
        f0, f1 = windowed_first(codes, marks, P, chunk, window)
        got = orc.grade(net, bits, P, chunk, robust, merge_window=window)
        assert np.array_equal(got["first_t0"], f0) and np.array_equal(got["first_t1"], f1)
        full = orc.grade(net, bits, P, chunk, robust)
        assert orc.grade(net, bits, P, chunk, robust, merge_window=chunk)["total"] == full["total"]
        # the window changes which pairs detect first
        if seed == 0:
            assert not np.array_equal(full["first_t0"], f0) or not np.array_equal(full["first_t1"], f1)


def fan_tree(width: int, extra_root: bool = False):
    """PI -> [BUFF root] -> `width` BUFFs -> `width` BUFF primary outputs each:
    width^2 PI->PO paths, all marked whenever the PI toggles."""
    types, fanins, po = [pg.INPT], [[]], [0]
    src = 0
    if extra_root:
        types.append(pg.BUFF); fanins.append([0]); po.append(0)
        src = 1
    mids = []
    for _ in range(width):
        types.append(pg.BUFF); fanins.append([src]); po.append(0)
        mids.append(len(types) - 1)
    for m in mids:
        for _ in range(width):
            types.append(pg.BUFF); fanins.append([m]); po.append(1)
    return pg.Netlist.from_arrays(1, types, po, fanins)


def toggle_vectors(pattern, P):
    vals = np.array([[pattern[i % len(pattern)]] for i in range(P)], np.uint8)
    from helpers import pack
    return pack(vals)


def test_u16_flags_on_fan_trees():
    # the PI rises at pair 0 and falls at pair 1: every node is new there, and
    # c(PI) = width^2
    P = 64
    bits = toggle_vectors([0, 1, 0, 0], P)
    big = orc.grade(orc.OracleNet(fan_tree(260).arrays()), bits, P, P, True)
    assert big["total"] == 2 * 260 * 260
    assert big["flags"] & orc.FLAG_EXCEEDS_U16 and big["flags"] & orc.FLAG_EXCEEDS_U16_DP
    small = orc.grade(orc.OracleNet(fan_tree(200).arrays()), bits, P, P, True)
    assert small["total"] == 2 * 200 * 200 and small["flags"] == 0
    # a root BUFF under the PI, toggling again and again: at the later
    # toggles nothing is new and the root's h (marked paths without a new
    # node) is width^2, checked by the full flag only; here the first toggle
    # already trips both
    bits = toggle_vectors([0, 1, 0, 1], P)
    rooted = orc.grade(orc.OracleNet(fan_tree(200, extra_root=True).arrays()), bits, P, P, True)
    assert rooted["flags"] == 0
    rooted = orc.grade(orc.OracleNet(fan_tree(260, extra_root=True).arrays()), bits, P, P, True)
    assert rooted["flags"] & orc.FLAG_EXCEEDS_U16 and rooted["flags"] & orc.FLAG_EXCEEDS_U16_DP
