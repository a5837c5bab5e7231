"""CPU restatement of the learned pair ranking's grader-side pieces: the 512
pair features, each pair's exact solo gain, and BF16 rounding.

Test infrastructure only: imported by tests/ as the checker of the HIP path.
None of this has a counterpart in the reference (it feeds the learned
ranking built on top of the grading path).  Parity is pinned by:
  * bf16_round  -- torch's float32 -> bfloat16 conversion (an independent
                   round-to-nearest-even implementation);
  * features    -- closed forms (all-stable pairs give 0; all-toggling pairs
                   give 1 on every PI bucket and on level 0, forward marks
                   included) and a pure-Python loop over brute-force frames
                   and forward marks (brute.frames, brute.sensitizes);
  * pair_gains  -- brute-force enumeration of fully marked PI->PO paths
                   (brute.marks_for_pattern + brute.suffix_paths).
"""
from __future__ import annotations

import numpy as np

from . import oracle as orc

FEATURES = 512
LEVELS = 128          # levels with a transition feature and a forward-mark feature
PI_BUCKETS = 128


def bf16_round(x: np.ndarray) -> np.ndarray:
    """float32 -> bfloat16 bits (uint16), round to nearest even (finite x)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_to_f32(h: np.ndarray) -> np.ndarray:
    return (np.asarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def pi_bucket_bounds(num_pi: int, b: int):
    """PIs [floor(b I / 128), floor((b + 1) I / 128)) form bucket b (empty
    buckets when I < 128)."""
    return b * num_pi // PI_BUCKETS, (b + 1) * num_pi // PI_BUCKETS


def features(arrs: dict, codes: np.ndarray, fwd_marks: np.ndarray, max_level: int) -> np.ndarray:
    """Feature bits (uint16, BF16) of shape (pairs, 512) from the node codes
    (n, pairs) of a simulation (code bit 1 = transition) and the forward
    marks (n, pairs) of the mark pass before removal:

      [0, 128)    transitions / nodes on level f
      [128, 256)  forward marks / nodes on level f - 128
                  (both for levels <= min(max_level, L, 127))
      [256, 384)  toggling PIs / PIs of bucket b
      [384, 512)  sum of fanout counts of toggling PIs / that of bucket b
    Each value is float32(count) * float32(1 / float32(denominator)).
    """
    n, P = codes.shape
    t = ((codes >> 1) & 1).astype(np.int64)
    level_off = np.asarray(arrs["level_off"], dtype=np.int64)
    L = len(level_off) - 2
    num_pi = int(arrs["num_pi"])
    fout_off = np.asarray(arrs["fout_off"], dtype=np.int64)
    nfo = (fout_off[1:] - fout_off[:-1])[:num_pi]
    out = np.zeros((P, FEATURES), np.float32)

    def put(f, count, denom):
        scale = np.float32(1.0) / np.float32(denom) if denom > 0 else np.float32(0.0)
        out[:, f] = count.astype(np.float32) * scale

    m = np.asarray(fwd_marks, dtype=np.int64)
    for f in range(min(max_level, L, LEVELS - 1) + 1):
        lo, hi = level_off[f], level_off[f + 1]
        put(f, t[lo:hi].sum(axis=0), hi - lo)
        put(LEVELS + f, m[lo:hi].sum(axis=0), hi - lo)
    for b in range(PI_BUCKETS):
        lo, hi = pi_bucket_bounds(num_pi, b)
        put(256 + b, t[lo:hi].sum(axis=0), hi - lo)
        w = nfo[lo:hi]
        put(384 + b, (w[:, None] * t[lo:hi]).sum(axis=0), int(w.sum()))
    return bf16_round(out)


def pair_gains(arrs: dict, a_bits, num_pairs: int, robust: bool, b_bits=None) -> np.ndarray:
    """Each pair graded alone against an empty detection history: its number
    of marked PI->PO paths (the C oracle's cover DP with every marked node
    new), saturated to uint32."""
    net = orc.OracleNet(arrs)
    codes = orc.sim_chunk(net, a_bits, num_pairs, 0, num_pairs, b_bits=b_bits)
    marks, _ = orc.mark_chunk(net, codes, robust)
    out = np.zeros(num_pairs, np.uint32)
    gc = np.zeros(net.n, np.uint64)
    for p in range(num_pairs):
        first = np.full(net.n, p, np.uint32)
        tot, _ = orc.cover_chunk(net, np.ascontiguousarray(marks[:, p:p + 1]), p, first, first.copy(), gc)
        out[p] = min(tot, 0xFFFFFFFF)
    return out
