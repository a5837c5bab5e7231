"""Synthetic workloads graded by the tests and the benchmark.

Each entry gives the netlist recipe, the vector seed, the number of pairs and
the pairs per chunk.  Pairs are consecutive vectors (V[p], V[(p+1) mod P]) as
in simkernel.cuh:115-124 unless a workload says otherwise.
"""
from __future__ import annotations

from dataclasses import dataclass

from . import GenConfig

# fan-in distributions
TINY_FANIN = (0.1, 0.9, 0.0, 0.0)          # 2-input gates plus 10% NOT/BUFF
WIDE_FANIN = (0.1, 0.6, 0.2, 0.1)          # fan-in 1..4


@dataclass(frozen=True)
class Workload:
    name: str
    netlist: GenConfig
    vector_seed: int
    pairs: int
    chunk: int
    robust: bool = True
    lanes: int = 1  # chunks graded concurrently (pdf_grader_autotune's choice, see bench.py)
    words_per_lane: int = 1  # PDF_TUNE_WORDS_PER_LANE (2 needs chunks of a multiple of 65,536 pairs)
    use_graph: int = 1  # PDF_TUNE_USE_GRAPH: replay a captured graph (1) or launch directly (0)
    walk: int = 0  # PDF_TUNE_WALK: words per level-walk tile (0: one launch per level and pass)
    rev_compact: int = 0  # PDF_TUNE_REV_COMPACT: reverse levels over the forward pass's marked-item lists


TINY = Workload(
    name="tiny-200g",
    netlist=GenConfig(seed=1, num_pi=32, num_po=16, num_gates=200, levels=12, fanin_p=TINY_FANIN,
                      prev_level_frac=0.7),
    vector_seed=101, pairs=1024, chunk=1024)

GATES_50K = Workload(
    name="single-50k-gates",
    netlist=GenConfig(seed=2, num_pi=256, num_po=128, num_gates=50_000, levels=60, fanin_p=WIDE_FANIN,
                      prev_level_frac=0.7, fanout_alpha=2.2, fanout_cap=64),
    # the level walk, one 65,536-pair chunk in 4-word tiles (256 workgroups,
    # one per CU): 0.54-0.55 ms a grade on an MI355X against 0.75-0.77 ms for
    # per-level launches on 2 lanes x 32,768 pairs (graph replay), the
    # fastest configuration without the walk
    vector_seed=102, pairs=65_536, chunk=65_536, lanes=1, use_graph=1, walk=4)

GATES_2M = Workload(
    name="single-2m-gates",
    # 200 levels put the structural PI->PO path count far above 2^64; only
    # sensitized (marked) paths are counted, and the uint64 accumulation is
    # overflow-checked, so no limit is applied here
    netlist=GenConfig(seed=4, num_pi=4096, num_po=2048, num_gates=2_000_000, levels=200, fanin_p=WIDE_FANIN,
                      prev_level_frac=0.7, fanout_alpha=2.2, fanout_cap=64, path_limit_log2=0),
    # 2 lanes x 131,072-pair chunks x 2 words per lane, the compacted reverse
    # pass: 157-161 ms a grade on an MI355X against 178-186 ms with dense
    # reverse launches and 291 ms for 1 lane x 32,768 x 1 (the chunk buffers
    # take 234 GB)
    vector_seed=104, pairs=1_048_576, chunk=131_072, lanes=2, words_per_lane=2, rev_compact=1)


@dataclass(frozen=True)
class BatchWorkload:
    """Many circuits graded together: circuit i has netlist seed 3000 + i and
    vector seed 13000 + i; its gate count (log-uniform in [min_gates,
    max_gates]) and pair count (uniform in [min_pairs, max_pairs], rounded to
    a multiple of 64) come from a numpy Generator seeded with draw_seed;
    levels = round(8 log2 gates)."""
    name: str
    count: int
    min_gates: int = 1000
    max_gates: int = 200_000
    min_pairs: int = 4096
    max_pairs: int = 65536
    draw_seed: int = 3
    robust: bool = True

    def members(self):
        """[(GenConfig, vector_seed, pairs)] of every circuit."""
        import numpy as np

        rng = np.random.default_rng(self.draw_seed)
        gates = np.exp(rng.uniform(np.log(self.min_gates), np.log(self.max_gates), self.count)).round().astype(int)
        pairs = (np.round(rng.integers(self.min_pairs, self.max_pairs + 1, self.count) / 64) * 64).astype(int)
        out = []
        for i in range(self.count):
            g = int(gates[i])
            levels = int(round(8 * np.log2(g)))
            # above ~100k gates and 130 levels the structural path count passes
            # 2^63; only marked paths are counted (overflow-checked), so no limit
            cfg = GenConfig(seed=3000 + i, num_pi=256, num_po=128, num_gates=g, levels=levels, fanin_p=WIDE_FANIN,
                            prev_level_frac=0.7, fanout_alpha=2.2, fanout_cap=64, path_limit_log2=0)
            out.append((cfg, 13000 + i, int(pairs[i])))
        return out


BATCH_256 = BatchWorkload(name="batch-256-circuits", count=256)

COVER_HEAVY = Workload(
    name="cover-heavy-50k-gates",
    # 50,000 gates in 16 levels, 70% of them NOT/BUFF (a fan-in-1 gate passes
    # its input's mark on), 4,096 POs: ~2,800 robust paths over 65,536 iid
    # pairs against the 50k workload's 4, so the merge and the cover DP
    # carry the load (the cover takes ~0.4 of ~1.2 ms a grade)
    netlist=GenConfig(seed=5, num_pi=256, num_po=4096, num_gates=50_000, levels=16,
                      fanin_p=(0.7, 0.25, 0.05, 0.0), prev_level_frac=0.7, fanout_alpha=2.2, fanout_cap=64),
    # pdf_grader_autotune's choice on an MI355X: one chunk, per-level
    # launches (the level walk is slower here: 3,100 gates a level), direct
    # launches (1.22 ms a grade)
    vector_seed=7, pairs=65_536, chunk=65_536, lanes=1, use_graph=0)

WORKLOADS = {w.name: w for w in (TINY, GATES_50K, GATES_2M, COVER_HEAVY)}
BATCH_WORKLOADS = {BATCH_256.name: BATCH_256}
