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


TINY = Workload(
    name="tiny-200g",
    netlist=GenConfig(seed=1, num_pi=32, num_po=16, num_gates=200, levels=12, fanin_p=TINY_FANIN,
                      prev_level_frac=0.7),
    vector_seed=101, pairs=1024, chunk=1024)

GATES_50K = Workload(
    name="single-50k-gates",
    netlist=GenConfig(seed=2, num_pi=256, num_po=128, num_gates=50_000, levels=60, fanin_p=WIDE_FANIN,
                      prev_level_frac=0.7, fanout_alpha=2.2, fanout_cap=64),
    vector_seed=102, pairs=65_536, chunk=65_536)

GATES_2M = Workload(
    name="single-2m-gates",
    # 200 levels put the structural PI->PO path count far above 2^64; only
    # sensitized (marked) paths are counted, and the uint64 accumulation is
    # overflow-checked, so no limit is applied here
    netlist=GenConfig(seed=4, num_pi=4096, num_po=2048, num_gates=2_000_000, levels=200, fanin_p=WIDE_FANIN,
                      prev_level_frac=0.7, fanout_alpha=2.2, fanout_cap=64, path_limit_log2=0),
    vector_seed=104, pairs=1_048_576, chunk=32_768)


def batch_member(i: int, rng_gates: int, levels: int) -> GenConfig:
    """Circuit i of the 256-circuit batch: the gate count and levels are drawn by
    the caller (log-uniform in [1k, 200k], levels = round(8 log2 gates))."""
    return GenConfig(seed=3000 + i, num_pi=256, num_po=128, num_gates=rng_gates, levels=levels,
                     fanin_p=WIDE_FANIN, prev_level_frac=0.7, fanout_alpha=2.2, fanout_cap=64)


WORKLOADS = {w.name: w for w in (TINY, GATES_50K, GATES_2M)}
