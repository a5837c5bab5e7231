"""Writes the golden fixtures of tests/golden/: exact grading outputs.

* tiny_200g.json: the 200-gate workload (2-input gates, 1,024 iid pairs);
* wide_600g.json: 632 nodes with fan-in 1..4 and every gate type, 2,048
  pairs over hold-biased vectors (a PI keeps its previous value with
  probability 0.6, so many paths are sensitized).  Robust and non-robust
  grading differ here (82 against 114 paths, 15 gates' counts differ), so the
  fixture pins the non-robust sensitization table as well as the robust one.

The values are produced twice -- by the brute-force path enumerator
(oracle/brute.py) and by the C oracle at several chunk sizes -- and written
only if all agree.  The netlist and vector digests pin the seeded generators
as well.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "gpu-path-delay-coverage_amd"), os.path.join(ROOT, "tests")]

import numpy as np  # noqa: E402

import pdfgrade as pg  # noqa: E402
from pdfgrade.configs import TINY  # noqa: E402

from helpers import biased_vectors, brute_grade, oracle_grade  # noqa: E402

# the wide fixture's recipe: netlist, vectors, pairs, the chunk whose per-chunk
# totals are recorded
WIDE_NETLIST = pg.GenConfig(seed=60, num_pi=32, num_po=96, num_gates=600, levels=10,
                            fanin_p=(0.45, 0.35, 0.12, 0.08), prev_level_frac=0.7, fanout_alpha=2.2,
                            fanout_cap=16)
WIDE_VECTOR_SEED, WIDE_P_HOLD, WIDE_PAIRS, WIDE_CHUNK = 60, 0.6, 2048, 512


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def netlist_digest(arrs) -> str:
    return digest(arrs["type"], arrs["po"], arrs["fin_off"], arrs["fin"])


def tiny_inputs():
    nl = pg.Netlist.generate(TINY.netlist)
    arrs = nl.arrays()
    return nl, arrs, pg.gen_vectors(TINY.vector_seed, arrs["num_pi"], TINY.pairs), TINY.pairs, 256


def wide_inputs():
    nl = pg.Netlist.generate(WIDE_NETLIST)
    arrs = nl.arrays()
    bits = biased_vectors(WIDE_VECTOR_SEED, arrs["num_pi"], WIDE_PAIRS, p_one=0.5, p_hold=WIDE_P_HOLD)
    return nl, arrs, bits, WIDE_PAIRS, WIDE_CHUNK


# fixture file -> (workload name, inputs)
FIXTURES = {"tiny_200g.json": (TINY.name, tiny_inputs), "wide_600g.json": ("wide-600g-hold-biased", wide_inputs)}


def build(name: str, inputs) -> dict:
    nl, arrs, bits, pairs, chunk = inputs()
    out = {"workload": name, "netlist_sha256": netlist_digest(arrs), "vectors_sha256": digest(bits),
           "n": int(arrs["n"]), "edges": int(len(arrs["fin"])), "pairs": pairs, "chunk": chunk, "results": {}}
    for robust in (True, False):
        b = brute_grade(nl, bits, pairs, robust)
        for c in sorted({pairs, chunk, 64}):
            o = oracle_grade(nl, bits, pairs, c, robust)
            assert o["total"] == b["total"]
            assert list(o["gate_counts"]) == b["gate_counts"]
            assert list(o["first_t0"]) == b["first_t0"] and list(o["first_t1"]) == b["first_t1"]
            assert o["flags"] == 0
        key = "robust" if robust else "nonrobust"
        o = oracle_grade(nl, bits, pairs, chunk, robust)
        out["results"][key] = {"total": b["total"], "gate_counts": b["gate_counts"],
                               "first_t0": b["first_t0"], "first_t1": b["first_t1"],
                               "chunk_totals": [int(x) for x in o["chunk_totals"]],
                               "per_pattern_total": b["per_pattern"]}
    return out


def main():
    for fname, (name, inputs) in FIXTURES.items():
        out = build(name, inputs)
        path = os.path.join(HERE, fname)
        with open(path, "w") as f:
            json.dump(out, f, separators=(",", ":"))
        print("wrote", path, {k: v["total"] for k, v in out["results"].items()})


if __name__ == "__main__":
    main()
