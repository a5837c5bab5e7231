"""The committed golden fixtures (tests/golden/*.json, written by
tests/golden/make_golden.py after the brute-force enumerator and the C oracle
agreed) against today's generators, oracle and -- on a GPU -- the HIP path:
the staged chunk loop, the fused grade with per-level launches and the level
walk."""
import json
import os

import pytest

import pdfgrade as pg

from golden.make_golden import FIXTURES, digest, netlist_digest
from helpers import oracle_grade

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = sorted(FIXTURES)
MODES = ["robust", "nonrobust"]


def load(fname):
    with open(os.path.join(HERE, fname)) as f:
        return json.load(f)


def inputs(fname):
    return FIXTURES[fname][1]()


@pytest.mark.parametrize("fname", NAMES)
def test_generator_reproduces_fixture_inputs(fname):
    fx = load(fname)
    nl, arrs, bits, pairs, chunk = inputs(fname)
    assert netlist_digest(arrs) == fx["netlist_sha256"]
    assert digest(bits) == fx["vectors_sha256"]
    assert arrs["n"] == fx["n"] and len(arrs["fin"]) == fx["edges"]
    assert pairs == fx["pairs"] and chunk == fx["chunk"]


def test_wide_fixture_separates_the_modes():
    """The wide fixture is there to pin non-robust grading: its results must
    differ from the robust ones (totals, counts and first detections)."""
    fx = load("wide_600g.json")["results"]
    r, n = fx["robust"], fx["nonrobust"]
    assert n["total"] > r["total"] > 50
    assert sum(a != b for a, b in zip(r["gate_counts"], n["gate_counts"])) >= 10
    assert r["first_t0"] != n["first_t0"]
    assert sum(1 for x in r["per_pattern_total"] if x) >= 20


@pytest.mark.parametrize("fname", NAMES)
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("which", ["all", "fixture", "64"])
def test_oracle_matches_fixture(fname, mode, which):
    fxa = load(fname)
    fx = fxa["results"][mode]
    nl, _, bits, pairs, chunk = inputs(fname)
    c = {"all": pairs, "fixture": chunk, "64": 64}[which]
    o = oracle_grade(nl, bits, pairs, c, mode == "robust")
    assert o["total"] == fx["total"]
    assert o["gate_counts"].tolist() == fx["gate_counts"]
    assert o["first_t0"].tolist() == fx["first_t0"] and o["first_t1"].tolist() == fx["first_t1"]
    if c == chunk:
        assert o["chunk_totals"].tolist() == fx["chunk_totals"]


@pytest.mark.gpu
@pytest.mark.parametrize("fname", NAMES)
@pytest.mark.parametrize("mode", MODES)
def test_gpu_matches_fixture(fname, mode):
    fx = load(fname)["results"][mode]
    nl, _, bits, pairs, chunk = inputs(fname)
    robust = mode == "robust"
    with pg.Grader(nl) as g:
        g.set_vectors(bits, pairs)
        g.set_chunk(chunk)
        total, per_chunk = g.grade_stages(robust)
        f0, f1 = g.first()
        assert total == fx["total"] and per_chunk == fx["chunk_totals"]
        assert g.gate_counts().tolist() == fx["gate_counts"]
        assert f0.tolist() == fx["first_t0"] and f1.tolist() == fx["first_t1"]
        for c, walk in ((pairs, 0), (pairs, 4), (chunk, 1), (chunk, 8)):
            g.set_chunk(c)
            g.set_tuning(pg.TUNE_WALK, walk)
            assert g.grade(robust) == fx["total"], (c, walk)
            assert g.gate_counts().tolist() == fx["gate_counts"], (c, walk)
            f0, f1 = g.first()
            assert f0.tolist() == fx["first_t0"] and f1.tolist() == fx["first_t1"], (c, walk)
