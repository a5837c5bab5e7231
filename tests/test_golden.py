"""The committed 200-gate fixture (tests/golden/tiny_200g.json, written by
tests/golden/make_golden.py after the brute-force enumerator and the C oracle
agreed) against today's generator, oracle and -- on a GPU -- the HIP path."""
import json
import os

import numpy as np
import pytest

import pdfgrade as pg
from pdfgrade.configs import TINY

from golden.make_golden import digest, netlist_digest
from helpers import oracle_grade

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tiny_200g.json")


def load():
    with open(FIXTURE) as f:
        return json.load(f)


def inputs():
    nl = pg.Netlist.generate(TINY.netlist)
    arrs = nl.arrays()
    return nl, arrs, pg.gen_vectors(TINY.vector_seed, arrs["num_pi"], TINY.pairs)


def test_generator_reproduces_fixture_inputs():
    fx = load()
    nl, arrs, bits = inputs()
    assert netlist_digest(arrs) == fx["netlist_sha256"]
    assert digest(bits) == fx["vectors_sha256"]
    assert arrs["n"] == fx["n"] and len(arrs["fin"]) == fx["edges"]


@pytest.mark.parametrize("mode", ["robust", "nonrobust"])
@pytest.mark.parametrize("chunk", [1024, 256, 64])
def test_oracle_matches_fixture(mode, chunk):
    fx = load()["results"][mode]
    nl, _, bits = inputs()
    o = oracle_grade(nl, bits, TINY.pairs, chunk, mode == "robust")
    assert o["total"] == fx["total"]
    assert o["gate_counts"].tolist() == fx["gate_counts"]
    assert o["first_t0"].tolist() == fx["first_t0"] and o["first_t1"].tolist() == fx["first_t1"]
    if chunk == 256:
        assert o["chunk_totals"].tolist() == fx["chunk_totals_256"]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["robust", "nonrobust"])
def test_gpu_matches_fixture(mode):
    fx = load()["results"][mode]
    nl, _, bits = inputs()
    with pg.Grader(nl) as g:
        g.set_vectors(bits, TINY.pairs)
        g.set_chunk(256)
        total, per_chunk = g.grade_stages(mode == "robust")
        f0, f1 = g.first()
        assert total == fx["total"] and per_chunk == fx["chunk_totals_256"]
        assert g.gate_counts().tolist() == fx["gate_counts"]
        assert f0.tolist() == fx["first_t0"] and f1.tolist() == fx["first_t1"]
        g.set_chunk(1024)
        assert g.grade(mode == "robust") == fx["total"]
        assert g.gate_counts().tolist() == fx["gate_counts"]
