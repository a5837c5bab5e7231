"""Seeded mutation fuzzing of the host parsers (.bench, .level, vector text):
every mutated file either parses into a consistent netlist / vector set or
is refused with a PdfError; nothing crashes."""
from __future__ import annotations

import random

import numpy as np
import pytest

import pdfgrade as pg
from helpers import C17_BENCH
from pdfgrade.configs import TINY

CHARS = "abcXYZ019()=,# \n\t_-.INPUTOUTPANDNOTXORBUFFDFFFROM"


def mutate(rng: random.Random, text: str) -> str:
    t = list(text)
    for _ in range(rng.randint(1, 12)):
        op = rng.random()
        if op < 0.3 and t:
            del t[rng.randrange(len(t))]
        elif op < 0.6:
            t.insert(rng.randrange(len(t) + 1), rng.choice(CHARS))
        elif op < 0.8 and t:
            t[rng.randrange(len(t))] = rng.choice(CHARS)
        else:
            lines = "".join(t).split("\n")
            if len(lines) > 1:
                i, j = rng.randrange(len(lines)), rng.randrange(len(lines))
                if rng.random() < 0.5:
                    lines[i], lines[j] = lines[j], lines[i]
                else:
                    lines.insert(i, lines[j])
            t = list("\n".join(lines))
    return "".join(t)


def check_netlist(nl: pg.Netlist):
    a = nl.arrays()
    n = a["n"]
    fo, fi = a["fin_off"], a["fin"]
    assert fo[0] == 0 and np.all(np.diff(fo.astype(np.int64)) >= 0)
    lvl = a["level"]
    for g in range(n):
        ins = fi[fo[g]:fo[g + 1]]
        assert np.all(ins < g)                       # fanins precede the node
        if a["type"][g] != pg.INPT:
            assert len(ins) > 0 and lvl[g] == 1 + max(lvl[ins])
    # fanouts mirror fanins
    assert int(a["fout_off"][-1]) == int(fo[-1])


@pytest.mark.parametrize("seed", range(3))
def test_bench_and_level_parsers_survive_mutations(tmp_path, seed):
    rng = random.Random(seed)
    level_path = tmp_path / "tiny.level"
    pg.Netlist.generate(TINY.netlist).write_level(str(level_path))
    level_text = level_path.read_text()
    parsed = 0
    for it in range(150):
        p = tmp_path / ("x.bench" if it % 2 == 0 else "x.level")
        p.write_text(mutate(rng, C17_BENCH if it % 2 == 0 else level_text))
        try:
            nl = pg.Netlist.read_bench(str(p)) if it % 2 == 0 else pg.Netlist.read_level(str(p))
        except pg.PdfError:
            continue
        check_netlist(nl)
        parsed += 1
    assert parsed > 0


@pytest.mark.parametrize("seed", range(3))
def test_vector_reader_survives_random_text(tmp_path, seed):
    rng = random.Random(100 + seed)
    p = tmp_path / "v.vec"
    for _ in range(150):
        num_pi = rng.randint(1, 70)
        p.write_text("".join(rng.choice("01 \n#x\t") for _ in range(rng.randint(0, 400))))
        try:
            bits, count = pg.read_vectors(str(p), num_pi)
        except pg.PdfError:
            continue
        assert bits.shape == (num_pi, max(1, (count + 63) // 64))
        if count % 64:
            assert not np.any(bits[:, -1] >> np.uint64(count % 64))  # nothing past the last vector
