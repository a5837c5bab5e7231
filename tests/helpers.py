"""Shared helpers for the tests: small netlists and input conversions."""
from __future__ import annotations

import numpy as np

import pdfgrade as pg
from oracle import brute
from oracle import oracle as orc

C17_BENCH = """# ISCAS-85 c17
INPUT(1)
INPUT(2)
INPUT(3)
INPUT(6)
INPUT(7)
OUTPUT(22)
OUTPUT(23)
10 = NAND(1, 3)
11 = NAND(3, 6)
16 = NAND(2, 11)
19 = NAND(11, 7)
22 = NAND(10, 16)
23 = NAND(16, 19)
"""


def small_config(seed: int, num_pi=6, num_po=12, num_gates=24, levels=5, wide=True, alpha=0.0):
    fanin = (0.15, 0.45, 0.25, 0.15) if wide else (0.1, 0.9, 0.0, 0.0)
    return pg.GenConfig(seed=seed, num_pi=num_pi, num_po=num_po, num_gates=num_gates, levels=levels,
                        fanin_p=fanin, prev_level_frac=0.7, fanout_alpha=alpha, fanout_cap=16 if alpha else 0)


def biased_vectors(seed: int, num_pi: int, num_vectors: int, p_one: float = 0.5, p_hold: float = 0.0):
    """Vectors with a chosen density and a chance of repeating the previous
    vector's value per PI (more transitions held stable -> more marks)."""
    rng = np.random.default_rng(seed)
    vals = (rng.random((num_vectors, num_pi)) < p_one).astype(np.uint8)
    if p_hold:
        hold = rng.random((num_vectors, num_pi)) < p_hold
        for v in range(1, num_vectors):
            vals[v][hold[v]] = vals[v - 1][hold[v]]
    return pack(vals)


def pack(vals: np.ndarray) -> np.ndarray:
    """(num_vectors, num_pi) 0/1 -> PI-major packed (num_pi, words) uint64."""
    V, I = vals.shape
    words = (V + 63) // 64
    padded = np.zeros((words * 64, I), np.uint8)
    padded[:V] = vals
    bits = np.packbits(padded.T.reshape(I, words, 64), axis=2, bitorder="little")
    return np.ascontiguousarray(bits.reshape(I, words * 8)).view(np.uint64).reshape(I, words).copy()


def unpack(bits: np.ndarray, num_vectors: int) -> np.ndarray:
    I, words = bits.shape
    u = np.unpackbits(np.ascontiguousarray(bits).view(np.uint8).reshape(I, words * 8), axis=1, bitorder="little")
    return u[:, :num_vectors].T.copy()


def netlist_lists(arrs: dict):
    n = arrs["n"]
    fo, fi = arrs["fin_off"], arrs["fin"]
    return ([int(t) for t in arrs["type"]], [[int(x) for x in fi[fo[g]:fo[g + 1]]] for g in range(n)],
            [int(x) for x in arrs["po"]])


def brute_grade(nl: pg.Netlist, a_bits, num_pairs, robust, b_bits=None):
    arrs = nl.arrays()
    types, fanins, po = netlist_lists(arrs)
    va = unpack(a_bits, num_pairs)
    vb = unpack(b_bits, num_pairs) if b_bits is not None else np.roll(va, -1, axis=0)
    return brute.grade(types, fanins, po, arrs["num_pi"], va.tolist(), vb.tolist(), robust)


def oracle_grade(nl: pg.Netlist, a_bits, num_pairs, chunk, robust, b_bits=None, merge_window=0):
    return orc.grade(orc.OracleNet(nl.arrays()), a_bits, num_pairs, chunk, robust, b_bits=b_bits,
                     merge_window=merge_window)


def host_threads() -> int:
    """Threads for the bit-sliced CPU grader: the cores this process may use,
    capped by OMP_NUM_THREADS when that is set."""
    import os
    cores = len(os.sched_getaffinity(0))
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(cores, env) if env > 0 else cores)


def bitsliced_grade(nl: pg.Netlist, a_bits, num_pairs, chunk, robust, b_bits=None):
    return orc.grade_bitsliced(orc.OracleNet(nl.arrays()), a_bits, num_pairs, chunk, robust, b_bits=b_bits,
                               threads=host_threads())


def edge_netlist():
    """Every gate type with 1 and >= 2 fanins, a 12-input AND/OR (more fanins
    than the mark pass keeps in registers), POs with fanouts, an unused PI."""
    I = 8
    types = [pg.INPT] * I
    fanins = [[] for _ in range(I)]
    po = [0] * I

    def add(t, ins, out=0):
        types.append(t)
        fanins.append(ins)
        po.append(out)
        return len(types) - 1

    l1 = [add(t, ins) for t, ins in [
        (pg.AND, [0, 1]), (pg.NAND, [1, 2, 3]), (pg.OR, [2, 3]), (pg.NOR, [3, 4, 5]), (pg.XOR, [4, 5]),
        (pg.XNOR, [5]), (pg.NOT, [6]), (pg.BUFF, [0, 1]), (pg.FROM, [2]), (pg.DFF, [3]), (pg.NAND, [4]),
        (pg.UNKN, [5, 6]), (pg.NOT, [1, 2]), (pg.XOR, [6]), (pg.UNKN, [0]), (pg.OR, [0, 1, 2, 3, 4])]]
    big_and = add(pg.AND, [0, 1, 2, 3, 4, 5, 6, l1[0], l1[2], l1[6], l1[8], l1[9]], out=1)
    big_or = add(pg.NOR, [l1[1], l1[3], 0, 1, 2, 3, 4, 5, 6, l1[13], l1[14], l1[7]], out=1)
    l2 = [add(t, ins, out) for t, ins, out in [
        (pg.NAND, [l1[0], l1[1]], 1), (pg.OR, [l1[2], l1[3], l1[5]], 1), (pg.AND, [l1[6], l1[7]], 0),
        (pg.BUFF, [l1[8]], 1), (pg.NOR, [l1[9], l1[10], l1[11]], 0), (pg.NOT, [l1[12]], 1),
        (pg.XNOR, [l1[13], l1[14]], 1), (pg.AND, [l1[4], l1[6]], 1)]]
    add(pg.OR, [l2[2], l2[4], big_and], out=1)
    add(pg.NAND, [l2[0], l2[2]], out=1)
    return pg.Netlist.from_arrays(I, types, po, fanins)
