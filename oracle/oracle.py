"""ctypes wrapper of the CPU oracle (oracle.c).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker of the HIP path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
BSL_PATH = os.path.join(HERE, "libbitsliced.so")

FLAG_EXCEEDS_U16 = 1
FLAG_EXCEEDS_U32 = 2
FLAG_U64_OVERFLOW = 4
FLAG_EXCEEDS_U16_DP = 8
REFERENCE_MERGE_WINDOW = 4224  # patterns examined per chunk by mergekernel.cuh:125-130


class _Net(C.Structure):
    _fields_ = [("n", C.c_uint32), ("num_pi", C.c_uint32), ("L", C.c_uint32)] + [
        (k, C.c_void_p) for k in ("type", "po", "fin_off", "fin", "fout_off", "fout", "level_off")]


class _Opt(C.Structure):
    _fields_ = [("robust", C.c_int), ("merge_window", C.c_uint32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run oracle/build.py (or __graft_entry__.build())")
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_table.restype = C.POINTER(C.c_uint8)
        _lib.orc_table.argtypes = [C.c_int]
        _lib.orc_mark_lut.restype = C.c_uint16
        _lib.orc_mark_lut.argtypes = [C.c_int, C.c_int]
    return _lib


_bsl = None


def bsl_lib():
    global _bsl
    if _bsl is None:
        if not os.path.exists(BSL_PATH):
            raise ImportError(f"{BSL_PATH} is not built; run oracle/build.py (or __graft_entry__.build())")
        _bsl = C.CDLL(BSL_PATH)
    return _bsl


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class OracleNet:
    """Keeps the numpy CSR arrays alive behind an orc_netlist."""

    def __init__(self, arrs: dict):
        self.a = {k: np.ascontiguousarray(arrs[k]) for k in
                  ("type", "po", "fin_off", "fin", "fout_off", "fout", "level_off")}
        for k in ("fin", "fout"):
            if self.a[k].size == 0:
                self.a[k] = np.zeros(1, np.uint32)
        self.n = int(arrs["n"])
        self.s = _Net(self.n, int(arrs["num_pi"]), int(arrs["L"]), *[_p(self.a[k]).value for k in
                                                                    ("type", "po", "fin_off", "fin", "fout_off",
                                                                     "fout", "level_off")])

    @property
    def ref(self):
        return C.byref(self.s)


def table(which: int, size: int) -> np.ndarray:
    """0..5 = and2, nand2, or2, nor2, xor2, xnor2 (16 entries); 6 = stable (4); 7 = not (4)."""
    ptr = lib().orc_table(which)
    return np.ctypeslib.as_array(ptr, shape=(size,)).copy()


def mark_lut(robust: bool, idx: int) -> int:
    return int(lib().orc_mark_lut(int(robust), idx))


def grade(net: OracleNet, a_bits: np.ndarray, num_pairs: int, chunk: int, robust: bool,
          b_bits: np.ndarray | None = None, merge_window: int = 0) -> dict:
    a_bits = np.ascontiguousarray(a_bits, dtype=np.uint64)
    if b_bits is not None:
        b_bits = np.ascontiguousarray(b_bits, dtype=np.uint64)
    chunk = num_pairs if chunk <= 0 or chunk > num_pairs else chunk
    nchunks = (num_pairs + chunk - 1) // chunk
    f0 = np.full(net.n, 0xFFFFFFFF, np.uint32)
    f1 = np.full(net.n, 0xFFFFFFFF, np.uint32)
    gc = np.zeros(net.n, np.uint64)
    ct = np.zeros(nchunks, np.uint64)
    tot = C.c_uint64()
    flags = C.c_uint32()
    opt = _Opt(int(robust), merge_window)
    rc = lib().orc_grade(net.ref, _p(a_bits), _p(b_bits), C.c_uint64(num_pairs), C.c_uint64(chunk), C.byref(opt),
                         _p(f0), _p(f1), _p(gc), _p(ct), C.byref(tot), C.byref(flags))
    if rc:
        raise RuntimeError(f"orc_grade failed ({rc})")
    return dict(first_t0=f0, first_t1=f1, gate_counts=gc, chunk_totals=ct, total=tot.value, flags=flags.value)


def segments(net: OracleNet, K: int) -> np.ndarray:
    """Segments of length K (orc_segments): (S, K) uint32, 0xFFFFFFFF padded."""
    cnt = C.c_uint64()
    if lib().orc_segments(net.ref, C.c_uint32(K), None, C.c_uint64(0), C.byref(cnt)):
        raise ValueError(f"bad segment length {K}")
    out = np.zeros((max(cnt.value, 1), K), np.uint32)
    lib().orc_segments(net.ref, C.c_uint32(K), _p(out), C.c_uint64(cnt.value), C.byref(cnt))
    return out[:cnt.value]


def grade_segs(net: OracleNet, a_bits: np.ndarray, num_pairs: int, chunk: int, robust: bool, K: int,
               b_bits: np.ndarray | None = None, merge_window: int = 0) -> dict:
    """The chunk loop with segments of length K (orc_grade_segs)."""
    a_bits = np.ascontiguousarray(a_bits, dtype=np.uint64)
    if b_bits is not None:
        b_bits = np.ascontiguousarray(b_bits, dtype=np.uint64)
    segs = np.ascontiguousarray(segments(net, K))
    S = segs.shape[0]
    chunk = num_pairs if chunk <= 0 or chunk > num_pairs else chunk
    nchunks = (num_pairs + chunk - 1) // chunk
    s0 = np.full(max(S, 1), 0xFFFFFFFF, np.uint32)
    s1 = np.full(max(S, 1), 0xFFFFFFFF, np.uint32)
    gc = np.zeros(net.n, np.uint64)
    ct = np.zeros(nchunks, np.uint64)
    tot = C.c_uint64()
    flags = C.c_uint32()
    opt = _Opt(int(robust), merge_window)
    rc = lib().orc_grade_segs(net.ref, _p(a_bits), _p(b_bits), C.c_uint64(num_pairs), C.c_uint64(chunk),
                              C.byref(opt), C.c_uint32(K), _p(segs), C.c_uint64(S), _p(s0), _p(s1), _p(gc), _p(ct),
                              C.byref(tot), C.byref(flags))
    if rc:
        raise RuntimeError(f"orc_grade_segs failed ({rc})")
    return dict(segments=segs, seg_t0=s0[:S], seg_t1=s1[:S], gate_counts=gc, chunk_totals=ct, total=tot.value,
                flags=flags.value)


def sim_chunk(net: OracleNet, a_bits, num_pairs: int, start: int, width: int, b_bits=None) -> np.ndarray:
    a_bits = np.ascontiguousarray(a_bits, dtype=np.uint64)
    if b_bits is not None:
        b_bits = np.ascontiguousarray(b_bits, dtype=np.uint64)
    codes = np.zeros((net.n, width), np.uint8)
    lib().orc_sim_chunk(net.ref, _p(a_bits), _p(b_bits), C.c_uint64(num_pairs), C.c_uint64(start),
                        C.c_uint64(width), _p(codes))
    return codes


def mark_chunk(net: OracleNet, codes: np.ndarray, robust: bool):
    """Returns (final marks, forward marks) as (n, width) uint8."""
    codes = np.ascontiguousarray(codes)
    width = codes.shape[1]
    marks = np.zeros_like(codes)
    fwd = np.zeros_like(codes)
    lib().orc_mark_chunk(net.ref, _p(codes), C.c_uint64(width), int(robust), _p(marks), _p(fwd))
    return marks, fwd


def merge_chunk(net: OracleNet, codes, marks, start: int, first_t0, first_t1, merge_window: int = 0):
    codes = np.ascontiguousarray(codes)
    marks = np.ascontiguousarray(marks)
    lib().orc_merge_chunk(net.ref, _p(codes), _p(marks), C.c_uint64(codes.shape[1]), C.c_uint64(start),
                          C.c_uint32(merge_window), _p(first_t0), _p(first_t1))


def cover_chunk(net: OracleNet, marks, start: int, first_t0, first_t1, gate_counts):
    marks = np.ascontiguousarray(marks)
    tot = C.c_uint64()
    flags = C.c_uint32()
    lib().orc_cover_chunk(net.ref, _p(marks), C.c_uint64(marks.shape[1]), C.c_uint64(start), _p(first_t0),
                          _p(first_t1), _p(gate_counts), C.byref(tot), C.byref(flags))
    return tot.value, flags.value


def grade_bitsliced(net: OracleNet, a_bits: np.ndarray, num_pairs: int, chunk: int, robust: bool,
                    b_bits: np.ndarray | None = None, threads: int = 1) -> dict:
    """The bit-sliced grader of bitsliced.c: same results as grade() with the
    whole-chunk merge (chunk a multiple of 64, or the whole run); `flags`
    carries only FLAG_U64_OVERFLOW."""
    a_bits = np.ascontiguousarray(a_bits, dtype=np.uint64)
    if b_bits is not None:
        b_bits = np.ascontiguousarray(b_bits, dtype=np.uint64)
    f0 = np.full(net.n, 0xFFFFFFFF, np.uint32)
    f1 = np.full(net.n, 0xFFFFFFFF, np.uint32)
    gc = np.zeros(net.n, np.uint64)
    tot = C.c_uint64()
    flags = C.c_uint32()
    rc = bsl_lib().bsl_grade(net.ref, _p(a_bits), _p(b_bits), C.c_uint64(num_pairs), C.c_uint64(chunk),
                             int(robust), int(threads), _p(f0), _p(f1), _p(gc), C.byref(tot), C.byref(flags))
    if rc:
        raise RuntimeError(f"bsl_grade failed ({rc})")
    return dict(first_t0=f0, first_t1=f1, gate_counts=gc, total=tot.value, flags=flags.value)


def cover_new_chunk(net: OracleNet, marks, newm, gate_counts, num_pi: int):
    """orc_cover_new_chunk: (chunk total, flags, c of every PI [num_pi][W])."""
    marks = np.ascontiguousarray(marks, dtype=np.uint8)
    newm = np.ascontiguousarray(newm, dtype=np.uint8)
    W = marks.shape[1]
    pic = np.zeros((num_pi, W), np.uint64)
    tot = C.c_uint64()
    flags = C.c_uint32()
    lib().orc_cover_new_chunk(net.ref, _p(marks), _p(newm), C.c_uint64(W), _p(gate_counts), C.byref(tot),
                              C.byref(flags), _p(pic))
    return tot.value, flags.value, pic


def new_from_firsts(n: int, W: int, start: int, first_t0, first_t1) -> np.ndarray:
    """K = 1 new-node matrix (node g new at pattern start+p when its first T0
    or T1 is start+p), [n][W] bytes."""
    newm = np.zeros((n, W), np.uint8)
    for arr in (first_t0, first_t1):
        arr = np.asarray(arr, dtype=np.int64)
        sel = (arr != 0xFFFFFFFF) & (arr >= start) & (arr < start + W)
        newm[np.nonzero(sel)[0], arr[sel] - start] = 1
    return newm


def ref_unmerged_segments(S: int) -> int:
    """orc_ref_unmerged_segments: segments the reference's merge launch loop
    never examines (mergekernel.cuh:229-250)."""
    f = lib().orc_ref_unmerged_segments
    f.restype = C.c_uint64
    f.argtypes = [C.c_uint64]
    return int(f(C.c_uint64(S)))
