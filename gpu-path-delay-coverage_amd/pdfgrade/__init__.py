"""Python host side of the MI355X path-delay fault grader (ctypes over the C ABI
declared in include/pdfgrade.h).

The grader mirrors the reference's four host wrappers and its chunk loop:

    Grader.sim(chunk, start)            gpuRunSimulation   simkernel.cuh:200-249
    Grader.mark(chunk, start, robust)   gpuMarkPaths       markkernel.cuh:295-355
    Grader.merge(chunk, start, segs)    gpuMergeSegments   mergekernel.cuh:200-267
    Grader.cover(chunk, start, segs)    gpuCountPaths      coverkernel.cuh:233-333
    Grader.grade(robust)                chunk loop         main.cu:171-231

Every call runs the HIP kernels in libpdfgrade.so; there is no CPU fallback.
Loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import sys
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpdfgrade.so")

# node types / codes (include/pdfgrade.h)
UNKN, INPT, AND, NAND, OR, NOR, XOR, XNOR, BUFF, NOT, FROM, DFF = range(12)
TYPE_NAMES = ["UNKN", "INPT", "AND", "NAND", "OR", "NOR", "XOR", "XNOR", "BUFF", "NOT", "FROM", "DFF"]
S0, S1, T0, T1 = 0, 1, 2, 3
NONE = 0xFFFFFFFF
TUNE_WORDS_PER_LANE, TUNE_USE_GRAPH, TUNE_COVER_WAVES, TUNE_SEGMENT_ORDER, TUNE_LANES, TUNE_REV_COMPACT = 1, 2, 3, 4, 5, 6
TUNE_REV_TILES = 7
TUNE_FWD_ORDER = 12
TUNE_WALK = 8
TUNE_COVER_NODES, TUNE_COVER_POOL, TUNE_REV_WALK = 9, 10, 11
REF_EXCEEDS_U16, REF_EXCEEDS_U32, REF_SEGMENT_STRIDE, REF_MERGE_WINDOW = 1, 2, 4, 8
REF_MERGE_PATTERNS = 4224  # the reference's merge window (mergekernel.cuh:125-130)
DEBUG_FORWARD_MARKS, DEBUG_PI_COVER = 1, 2

OK, ERR_ARG, ERR_NETLIST, ERR_IO, ERR_HIP, ERR_NOMEM, ERR_UNSUPPORTED, ERR_STATE, ERR_OVERFLOW = (
    0, -1, -2, -3, -4, -5, -6, -7, -8)


class PdfError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class _GenParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("num_pi", C.c_uint32), ("num_po", C.c_uint32),
                ("num_gates", C.c_uint32), ("levels", C.c_uint32), ("fanin_p", C.c_double * 4),
                ("prev_level_frac", C.c_double), ("fanout_alpha", C.c_double),
                ("fanout_cap", C.c_uint32), ("max_attempts", C.c_uint32), ("path_limit_log2", C.c_uint32)]


class _Desc(C.Structure):
    _fields_ = [("n", C.c_uint32), ("num_pi", C.c_uint32), ("type", C.c_void_p), ("po", C.c_void_p),
                ("fin_off", C.c_void_p), ("fin", C.c_void_p)]


class _Info(C.Structure):
    _fields_ = [("n", C.c_uint32), ("num_pi", C.c_uint32), ("num_po", C.c_uint32), ("levels", C.c_uint32),
                ("edges", C.c_uint64), ("structural_paths", C.c_uint64)]


class _GraderTuning(C.Structure):
    _fields_ = [("lanes", C.c_uint32), ("chunk_pairs", C.c_uint64), ("use_graph", C.c_uint32),
                ("words_per_lane", C.c_uint32), ("walk", C.c_uint32), ("ms", C.c_float)]


class KernelTimes(C.Structure):
    _fields_ = [("fwd_ms", C.c_double), ("reverse_ms", C.c_double), ("events_ms", C.c_double),
                ("cover_ms", C.c_double), ("fill_ms", C.c_double), ("fwd_launches", C.c_uint32),
                ("reverse_launches", C.c_uint32), ("events_launches", C.c_uint32),
                ("cover_launches", C.c_uint32), ("fill_launches", C.c_uint32), ("wall_ms", C.c_double),
                ("merge_ms", C.c_double), ("merge_launches", C.c_uint32)]


# exported symbols of include/pdfgrade.h with their argument types
_P = C.c_void_p
_U64 = C.c_uint64
_U64P = C.POINTER(C.c_uint64)
_FP = C.POINTER(C.c_float)
SIGNATURES = {
    "pdf_last_error": ([], C.c_char_p),
    "pdf_abi_version": ([], C.c_int),
    "pdf_netlist_create": ([C.POINTER(_Desc), C.POINTER(_P)], C.c_int),
    "pdf_netlist_read_bench": ([C.c_char_p, C.POINTER(_P)], C.c_int),
    "pdf_netlist_read_level": ([C.c_char_p, C.POINTER(_P)], C.c_int),
    "pdf_netlist_write_level": ([_P, C.c_char_p], C.c_int),
    "pdf_netlist_free": ([_P], None),
    "pdf_netlist_get_info": ([_P, C.POINTER(_Info)], C.c_int),
    "pdf_netlist_export": ([_P] + [C.c_void_p] * 8, C.c_int),
    "pdf_netlist_segments": ([_P, C.c_uint32, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)], C.c_int),
    "pdf_netlist_generate": ([C.POINTER(_GenParams), C.POINTER(_P)], C.c_int),
    "pdf_gen_vectors": ([_U64, C.c_uint32, _U64, C.c_void_p], C.c_int),
    "pdf_vectors_count": ([C.c_char_p, C.c_uint32, _U64P], C.c_int),
    "pdf_vectors_read": ([C.c_char_p, C.c_uint32, C.c_void_p, _U64, _U64P], C.c_int),
    "pdf_vectors_write": ([C.c_char_p, C.c_uint32, C.c_void_p, _U64], C.c_int),
    "pdf_gen_toggle_pairs": ([_U64, C.c_uint32, _U64, C.c_double, C.c_double, C.c_void_p, C.c_void_p], C.c_int),
    "pdf_grader_create": ([C.c_int, _P, C.POINTER(_P)], C.c_int),
    "pdf_grader_destroy": ([_P], None),
    "pdf_grader_set_vectors": ([_P, C.c_void_p, _U64], C.c_int),
    "pdf_grader_set_pairs": ([_P, C.c_void_p, C.c_void_p, _U64], C.c_int),
    "pdf_grader_set_chunk": ([_P, _U64], C.c_int),
    "pdf_grader_num_chunks": ([_P, _U64P], C.c_int),
    "pdf_grader_chunk_width": ([_P, _U64, _U64P], C.c_int),
    "pdf_grader_reset": ([_P], C.c_int),
    "pdf_sim": ([_P, _U64, _U64, _FP], C.c_int),
    "pdf_mark": ([_P, _U64, _U64, C.c_int, _FP], C.c_int),
    "pdf_merge": ([_P, _U64, _U64, C.c_int, _FP], C.c_int),
    "pdf_cover": ([_P, _U64, _U64, C.c_int, _U64P, _FP], C.c_int),
    "pdf_grade": ([_P, C.c_int, _U64P], C.c_int),
    "pdf_grade_timed": ([_P, C.c_int, _U64P, C.POINTER(KernelTimes)], C.c_int),
    "pdf_grade_segs": ([_P, C.c_int, C.c_int, _U64P, C.POINTER(KernelTimes)], C.c_int),
    "pdf_grade_async": ([_P, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "pdf_grader_reset_async": ([_P, C.c_void_p], C.c_int),
    "pdf_sim_async": ([_P, C.c_uint64, C.c_uint64, C.c_void_p], C.c_int),
    "pdf_mark_async": ([_P, C.c_uint64, C.c_uint64, C.c_int, C.c_void_p], C.c_int),
    "pdf_merge_async": ([_P, C.c_uint64, C.c_uint64, C.c_int, C.c_void_p], C.c_int),
    "pdf_cover_async": ([_P, C.c_uint64, C.c_uint64, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
                        C.c_int),
    "pdf_grade_noalloc": ([_P, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "pdf_grader_set_tuning": ([_P, C.c_int, C.c_int64], C.c_int),
    "pdf_grader_set_merge_window": ([_P, C.c_uint32], C.c_int),
    "pdf_grader_ref_flags": ([_P, C.POINTER(C.c_uint32), _U64P], C.c_int),
    "pdf_ref_unmerged_segments": ([_U64], _U64),
    "pdf_get_sim_planes": ([_P, C.c_void_p, C.c_void_p], C.c_int),
    "pdf_get_mark_plane": ([_P, C.c_void_p], C.c_int),
    "pdf_get_first": ([_P, C.c_void_p, C.c_void_p], C.c_int),
    "pdf_get_segment_first": ([_P, C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)], C.c_int),
    "pdf_grader_set_debug": ([_P, C.c_uint32], C.c_int),
    "pdf_get_forward_mark_plane": ([_P, C.c_void_p], C.c_int),
    "pdf_get_pi_cover": ([_P, C.c_void_p], C.c_int),
    "pdf_get_gate_counts": ([_P, C.c_void_p], C.c_int),
    "pdf_get_total": ([_P, _U64P], C.c_int),
    "pdf_pair_features": ([_P, C.c_uint32, C.c_int, C.c_void_p, _U64], C.c_int),
    "pdf_score_pairs": ([_P, _P, C.c_uint32, C.c_int, C.c_void_p, _U64, C.c_void_p], C.c_int),
    "pdf_pair_gains": ([_P, C.c_int, C.c_void_p], C.c_int),
    "pdf_grader_rank_pairs": ([_P, C.c_void_p, C.c_void_p], C.c_int),
    "pdf_grade_curve": ([_P, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "pdf_mlp_create": ([C.c_int, _U64, C.POINTER(_P)], C.c_int),
    "pdf_mlp_destroy": ([_P], None),
    "pdf_mlp_set_params": ([_P, C.c_void_p], C.c_int),
    "pdf_mlp_get_params": ([_P, C.c_void_p], C.c_int),
    "pdf_mlp_forward": ([_P, C.c_void_p, _U64, _U64, C.c_void_p], C.c_int),
    "pdf_mlp_gradients": ([_P, C.c_void_p, _U64, C.c_void_p, _U64, _FP, C.c_void_p], C.c_int),
    "pdf_mlp_adam": ([_P, C.c_float, C.c_float, C.c_float, C.c_float], C.c_int),
    "pdf_mlp_train_step": ([_P, C.c_void_p, _U64, C.c_void_p, _U64, C.c_float, _FP], C.c_int),
    "pdf_mlp_sync": ([_P], C.c_int),
    "pdf_grader_autotune": ([_P, C.c_int, C.c_uint32, C.c_void_p], C.c_int),
    "pdf_batch_create": ([C.c_int, C.c_uint32, C.POINTER(_P), C.POINTER(_P), C.POINTER(C.c_uint64), C.POINTER(_P)],
                         C.c_int),
    "pdf_batch_destroy": ([_P], None),
    "pdf_batch_grade": ([_P, C.c_int, C.c_void_p, _FP], C.c_int),
    "pdf_batch_member": ([_P, C.c_uint32, C.POINTER(_P)], C.c_int),
    "pdf_batch_set_tuning": ([_P, C.c_int, C.c_int64], C.c_int),
    "pdf_batch_autotune": ([_P, C.c_int, C.c_uint32, C.c_void_p], C.c_int),
}

_lib = None


def lib():
    """Load libpdfgrade.so (built by __graft_entry__.build()); raise if absent.

    torch ships its own HIP runtime under the same soname as the one the
    library links, and whichever is loaded first serves the whole process;
    torch's device layer only works on its own.  A process that passes torch
    tensors to the library (the learned ranking) must import torch before the
    library is loaded; otherwise the library runs on the system ROCm runtime
    (the faster one for grading)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        h = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(h, name)
            f.argtypes = args
            f.restype = res
        _lib = h
    return _lib


def _check(rc: int):
    if rc != OK:
        raise PdfError(rc, lib().pdf_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


@dataclass(frozen=True)
class GenConfig:
    """Synthetic netlist + vector recipe (see configs.py)."""
    seed: int
    num_pi: int
    num_po: int
    num_gates: int
    levels: int
    fanin_p: tuple
    prev_level_frac: float = 0.7
    fanout_alpha: float = 0.0
    fanout_cap: int = 0
    max_attempts: int = 16
    path_limit_log2: int = 63  # reject netlists with >= 2^this PI->PO paths (0 = keep any)


class Netlist:
    """Levelized netlist held by the C library (host memory)."""

    def __init__(self, handle):
        self._h = handle

    def __del__(self, _finalizing=sys.is_finalizing):
        # at interpreter exit the module's globals are gone; the process frees everything
        h = getattr(self, "_h", None)
        if h and not _finalizing():
            lib().pdf_netlist_free(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    @classmethod
    def generate(cls, cfg: GenConfig) -> "Netlist":
        p = _GenParams()
        p.seed, p.num_pi, p.num_po, p.num_gates, p.levels = cfg.seed, cfg.num_pi, cfg.num_po, cfg.num_gates, cfg.levels
        fp = list(cfg.fanin_p) + [0.0] * (4 - len(cfg.fanin_p))
        for i in range(4):
            p.fanin_p[i] = fp[i]
        p.prev_level_frac, p.fanout_alpha = cfg.prev_level_frac, cfg.fanout_alpha
        p.fanout_cap, p.max_attempts = cfg.fanout_cap, cfg.max_attempts
        p.path_limit_log2 = cfg.path_limit_log2
        h = C.c_void_p()
        _check(lib().pdf_netlist_generate(C.byref(p), C.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, num_pi: int, types, po, fanins) -> "Netlist":
        """types/po: per node; fanins: list of fanin-id lists (ordered)."""
        n = len(types)
        t = np.asarray(types, dtype=np.uint8)
        o = np.asarray(po, dtype=np.uint8)
        off = np.zeros(n + 1, dtype=np.uint32)
        off[1:] = np.cumsum([len(f) for f in fanins])
        fin = np.asarray([x for f in fanins for x in f] or [0], dtype=np.uint32)
        d = _Desc(n, num_pi, _ptr(t).value, _ptr(o).value, _ptr(off).value, _ptr(fin).value)
        h = C.c_void_p()
        _check(lib().pdf_netlist_create(C.byref(d), C.byref(h)))
        return cls(h)

    @classmethod
    def read_bench(cls, path: str) -> "Netlist":
        h = C.c_void_p()
        _check(lib().pdf_netlist_read_bench(path.encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def read_level(cls, path: str) -> "Netlist":
        h = C.c_void_p()
        _check(lib().pdf_netlist_read_level(path.encode(), C.byref(h)))
        return cls(h)

    def write_level(self, path: str):
        _check(lib().pdf_netlist_write_level(self._h, path.encode()))

    def info(self) -> dict:
        i = _Info()
        _check(lib().pdf_netlist_get_info(self._h, C.byref(i)))
        return dict(n=i.n, num_pi=i.num_pi, num_po=i.num_po, levels=i.levels, edges=i.edges,
                    structural_paths=i.structural_paths)

    def segments(self, segs: int) -> np.ndarray:
        """Segments of length segs (pdf_netlist_segments): (S, segs) uint32,
        NONE padded."""
        cnt = C.c_uint64()
        _check(lib().pdf_netlist_segments(self._h, segs, None, 0, C.byref(cnt)))
        out = np.zeros((max(cnt.value, 1), segs), np.uint32)
        _check(lib().pdf_netlist_segments(self._h, segs, _ptr(out), cnt.value, C.byref(cnt)))
        return out[:cnt.value]

    def arrays(self) -> dict:
        """CSR arrays as numpy: type, po, level, fin_off, fin, fout_off, fout, level_off."""
        inf = self.info()
        n, e, L = inf["n"], inf["edges"], inf["levels"]
        a = dict(type=np.zeros(n, np.uint8), po=np.zeros(n, np.uint8), level=np.zeros(n, np.uint32),
                 fin_off=np.zeros(n + 1, np.uint32), fin=np.zeros(max(e, 1), np.uint32),
                 fout_off=np.zeros(n + 1, np.uint32), fout=np.zeros(max(e, 1), np.uint32),
                 level_off=np.zeros(L + 2, np.uint32))
        _check(lib().pdf_netlist_export(self._h, *[_ptr(a[k]) for k in
                                                  ("type", "po", "level", "fin_off", "fin", "fout_off", "fout",
                                                   "level_off")]))
        a["fin"] = a["fin"][:e]
        a["fout"] = a["fout"][:e]
        a["num_pi"] = inf["num_pi"]
        a["n"] = n
        a["L"] = L
        return a


def hold_vectors(seed: int, num_pi: int, num_vectors: int, p_hold: float) -> np.ndarray:
    """Hold-biased vectors: vector 0 iid Bernoulli(1/2), then each PI keeps
    the previous vector's value with probability p_hold and draws a fresh one
    otherwise (fewer side inputs switch, so many more paths are sensitized
    than with iid vectors).  PI-major packed like gen_vectors (numpy
    Generator `seed`)."""
    rng = np.random.default_rng(seed)
    fresh = rng.random((num_vectors, num_pi)) < 0.5
    held = rng.random((num_vectors, num_pi)) < p_hold
    held[0] = False
    src = np.where(held, 0, np.arange(num_vectors)[:, None])
    np.maximum.accumulate(src, axis=0, out=src)  # the last vector that drew PI i fresh
    vals = np.take_along_axis(fresh, src, axis=0)
    words = (num_vectors + 63) // 64
    padded = np.zeros((words * 64, num_pi), dtype=bool)
    padded[:num_vectors] = vals
    packed = np.packbits(padded.T.reshape(num_pi, words, 64), axis=2, bitorder="little")
    return np.ascontiguousarray(packed.reshape(num_pi, words * 8)).view("<u8").astype(np.uint64)


def gen_vectors(seed: int, num_pi: int, num_vectors: int) -> np.ndarray:
    """iid Bernoulli(1/2) vectors, PI-major packed: shape (num_pi, ceil(V/64)) uint64."""
    words = (num_vectors + 63) // 64
    out = np.zeros((num_pi, words), dtype=np.uint64)
    _check(lib().pdf_gen_vectors(seed, num_pi, num_vectors, _ptr(out)))
    return out


def read_vectors(path: str, num_pi: int):
    """(PI-major packed bits, count) from a vector text file."""
    n = C.c_uint64()
    _check(lib().pdf_vectors_count(path.encode(), num_pi, C.byref(n)))
    bits = np.zeros((num_pi, max(1, (n.value + 63) // 64)), np.uint64)
    got = C.c_uint64()
    _check(lib().pdf_vectors_read(path.encode(), num_pi, _ptr(bits), n.value, C.byref(got)))
    return bits, got.value


def write_vectors(path: str, bits: np.ndarray, num_vectors: int):
    """Write PI-major packed vectors as a vector text file."""
    bits = np.ascontiguousarray(bits, dtype=np.uint64)
    _check(lib().pdf_vectors_write(path.encode(), bits.shape[0], _ptr(bits), num_vectors))


def gen_toggle_pairs(seed: int, num_pi: int, count: int, q_lo: float, q_hi: float):
    """Explicit pairs (A, B): A uniform, B = A with each PI flipped with a
    per-pair probability log-uniform in [q_lo, q_hi]."""
    words = (count + 63) // 64
    a = np.zeros((num_pi, words), np.uint64)
    b = np.zeros((num_pi, words), np.uint64)
    _check(lib().pdf_gen_toggle_pairs(seed, num_pi, count, q_lo, q_hi, _ptr(a), _ptr(b)))
    return a, b


class Grader:
    """One netlist resident in HBM plus the chunk bit-planes of one vector set."""

    def __init__(self, netlist: Netlist, device: int = 0):
        self._h = None
        self._owned = True
        h = C.c_void_p()
        _check(lib().pdf_grader_create(device, netlist.handle, C.byref(h)))
        self._h = h
        self.n = netlist.info()["n"]

    @classmethod
    def _view(cls, handle, n: int, num_pairs: int, keepalive) -> "Grader":
        """A non-owning view of a grader owned by a batch."""
        g = cls.__new__(cls)
        g._h, g._owned, g.n, g._num_pairs, g._keepalive = handle, False, n, num_pairs, keepalive
        return g

    def close(self):
        if self._h and self._owned:
            lib().pdf_grader_destroy(self._h)
        self._h = None

    def __del__(self, _finalizing=sys.is_finalizing):
        if not _finalizing():
            self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- inputs
    def set_vectors(self, bits: np.ndarray, num_vectors: int):
        bits = np.ascontiguousarray(bits, dtype=np.uint64)
        _check(lib().pdf_grader_set_vectors(self._h, _ptr(bits), num_vectors))
        self._num_pairs = num_vectors

    def set_pairs(self, a_bits: np.ndarray, b_bits: np.ndarray, num_pairs: int):
        a_bits = np.ascontiguousarray(a_bits, dtype=np.uint64)
        b_bits = np.ascontiguousarray(b_bits, dtype=np.uint64)
        _check(lib().pdf_grader_set_pairs(self._h, _ptr(a_bits), _ptr(b_bits), num_pairs))
        self._num_pairs = num_pairs

    def set_chunk(self, chunk_pairs: int):
        _check(lib().pdf_grader_set_chunk(self._h, chunk_pairs))

    def num_chunks(self) -> int:
        v = C.c_uint64()
        _check(lib().pdf_grader_num_chunks(self._h, C.byref(v)))
        return v.value

    def chunk_width(self, chunk: int) -> int:
        v = C.c_uint64()
        _check(lib().pdf_grader_chunk_width(self._h, chunk, C.byref(v)))
        return v.value

    def reset(self):
        _check(lib().pdf_grader_reset(self._h))

    def set_tuning(self, key: int, value: int):
        """TUNE_WORDS_PER_LANE (1/2), TUNE_USE_GRAPH (0/1), TUNE_LANES (1..4),
        TUNE_REV_TILES (0 automatic, 1/2/4/8), TUNE_REV_COMPACT (0/1), TUNE_WALK
        and TUNE_REV_WALK (0/1/2/4/8) or the cover caps; results never change."""
        _check(lib().pdf_grader_set_tuning(self._h, key, value))

    def autotune(self, robust: bool, reps: int = 5) -> dict:
        """Time the knob configurations, keep the fastest; returns it."""
        t = _GraderTuning()
        _check(lib().pdf_grader_autotune(self._h, int(robust), reps, C.byref(t)))
        return {k: getattr(t, k) for k, _ in _GraderTuning._fields_}

    # -- the four stages (reference host wrappers)
    def sim(self, chunk: int, start: int) -> float:
        ms = C.c_float()
        _check(lib().pdf_sim(self._h, chunk, start, C.byref(ms)))
        return ms.value

    def mark(self, chunk: int, start: int, robust: bool) -> float:
        ms = C.c_float()
        _check(lib().pdf_mark(self._h, chunk, start, int(robust), C.byref(ms)))
        return ms.value

    def merge(self, chunk: int, start: int, segs: int = 1) -> float:
        ms = C.c_float()
        _check(lib().pdf_merge(self._h, chunk, start, segs, C.byref(ms)))
        return ms.value

    def cover(self, chunk: int, start: int, segs: int = 1):
        ms = C.c_float()
        cov = C.c_uint64()
        _check(lib().pdf_cover(self._h, chunk, start, segs, C.byref(cov), C.byref(ms)))
        return cov.value, ms.value

    # -- the stages and the reset queued on a caller's HIP stream handle
    # (e.g. torch.cuda.Stream().cuda_stream); device pointers as ints
    def reset_async(self, stream: int):
        _check(lib().pdf_grader_reset_async(self._h, C.c_void_p(stream)))

    def sim_async(self, chunk: int, start: int, stream: int):
        _check(lib().pdf_sim_async(self._h, chunk, start, C.c_void_p(stream)))

    def mark_async(self, chunk: int, start: int, robust: bool, stream: int):
        _check(lib().pdf_mark_async(self._h, chunk, start, int(robust), C.c_void_p(stream)))

    def merge_async(self, chunk: int, start: int, segs: int, stream: int):
        _check(lib().pdf_merge_async(self._h, chunk, start, segs, C.c_void_p(stream)))

    def cover_async(self, chunk: int, start: int, segs: int, stream: int, chunk_total_dev: int,
                    per_gate_accum_dev: int = 0, err_dev: int = 0):
        _check(lib().pdf_cover_async(self._h, chunk, start, segs, C.c_void_p(stream), C.c_void_p(chunk_total_dev),
                                     C.c_void_p(per_gate_accum_dev) if per_gate_accum_dev else None,
                                     C.c_void_p(err_dev) if err_dev else None))

    def grade(self, robust: bool) -> int:
        tot = C.c_uint64()
        _check(lib().pdf_grade(self._h, int(robust), C.byref(tot)))
        return tot.value

    def grade_timed(self, robust: bool):
        """pdf_grade with every launch timed by HIP events: (total, times dict)."""
        tot = C.c_uint64()
        t = KernelTimes()
        _check(lib().pdf_grade_timed(self._h, int(robust), C.byref(tot), C.byref(t)))
        return tot.value, {k: getattr(t, k) for k, _ in KernelTimes._fields_}

    def grade_segs(self, robust: bool, segs: int, timed: bool = False):
        """pdf_grade_segs: the fused grade with segments of length segs;
        returns the total, or (total, times dict) when timed."""
        tot = C.c_uint64()
        t = KernelTimes() if timed else None
        _check(lib().pdf_grade_segs(self._h, int(robust), segs, C.byref(tot), C.byref(t) if timed else None))
        return (tot.value, {k: getattr(t, k) for k, _ in KernelTimes._fields_}) if timed else tot.value

    def grade_async(self, robust: bool, segs: int, stream: int, total_dev: int, err_dev: int = 0):
        """pdf_grade_async: queue the fused grade on a HIP stream handle
        (e.g. torch.cuda.Stream().cuda_stream); the total lands in the device
        u64 at total_dev once the stream has run it."""
        _check(lib().pdf_grade_async(self._h, int(robust), segs, C.c_void_p(stream), C.c_void_p(total_dev),
                                     C.c_void_p(err_dev) if err_dev else None))

    def grade_noalloc(self, robust: bool, stream: int, total_dev: int, err_dev: int = 0):
        """pdf_grade_noalloc: pdf_grade_async (segments of length 1) with every
        chunk buffer taken from the library's static per-GPU scratch array;
        nothing is allocated."""
        _check(lib().pdf_grade_noalloc(self._h, int(robust), C.c_void_p(stream), C.c_void_p(total_dev),
                                       C.c_void_p(err_dev) if err_dev else None))

    def set_merge_window(self, pairs: int):
        """Take first detections only from the first `pairs` pairs of each
        chunk (REF_MERGE_PATTERNS reproduces the reference; 0 = whole chunk)."""
        _check(lib().pdf_grader_set_merge_window(self._h, pairs))

    def ref_flags(self):
        """(REF_* flags, segments the reference's merge would leave
        unmerged) of the last grade."""
        f = C.c_uint32()
        un = C.c_uint64()
        _check(lib().pdf_grader_ref_flags(self._h, C.byref(f), C.byref(un)))
        return f.value, un.value

    def grade_stages(self, robust: bool, segs: int = 1):
        """The reference's chunk loop (main.cu:171-231) through the four stage
        calls; returns (total, per-chunk coverage list)."""
        self.reset()
        totals, per_chunk = 0, []
        start = 0
        for chunk in range(self.num_chunks()):
            self.sim(chunk, start)
            self.mark(chunk, start, robust)
            self.merge(chunk, start, segs)
            cov, _ = self.cover(chunk, start, segs)
            per_chunk.append(cov)
            totals += cov
            start += self.chunk_width(chunk)
        return totals, per_chunk

    # -- parity getters
    def sim_planes(self, width: int):
        words = (width + 63) // 64
        a = np.zeros((self.n, words), np.uint64)
        b = np.zeros((self.n, words), np.uint64)
        _check(lib().pdf_get_sim_planes(self._h, _ptr(a), _ptr(b)))
        return a, b

    def mark_plane(self, width: int):
        m = np.zeros((self.n, (width + 63) // 64), np.uint64)
        _check(lib().pdf_get_mark_plane(self._h, _ptr(m)))
        return m

    def first(self):
        f0 = np.zeros(self.n, np.uint32)
        f1 = np.zeros(self.n, np.uint32)
        _check(lib().pdf_get_first(self._h, _ptr(f0), _ptr(f1)))
        return f0, f1

    def set_debug(self, flags: int):
        """DEBUG_FORWARD_MARKS | DEBUG_PI_COVER (pdf_grader_set_debug)."""
        _check(lib().pdf_grader_set_debug(self._h, flags))

    def forward_mark_plane(self, width: int) -> np.ndarray:
        """Forward marks (before removal) of the chunk last marked: (n, words)."""
        out = np.zeros((self.n, (width + 63) // 64), np.uint64)
        _check(lib().pdf_get_forward_mark_plane(self._h, _ptr(out)))
        return out

    def pi_cover(self, num_pi: int, width: int) -> np.ndarray:
        """c of every PI per pair of the chunk last covered: (num_pi, width)."""
        out = np.zeros((num_pi, width), np.uint64)
        _check(lib().pdf_get_pi_cover(self._h, _ptr(out)))
        return out

    def segment_first(self):
        """First T0 / T1 pair of every segment of the last segs >= 2 merge."""
        cnt = C.c_uint64()
        _check(lib().pdf_get_segment_first(self._h, None, None, 0, C.byref(cnt)))
        f0 = np.zeros(max(cnt.value, 1), np.uint32)
        f1 = np.zeros(max(cnt.value, 1), np.uint32)
        _check(lib().pdf_get_segment_first(self._h, _ptr(f0), _ptr(f1), cnt.value, C.byref(cnt)))
        return f0[:cnt.value], f1[:cnt.value]

    def gate_counts(self) -> np.ndarray:
        c = np.zeros(self.n, np.uint64)
        _check(lib().pdf_get_gate_counts(self._h, _ptr(c)))
        return c

    def total(self) -> int:
        t = C.c_uint64()
        _check(lib().pdf_get_total(self._h, C.byref(t)))
        return t.value

    # -- learned pair ranking (device buffers: torch tensors on the grader's GPU)
    def pair_features(self, max_level: int, robust: bool, feat):
        """Write the 512 BF16 features of every loaded pair into `feat`, a
        feature-major int16/bfloat16 device tensor (512, ld) or a column
        slice of one (row stride ld a multiple of 64, >= num_pairs columns)."""
        ptr, ld, cols = _feat_arg(feat)
        if cols < self._num_pairs:
            raise ValueError("feature tensor has fewer columns than pairs")
        _check(lib().pdf_pair_features(self._h, max_level, int(robust), C.c_void_p(ptr), ld))

    def score_pairs(self, mlp: "Mlp", max_level: int, robust: bool, feat, scores):
        """pair_features and mlp.forward over every loaded pair, pipelined
        chunk group by chunk group (pdf_score_pairs): same features in
        `feat`, same scores in `scores` (float32 device tensor, num_pairs)."""
        ptr, ld, cols = _feat_arg(feat)
        if cols < self._num_pairs:
            raise ValueError("feature tensor has fewer columns than pairs")
        _need_device(scores, self._num_pairs, 4)
        _check(lib().pdf_score_pairs(self._h, mlp._h, max_level, int(robust), C.c_void_p(ptr), ld,
                                     C.c_void_p(scores.data_ptr())))

    def pair_gains(self, robust: bool, gains):
        """Each pair's marked PI->PO path count graded alone (int32 device
        tensor of num_pairs entries, read as uint32)."""
        _need_device(gains, self._num_pairs, 4)
        _check(lib().pdf_pair_gains(self._h, int(robust), C.c_void_p(gains.data_ptr())))

    def rank_pairs(self, scores, perm=None):
        """Reorder the loaded explicit pairs by descending float32 score;
        `perm` (optional int32 device tensor) receives the old index of each
        new position."""
        _need_device(scores, self._num_pairs, 4)
        if perm is not None:
            _need_device(perm, self._num_pairs, 4)
        _check(lib().pdf_grader_rank_pairs(self._h, C.c_void_p(scores.data_ptr()),
                                           C.c_void_p(perm.data_ptr() if perm is not None else 0)))

    def grade_curve(self, robust: bool, pair_totals=None):
        """Grade the loaded pairs in order; per chunk the running total and the
        milliseconds since the start: (uint64 array, float32 array).  With
        `pair_totals` (int32 device tensor of num_pairs entries) each pair's
        contribution to the total is written there too."""
        nch = self.num_chunks()
        tot = np.zeros(nch, np.uint64)
        ms = np.zeros(nch, np.float32)
        if pair_totals is not None:
            _need_device(pair_totals, self._num_pairs, 4)
        _check(lib().pdf_grade_curve(self._h, int(robust), _ptr(tot), _ptr(ms),
                                     C.c_void_p(pair_totals.data_ptr() if pair_totals is not None else 0)))
        return tot, ms


FEATURES = 512
MLP_PARAMS = 197377


class _BatchTuning(C.Structure):
    _fields_ = [("segment_order", C.c_uint32), ("cover_waves", C.c_uint32), ("use_graph", C.c_uint32),
                ("rev_compact", C.c_uint32), ("words_per_lane", C.c_uint32), ("ms", C.c_float)]


class Batch:
    """Many circuits (one chunk of sequence-mode vectors each) graded by one
    captured graph of level launches over the union of their levels."""

    def __init__(self, netlists, vectors, num_vectors, device: int = 0):
        self._h = None
        count = len(netlists)
        if not (count == len(vectors) == len(num_vectors)) or count == 0:
            raise ValueError("one vector set per netlist")
        self._bits = [np.ascontiguousarray(v, dtype=np.uint64) for v in vectors]
        nets = (_P * count)(*[nl.handle for nl in netlists])
        bits = (_P * count)(*[b.ctypes.data for b in self._bits])
        nv = (C.c_uint64 * count)(*[int(x) for x in num_vectors])
        h = C.c_void_p()
        _check(lib().pdf_batch_create(device, count, nets, bits, nv, C.byref(h)))
        self._h = h
        self.count = count
        self._n = [nl.info()["n"] for nl in netlists]
        self._pairs = [int(x) for x in num_vectors]

    def close(self):
        if self._h:
            lib().pdf_batch_destroy(self._h)
            self._h = None

    def __del__(self, _finalizing=sys.is_finalizing):
        if not _finalizing():
            self.close()

    def grade(self, robust: bool):
        """(per-circuit totals as uint64 array, GPU milliseconds)."""
        tot = np.zeros(self.count, np.uint64)
        ms = C.c_float()
        _check(lib().pdf_batch_grade(self._h, int(robust), _ptr(tot), C.byref(ms)))
        return tot, ms.value

    def member(self, c: int) -> Grader:
        """Circuit c's grader (first(), gate_counts(), total() of the last grade)."""
        h = C.c_void_p()
        _check(lib().pdf_batch_member(self._h, c, C.byref(h)))
        return Grader._view(h, self._n[c], self._pairs[c], self)

    def set_tuning(self, key: int, value: int):
        _check(lib().pdf_batch_set_tuning(self._h, key, value))

    def autotune(self, robust: bool, reps: int = 3) -> dict:
        t = _BatchTuning()
        _check(lib().pdf_batch_autotune(self._h, int(robust), reps, C.byref(t)))
        return {k: getattr(t, k) for k, _ in _BatchTuning._fields_}


class Mlp:
    """The BF16 MFMA pair scorer (512 -> 256 -> 256 -> 1) of libpdfgrade.so.
    Device buffers are torch tensors on the scorer's GPU; features are BF16
    bits in an int16 (or bfloat16) tensor of shape (rows, 512)."""

    def __init__(self, device: int = 0, seed: int = 1):
        self._h = None
        h = C.c_void_p()
        _check(lib().pdf_mlp_create(device, seed, C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            lib().pdf_mlp_destroy(self._h)
            self._h = None

    def __del__(self, _finalizing=sys.is_finalizing):
        if not _finalizing():
            self.close()

    def params(self) -> np.ndarray:
        out = np.zeros(MLP_PARAMS, np.float32)
        _check(lib().pdf_mlp_get_params(self._h, _ptr(out)))
        return out

    def set_params(self, flat: np.ndarray):
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        if flat.size != MLP_PARAMS:
            raise ValueError(f"expected {MLP_PARAMS} parameters")
        _check(lib().pdf_mlp_set_params(self._h, _ptr(flat)))

    def forward(self, feat, scores):
        """scores[r] = output for column r of `feat` (feature-major (512, rows)
        tensor or column slice)."""
        ptr, ld, rows = _feat_arg(feat)
        _need_device(scores, rows, 4)
        _check(lib().pdf_mlp_forward(self._h, C.c_void_p(ptr), ld, rows, C.c_void_p(scores.data_ptr())))

    def gradients(self, feat, target):
        """(loss, flat float32 gradient) of the mean squared error; no update."""
        ptr, ld, rows = _feat_arg(feat)
        _need_device(target, rows, 4)
        loss = C.c_float()
        grad = np.zeros(MLP_PARAMS, np.float32)
        _check(lib().pdf_mlp_gradients(self._h, C.c_void_p(ptr), ld, C.c_void_p(target.data_ptr()), rows,
                                       C.byref(loss), _ptr(grad)))
        return loss.value, grad

    def adam(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
        _check(lib().pdf_mlp_adam(self._h, lr, beta1, beta2, eps))

    def train_step(self, feat, target, lr: float, want_loss: bool = False):
        """Gradient + Adam on one batch (a column slice of the feature matrix).
        Without want_loss the call returns before the GPU finishes: keep the
        tensors alive and call sync() before reading or changing them."""
        ptr, ld, rows = _feat_arg(feat)
        _need_device(target, rows, 4)
        loss = C.c_float()
        _check(lib().pdf_mlp_train_step(self._h, C.c_void_p(ptr), ld, C.c_void_p(target.data_ptr()), rows,
                                        lr, C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def sync(self):
        _check(lib().pdf_mlp_sync(self._h))


def _feat_arg(feat):
    """(pointer, ld, columns) of a feature-major (512, columns) device tensor
    whose rows are contiguous with a stride that is a multiple of 64."""
    if not getattr(feat, "is_cuda", False) or feat.dim() != 2 or feat.shape[0] != FEATURES:
        raise ValueError("expected a (512, columns) device tensor")
    if feat.element_size() != 2 or feat.stride(1) != 1 or feat.stride(0) % 64 or feat.data_ptr() % 16:
        raise ValueError("features: 2-byte elements, unit column stride, row stride a multiple of 64, "
                         "16-byte aligned start")
    import torch
    torch.cuda.current_stream(feat.device).synchronize()
    return feat.data_ptr(), feat.stride(0), feat.shape[1]


def _need_device(t, count: int, itemsize: int):
    """Check a device-tensor argument and wait for torch's pending work on it
    (the library runs on its own streams).  torch must have been imported
    before the library was loaded (see lib())."""
    if not getattr(t, "is_cuda", False) or not t.is_contiguous():
        raise ValueError("expected a contiguous device tensor")
    if t.numel() * t.element_size() < count * itemsize:
        raise ValueError(f"device tensor too small: need {count} x {itemsize} bytes")
    if t.data_ptr() % 16:
        raise ValueError("device tensor must be 16-byte aligned")
    import torch
    torch.cuda.current_stream(t.device).synchronize()


def planes_to_codes(a: np.ndarray, b: np.ndarray, width: Optional[int] = None) -> np.ndarray:
    """Unpack bit-planes (n, words) into byte codes (n, width)."""
    n, words = a.shape
    au = np.unpackbits(a.view(np.uint8).reshape(n, words * 8), axis=1, bitorder="little")
    bu = np.unpackbits(b.view(np.uint8).reshape(n, words * 8), axis=1, bitorder="little")
    return (bu | ((au ^ bu) << 1)).astype(np.uint8)[:, :width]


def plane_to_bytes(m: np.ndarray, width: Optional[int] = None) -> np.ndarray:
    n, words = m.shape
    return np.unpackbits(m.view(np.uint8).reshape(n, words * 8), axis=1,
                         bitorder="little").astype(np.uint8)[:, :width]
