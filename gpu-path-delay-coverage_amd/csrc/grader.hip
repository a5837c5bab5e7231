// MI355X grading path: sim -> mark -> merge -> cover on node-major bit-planes.
//
// Layout in HBM for one chunk of W pairs (words = ceil(W/64); rows padded to
// `pitch` words, a multiple of 16, so every row starts on a 128-byte line):
//   V[n][pitch]   value of each node under vector start+64k+j (bit j, word k).
//                 In sequence mode pair p is (V[p], V[p+1]) (simkernel.cuh:
//                 115-124), so one plane holds both frames: the frame-2 word
//                 is the frame-1 word shifted by one bit, and word `words`
//                 holds the 64 vectors after the chunk (for the last pair).
//   VB[n][pitch]  frame-2 values, explicit-pair mode only.
//   M[n][pitch]   path marks, written only where non-zero;
//   NZ[n][tiles]  bit (w & 63) of word (w >> 6): M[g][w] was written in this
//                 chunk (a mark word with its bit clear reads as zero).
// A node's 4-valued code is (b | (a^b) << 1): S0, S1, T0 (1->0), T1 (0->1),
// the encoding of the reference's tables (simkernel.cuh:158-165).
//
// Kernels (one launch per level where the reference launches per level):
//   k_fwd      simulation (simkernel.cuh:96-156) and forward marking
//              (markkernel.cuh:215-270), fused or alone
//   k_reverse  invalid-mark removal (markkernel.cuh:271-293) and the
//              first-detection merge (mergekernel.cuh:116-198), fused or alone;
//              also lists each word's marked nodes in descending level order
//   k_events   distinct first-detection pairs of the chunk
//   k_cover    exact uint64 reverse path-count DP (coverkernel.cuh:171-230)
//              over the nodes marked at each such pair only: a pair with no
//              first detection has c = 0 at every node (coverkernel.cuh:
//              214-220), so it contributes nothing and is skipped.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "mlp_internal.hpp"
#include "pdf_internal.hpp"

using pdf::Netlist;
using pdf::set_error;

namespace {

constexpr int kBlock = 256;
// Fanins kept in registers by the mark pass.  The copies are indexed only by
// compile-time constants (fully unrolled loops): a runtime index into a
// private array becomes an indexed register write, which for an index past
// the array overwrites unrelated live registers.
constexpr int kRegFanin = 4;

// Per-node record read with one 32-byte load (a scalar load when the node is
// the same for the whole wave): type, PO flag, the first four fanins inline and
// the CSR offsets of the full fanin and fanout lists.  reach: how far past the
// node's id the last node of the level of its last reader in the level walk's
// forward pass lies (a reader: a gate that reads it as fanin 0, or as any
// fanin of an AND/NAND/OR/NOR), at most 65,535; 0 when nothing reads it.
// The walk reads a fanin from its LDS window unless that distance exceeds
// the window, so a node whose reach fits the window is never read back.
struct alignas(32) GateRec {
  uint32_t fin[4];  // first four fanins (0 past nfi)
  uint32_t fin_off, fout_off;
  uint16_t nfi, nfo;
  uint8_t type, po;
  uint16_t reach;
};
static_assert(sizeof(GateRec) == 32, "GateRec is one 32-byte load");

struct DevNet {
  const GateRec *rec;
  const uint8_t *type;
  const uint8_t *po;
  const uint32_t *fin_off;
  const uint32_t *fin;
  const uint32_t *fout_off;
  const uint32_t *fout;
  const uint32_t *level;
  uint32_t n;
};

struct Planes {
  uint64_t *V, *VB, *M, *NZ;
  uint64_t *ACC;       // removal accumulator: OR of fanouts' final marks
  uint64_t pitch;      // words per plane row
  uint32_t words;      // pair words in this chunk
  uint32_t ntiles;     // NZ words per node
  uint64_t last_mask;  // valid pairs of the chunk's last word
  uint32_t merge_window = 0;  // 0: merge every pair of the chunk; else only its first merge_window pairs
  // after a level walk the walk's mark-present bytes stand in for NZ: bit
  // (w & (2^bt_shift - 1)) of bt[(w >> bt_shift) * bt_n + f] for word w of node f
  const uint8_t *bt = nullptr;
  uint32_t bt_shift = 0, bt_n = 0;
};

// Pairs of word w that the merge may take as first detections: all of them,
// or with a merge window only those whose index in the chunk is below it
// (the reference examines the first 4,224 patterns of a chunk only,
// mergekernel.cuh:125-130).
__device__ __forceinline__ uint64_t merge_mask(const Planes &pl, uint32_t w) {
  if (!pl.merge_window) return ~0ULL;
  const uint64_t lo = 64ull * w;
  if (lo >= pl.merge_window) return 0;
  const uint64_t room = pl.merge_window - lo;
  return room >= 64 ? ~0ULL : (1ULL << room) - 1;
}

// Bits of the per-grade error word: an exact uint64 count overflowed, and
// the flags of values that leave the reference's ranges (results stay exact)
constexpr uint32_t kErrOverflow = 1u;
constexpr uint32_t kErrRefU16 = 2u;  // a c / h the reference packs into 16 bits (coverkernel.cuh:41-43)
constexpr uint32_t kErrRefU32 = 4u;  // a chunk total past the reference's uint32 sums (coverkernel.cuh:79-101)

// Thread -> (gate, word) for a level launch.  A block of kBlock threads covers
// (kBlock >> twlog) gates x (1 << twlog) consecutive words; block b takes word
// tile b % T and gate group b / T, so with T a multiple of 8 a word tile stays
// on one XCD (blocks are dealt round-robin to the 8 XCDs) at every level and
// its fanin rows are re-read from that XCD's L2.  Blocks from `main_blocks`
// on (sequence-mode simulation only) compute the extra column `words`, one
// gate per thread.
// A (gate, 64-word tile) item of the compacted reverse pass: the gate, the
// tile and its mark-present bits (bit u: a forward mark at word 64 tile + u).
struct RevTile {
  uint32_t gate, tile;
  uint64_t bits;
};

struct LevelMap {
  uint32_t T;      // word tiles
  uint32_t twlog;  // log2 tile width in words (<= 6)
  uint32_t words;  // words in the chunk
  uint32_t lo, count;
  uint32_t main_blocks;
  // the forward pass appends the level's marked (gate, 64-word tile) items
  // (PDF_TUNE_REV_COMPACT; one gate per wave only) to items_nsub sub-lists,
  // tile t to sub-list t % items_nsub (or, with items_by_block, the
  // launch's block b to sub-list b % 64), each of items_cap entries at
  // items + sub * items_cap, its count at items_cnt[sub * kRevCntStride];
  // an item's tile field is t | items_tag (a batch's circuit index)
  RevTile *items = nullptr;
  uint32_t *items_cnt = nullptr;
  uint32_t items_cap = 0, items_nsub = 1;
  uint32_t items_tag = 0, items_by_block = 0;
  // the forward pass's records are the level's in fanin-locality order
  // (fanin_order), fout_off holding each one's node id
  uint32_t perm = 0;
};

// Sub-lists per level of the compacted reverse pass, at most: the forward
// pass's appends spread over one counter per tile (tiles past 64 share),
// each in a 64-byte line of its own, so they do not queue on one address.
constexpr uint32_t kRevSubLists = 64;
constexpr uint32_t kRevCntStride = 16;  // u32 words between counters

__device__ __forceinline__ bool map_item(const LevelMap &mp, uint32_t b, uint32_t &g, uint32_t &w) {
  const uint32_t grp = b / mp.T, tile = b - grp * mp.T;
  w = (tile << mp.twlog) + (threadIdx.x & ((1u << mp.twlog) - 1));
  const uint32_t gi = grp * (kBlock >> mp.twlog) + (threadIdx.x >> mp.twlog);
  g = mp.lo + gi;
  return w < mp.words && gi < mp.count;
}

__device__ __forceinline__ uint64_t shift_next(uint64_t cur, uint64_t nxt) { return (cur >> 1) | (nxt << 63); }

// Holds x in a register at this point of the program (an empty asm statement
// with x as input; no instruction is emitted): the loads x depends on complete
// here, so the compiler cannot sink them past a later early exit, and loads
// issued together before it share one round trip.
__device__ __forceinline__ void pin(uint64_t x) { asm volatile("" : : "v"(x)); }
__device__ __forceinline__ void pin(uint32_t x) { asm volatile("" : : "v"(x)); }
// Like pin, and every later use of x reads it after this point, so work on x
// cannot be moved up between loads issued before it.
__device__ __forceinline__ void hold(uint64_t &x) { asm volatile("" : "+v"(x)); }
__device__ __forceinline__ void hold(uint32_t &x) { asm volatile("" : "+v"(x)); }

// LDS words typed as LDS (address space 3), so every access through them is
// an LDS instruction with no generic-address check
typedef __attribute__((address_space(3))) uint64_t lds_u64;
typedef __attribute__((address_space(3))) uint32_t lds_u32;

// OR x into a removal accumulator word: a device-scope atomic, or for the
// level walk (whose accumulator words each belong to one workgroup) a
// workgroup-scope one.
template <bool WG>
__device__ __forceinline__ void push_acc(uint64_t *p, uint64_t x) {
  if (WG) __hip_atomic_fetch_or(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  else atomicOr((unsigned long long *)p, (unsigned long long)x);
}

// Frames (a, b) of node f at word w.
template <bool SEQ>
__device__ __forceinline__ void frames(const Planes &pl, uint32_t f, uint32_t w, uint64_t &a, uint64_t &b) {
  const uint64_t at = (uint64_t)f * pl.pitch + w;
  a = pl.V[at];
  b = SEQ ? shift_next(a, pl.V[at + 1]) : pl.VB[at];
}

// Whether node f's mark word w was written in this chunk (bit 0).
__device__ __forceinline__ uint32_t mark_present(const Planes &pl, uint32_t f, uint32_t w) {
  return pl.bt ? (uint32_t)(pl.bt[(uint64_t)(w >> pl.bt_shift) * pl.bt_n + f] >> (w & ((1u << pl.bt_shift) - 1)))
               : (uint32_t)(pl.NZ[(uint64_t)f * pl.ntiles + (w >> 6)] >> (w & 63));
}

// Mark word of node f (zero unless written in this chunk).
__device__ __forceinline__ uint64_t mark_of(const Planes &pl, uint32_t f, uint32_t w) {
  return (mark_present(pl, f, w) & 1) ? pl.M[(uint64_t)f * pl.pitch + w] : 0;
}

// Off-path sensitization predicate S(on, off) of markkernel.cuh:36-41 in
// bit-plane form (one bit per pair).  LUT_R: AND/NAND on in {T0,T1}, off = S1;
// OR/NOR off = S0.  LUT_NR: AND/NAND (on=T0, off in {S1,T1}) or (on=T1,
// off=S1); OR/NOR (on=T1, off in {S0,T0}) or (on=T0, off=S0).
template <bool ROBUST, bool ANDLIKE>
__device__ __forceinline__ uint64_t sens(uint64_t aon, uint64_t bon, uint64_t aoff, uint64_t boff) {
  if (ROBUST) return (aon ^ bon) & (ANDLIKE ? (aoff & boff) : ~(aoff | boff));
  if (ANDLIKE) return (aon & ~bon & boff) | (~aon & bon & aoff & boff);
  return (~aon & bon & ~boff) | (aon & ~bon & ~aoff & ~boff);
}

// OR_i AND_{j != i} S(fanin_i, fanin_j) over fanin positions, for
// nfi <= kRegFanin fanins held in registers.  Every S above factors as
// S(on, off) = (X(on) & U(off)) | (Y(on) & W(off)) with X and Y disjoint
// (robust: X = on transitions, U = the off-path value the gate needs, Y = 0;
// non-robust: X and Y the two transition directions, U and W what each asks
// of the off-path input).  So per bit AND_{j != i} S(i, j) is
// X_i & AND_{j != i} U_j  |  Y_i & AND_{j != i} W_j   (for nfi >= 2),
// and the exclusive ANDs come from prefix / suffix ANDs: O(nfi) work, not
// O(nfi^2).  nfi = 1 gives the empty AND (all ones), nfi = 0 nothing.
template <bool ROBUST, bool ANDLIKE>
__device__ __forceinline__ uint64_t side_ok(const uint64_t (&fa)[kRegFanin], const uint64_t (&fb)[kRegFanin],
                                            uint32_t nfi) {
  static_assert(kRegFanin == 4, "the exclusive ANDs below are written out for four fanins");
  if (nfi < 2) return nfi ? ~0ULL : 0ULL;
  uint64_t X[4], Y[4], U[4], W[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint64_t a = fa[j], b = fb[j];
    if (ROBUST) {
      X[j] = a ^ b;
      Y[j] = 0;
      U[j] = ANDLIKE ? (a & b) : ~(a | b);
      W[j] = ~0ULL;
    } else if (ANDLIKE) {
      X[j] = a & ~b;  // on T0 needs off in {S1, T1}
      U[j] = b;
      Y[j] = ~a & b;  // on T1 needs off = S1
      W[j] = a & b;
    } else {
      X[j] = ~a & b;  // on T1 needs off in {S0, T0}
      U[j] = ~b;
      Y[j] = a & ~b;  // on T0 needs off = S0
      W[j] = ~a & ~b;
    }
    if ((uint32_t)j >= nfi) { X[j] = Y[j] = 0; U[j] = W[j] = ~0ULL; }
  }
  const uint64_t u01 = U[0] & U[1], u23 = U[2] & U[3];
  uint64_t r = (X[0] & U[1] & u23) | (X[1] & U[0] & u23) | (X[2] & u01 & U[3]) | (X[3] & u01 & U[2]);
  if (!ROBUST) {
    const uint64_t w01 = W[0] & W[1], w23 = W[2] & W[3];
    r |= (Y[0] & W[1] & w23) | (Y[1] & W[0] & w23) | (Y[2] & w01 & W[3]) | (Y[3] & w01 & W[2]);
  }
  return r;
}

// side_ok for a gate whose AND/OR class differs between lanes: complementing
// every frame turns each OR-like predicate into the AND-like one (T0 <-> T1,
// S0 <-> S1 map LUT_R / LUT_NR's OR/NOR rows onto their AND/NAND rows,
// markkernel.cuh:36-37), so one branch-free evaluation serves both.
template <bool ROBUST>
__device__ __forceinline__ uint64_t side_ok_any(const uint64_t (&fa)[kRegFanin], const uint64_t (&fb)[kRegFanin],
                                                uint32_t nfi, uint64_t orlike_mask) {
  uint64_t ca[kRegFanin], cb[kRegFanin];
#pragma unroll
  for (int j = 0; j < kRegFanin; ++j) {
    ca[j] = fa[j] ^ orlike_mask;
    cb[j] = fb[j] ^ orlike_mask;
  }
  return side_ok<ROBUST, true>(ca, cb, nfi);
}

// Same predicate for gates with more fanins: operands re-read from the planes.
template <bool SEQ, bool ROBUST, bool ANDLIKE>
__device__ uint64_t side_ok_mem(const Planes &pl, const uint32_t *fins, uint32_t nfi, uint32_t w) {
  uint64_t r = 0;
  for (uint32_t i = 0; i < nfi; ++i) {
    uint64_t ai, bi;
    frames<SEQ>(pl, fins[i], w, ai, bi);
    uint64_t acc = ~0ULL;
    for (uint32_t j = 0; j < nfi && acc; ++j) {
      if (j == i) continue;
      uint64_t aj, bj;
      frames<SEQ>(pl, fins[j], w, aj, bj);
      acc &= sens<ROBUST, ANDLIKE>(ai, bi, aj, bj);
    }
    r |= acc;
  }
  return r;
}

// The chained 2-input fold of simkernel.cuh:142-148 for AND/NAND/OR/NOR.
__device__ __forceinline__ void fold(uint8_t ty, uint64_t &v, uint64_t x) {
  switch (ty) {
    case PDF_AND: v &= x; break;
    case PDF_NAND: v = ~(v & x); break;
    case PDF_OR: v |= x; break;
    default: v = ~(v | x); break;
  }
}

// The same fold without a branch on the type (for lanes of different gate
// types): om = all ones for OR/NOR, inv = all ones for NAND/NOR.
__device__ __forceinline__ void fold_masks(uint8_t ty, uint64_t &om, uint64_t &inv) {
  om = (ty == PDF_OR || ty == PDF_NOR) ? ~0ULL : 0ULL;
  inv = (ty == PDF_NAND || ty == PDF_NOR) ? ~0ULL : 0ULL;
}
__device__ __forceinline__ void fold_m(uint64_t &v, uint64_t x, uint64_t om, uint64_t inv) {
  v = ((v & x) | (om & (v | x))) ^ inv;
}

// One frame of one node at word w (used for the extra sequence column).
__device__ __forceinline__ uint64_t sim_one_frame(const DevNet &net, const Planes &pl, const GateRec &r,
                                                  uint32_t w) {
  const uint8_t ty = r.type;
  const bool multi = ty >= PDF_AND && ty <= PDF_NOR;
  // the register fanins' words, all issued before any is folded
  uint64_t x[kRegFanin];
#pragma unroll
  for (int j = 0; j < kRegFanin; ++j) {
    x[j] = 0;
    if (j == 0 || (multi && (uint32_t)j < r.nfi)) x[j] = pl.V[(uint64_t)r.fin[j] * pl.pitch + w];
  }
#pragma unroll
  for (int j = 0; j < kRegFanin; ++j) hold(x[j]);
  uint64_t v = x[0];
  if (ty == PDF_NOT) v = ~v;
  if (r.nfi > 1) {
    if (multi) {
#pragma unroll
      for (int j = 1; j < kRegFanin; ++j)
        if ((uint32_t)j < r.nfi) fold(ty, v, x[j]);
      for (uint32_t j = kRegFanin; j < r.nfi; ++j)
        fold(ty, v, pl.V[(uint64_t)net.fin[r.fin_off + j] * pl.pitch + w]);
    } else if (ty != PDF_BUFF && ty != PDF_DFF && ty != PDF_FROM) {
      v = 0;  // XOR/XNOR/NOT/others with >= 2 fanins give S0 (simkernel.cuh:89-92)
    }
  }
  return v;
}

// Forward pass over one level: simulation (DO_SIM) and/or forward marking
// (DO_MARK).  SEQ: consecutive-vector pairs from one value plane; otherwise
// explicit pairs with a second plane.  WIDE: 64-word tiles, so each wave
// works on one node and its record is read with scalar loads.  NW (WIDE
// only): words per lane, w and w + 64 of a 128-word tile; every load of both
// words is issued before any of them is used, so a level's waves fit in one
// resident round with twice the loads in flight.
// The body takes its block index `blk` of `nblk` explicitly, so one launch
// can also run the levels of several circuits (k_fwd_batch).
template <bool SEQ, bool DO_SIM, bool DO_MARK, bool ROBUST, bool WIDE, int NW>
__device__ __forceinline__ void fwd_body(const DevNet &net, const GateRec *__restrict__ recs, const Planes &pl,
                                         const uint64_t *__restrict__ va, const uint64_t *__restrict__ vb,
                                         uint64_t vstride, uint64_t wbase, const LevelMap &mp,
                                         uint32_t *__restrict__ clear32, uint32_t clear32_n, uint32_t blk,
                                         uint32_t nblk) {
  static_assert(NW == 1 || WIDE, "two words per lane needs whole-wave tiles");
  if (SEQ && DO_SIM && blk >= mp.main_blocks) {
    // extra column: the node values of the 64 vectors after the chunk
    const uint32_t gi = (blk - mp.main_blocks) * kBlock + threadIdx.x;
    // these threads also clear a caller-given u32 range (the next pass's
    // counters), striding over it
    const uint32_t nthreads = (nblk - mp.main_blocks) * kBlock;
    for (uint32_t i = gi; i < clear32_n; i += nthreads) clear32[i] = 0;
    if (gi >= mp.count) return;
    const GateRec r = recs[mp.lo + gi];
    const uint32_t g = mp.perm ? r.fout_off : mp.lo + gi;
    pin((uint32_t)r.type);
    pin((uint32_t)r.nfi);
    pin(r.fin[0]);
    const uint64_t v =
        r.type == PDF_INPT ? va[(uint64_t)g * vstride + wbase + mp.words] : sim_one_frame(net, pl, r, mp.words);
    pl.V[(uint64_t)g * pl.pitch + mp.words] = v;
    return;
  }
  uint32_t g, w;
  if (NW == 1) {
    if (!map_item(mp, blk, g, w)) return;
  } else {
    const uint32_t b = blk;
    const uint32_t grp = b / mp.T, tile = b - grp * mp.T;
    const uint32_t gi = grp * (kBlock >> 6) + (threadIdx.x >> 6);
    if (gi >= mp.count) return;
    g = mp.lo + gi;
    w = tile * 64 * NW + (threadIdx.x & 63);
  }
  if (WIDE) g = __builtin_amdgcn_readfirstlane(g);
  const GateRec r = recs[g];
  if (mp.perm) g = r.fout_off;  // the record's position was its place in the order
  // the whole record in the first round trip (the compiler would otherwise
  // fetch the fanin ids only after testing the type)
  pin((uint32_t)r.type);
  pin((uint32_t)r.nfi);
  pin(r.fin[0]);
  const uint8_t ty = r.type;
  bool live[NW];
#pragma unroll
  for (int u = 0; u < NW; ++u) live[u] = NW == 1 || w + 64 * u < mp.words;
  uint64_t a[NW], b[NW], m[NW];
#pragma unroll
  for (int u = 0; u < NW; ++u) { b[u] = 0; m[u] = 0; }
  if (ty == PDF_INPT) {
    // simkernel.cuh:115-124: pair (V[q], V[q+1 mod P])
#pragma unroll
    for (int u = 0; u < NW; ++u) {
      if (!live[u]) continue;
      const uint32_t wu = w + 64 * u;
      const uint64_t row = (uint64_t)g * pl.pitch + wu;
      const uint64_t q = (uint64_t)g * vstride + wbase + wu;
      if (DO_SIM) {
        a[u] = va[q];
        if (!SEQ) b[u] = vb[q];
        else if (DO_MARK) b[u] = shift_next(a[u], va[q + 1]);
        pl.V[row] = a[u];
        if (!SEQ) pl.VB[row] = b[u];
      } else if (DO_MARK) {
        frames<SEQ>(pl, g, wu, a[u], b[u]);
      }
      // pairs past the last one of the chunk never mark; every other mark
      // needs its fanin 0 marked, so this keeps them clear everywhere
      if (DO_MARK) m[u] = (a[u] ^ b[u]) & (wu + 1 == mp.words ? pl.last_mask : ~0ULL);
    }
  } else {
    const uint32_t nfi = r.nfi;
    const uint32_t *more = net.fin + r.fin_off;  // fanins from kRegFanin on
    const bool multi = ty >= PDF_AND && ty <= PDF_NOR;
    constexpr bool kBoth = DO_MARK || !SEQ;  // frame 2 of the fanins needed
    // issued with the value loads, so a marked fanin 0 costs one more round
    // trip (its mark word) instead of two
    uint64_t nz0[NW];
#pragma unroll
    for (int u = 0; u < NW; ++u)
      nz0[u] = DO_MARK && live[u] ? pl.NZ[(uint64_t)r.fin[0] * pl.ntiles + ((w + 64 * u) >> 6)] : 0;
    uint64_t fa[NW][kRegFanin], fb[NW][kRegFanin];  // constant indices only
    if (DO_SIM) {
      // the fanins' raw words, all issued before any is used (one round trip):
      // frame 1, and frame 2's source (the next word in sequence mode, the
      // second plane otherwise)
#pragma unroll
      for (int j = 0; j < kRegFanin; ++j) {
#pragma unroll
        for (int u = 0; u < NW; ++u) fa[u][j] = fb[u][j] = 0;
        if (j == 0 || (multi && (uint32_t)j < nfi)) {
#pragma unroll
          for (int u = 0; u < NW; ++u) {
            if (!live[u]) continue;
            const uint64_t at = (uint64_t)r.fin[j] * pl.pitch + w + 64 * u;
            fa[u][j] = pl.V[at];
            if (kBoth) fb[u][j] = SEQ ? pl.V[at + 1] : pl.VB[at];
          }
        }
      }
#pragma unroll
      for (int j = 0; j < kRegFanin; ++j)
#pragma unroll
        for (int u = 0; u < NW; ++u) {
          hold(fa[u][j]);
          if (kBoth) hold(fb[u][j]);
          if (kBoth && SEQ) fb[u][j] = shift_next(fa[u][j], fb[u][j]);
        }
#pragma unroll
      for (int u = 0; u < NW; ++u) {
        a[u] = fa[u][0];
        b[u] = fb[u][0];
        if (ty == PDF_NOT) { a[u] = ~a[u]; b[u] = ~b[u]; }
        if (nfi > 1) {
          if (multi) {
#pragma unroll
            for (int j = 1; j < kRegFanin; ++j)
              if ((uint32_t)j < nfi) { fold(ty, a[u], fa[u][j]); fold(ty, b[u], fb[u][j]); }
            for (uint32_t j = kRegFanin; j < nfi && live[u]; ++j) {
              uint64_t xa, xb = 0;
              if (kBoth) frames<SEQ>(pl, more[j], w + 64 * u, xa, xb);
              else xa = pl.V[(uint64_t)more[j] * pl.pitch + w + 64 * u];
              fold(ty, a[u], xa);
              fold(ty, b[u], xb);
            }
          } else if (ty != PDF_BUFF && ty != PDF_DFF && ty != PDF_FROM) {
            a[u] = 0;  // XOR/XNOR/NOT/others with >= 2 fanins give S0 (simkernel.cuh:89-92)
            b[u] = 0;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < NW; ++u) {
        if (!live[u]) continue;
        const uint64_t row = (uint64_t)g * pl.pitch + w + 64 * u;
        pl.V[row] = a[u];
        if (!SEQ) pl.VB[row] = b[u];
      }
    } else if (DO_MARK) {
#pragma unroll
      for (int u = 0; u < NW; ++u) {
        a[u] = b[u] = 0;
        if (live[u]) frames<SEQ>(pl, g, w + 64 * u, a[u], b[u]);
      }
    }
    if (DO_MARK) {
#pragma unroll
      for (int u = 0; u < NW; ++u) {
        const uint32_t wu = w + 64 * u;
        const uint64_t tr = live[u] ? a[u] ^ b[u] : 0;
        if (!tr) continue;
        const uint64_t m0 = ((nz0[u] >> (wu & 63)) & 1) ? pl.M[(uint64_t)r.fin[0] * pl.pitch + wu] : 0;
        if (ty == PDF_BUFF || ty == PDF_DFF || ty == PDF_FROM || ty == PDF_NOT) {
          m[u] = m0 & tr;
        } else if (ty >= PDF_AND && ty <= PDF_XNOR) {
          if (nfi == 1) {
            m[u] = m0 & tr;  // empty AND over off-path inputs
          } else if (multi && (m0 & tr)) {
            const bool andlike = ty == PDF_AND || ty == PDF_NAND;
            uint64_t rr;
            if (nfi <= kRegFanin) {
              if (!DO_SIM) {
#pragma unroll
                for (int j = 0; j < kRegFanin; ++j)
                  if ((uint32_t)j < nfi) frames<SEQ>(pl, r.fin[j], wu, fa[u][j], fb[u][j]);
              }
              rr = andlike ? side_ok<ROBUST, true>(fa[u], fb[u], nfi) : side_ok<ROBUST, false>(fa[u], fb[u], nfi);
            } else {
              rr = andlike ? side_ok_mem<SEQ, ROBUST, true>(pl, more, nfi, wu)
                           : side_ok_mem<SEQ, ROBUST, false>(pl, more, nfi, wu);
            }
            // markkernel.cuh:253-261: the on-path mark is always fanin 0's
            m[u] = rr & m0 & tr;
          }
          // XOR/XNOR with >= 2 fanins never sensitize (LUT entries 0)
        }
      }
    }
  }
  if (DO_MARK) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t tw = 1u << mp.twlog;
    uint64_t tile_bits[NW];  // one gate per wave: the wave's NZ words
#pragma unroll
    for (int u = 0; u < NW; ++u) {
      const uint32_t wu = w + 64 * u;
      if (live[u] && m[u]) pl.M[(uint64_t)g * pl.pitch + wu] = m[u];
      // one NZ word per (gate, 64-word tile): bits from this gate's lanes
      const uint64_t bal = __ballot(live[u] && m[u] != 0);
      tile_bits[u] = bal;
      if ((lane & (tw - 1)) == 0 && live[u]) {
        const uint64_t bits = mp.twlog == 6 ? bal : (bal >> lane) & ((1ULL << tw) - 1);
        pl.NZ[(uint64_t)g * pl.ntiles + (wu >> 6)] = bits;
      }
    }
    if (WIDE && mp.items && lane == 0) {
      // the compacted reverse pass's items (lane 0 holds the wave's first
      // word, so (w + 64 u) / 64 is tile u's index)
#pragma unroll
      for (int u = 0; u < NW; ++u) {
        if (!tile_bits[u]) continue;
        const uint32_t t = (w + 64u * u) >> 6;
        const uint32_t sub = mp.items_by_block ? blockIdx.x % kRevSubLists : t % mp.items_nsub;
        const uint32_t pos = atomicAdd(mp.items_cnt + sub * kRevCntStride, 1u);
        mp.items[(uint64_t)sub * mp.items_cap + pos] = RevTile{g, t | mp.items_tag, tile_bits[u]};
      }
    }
  }
}

// The robust two-words-per-lane pass is held to 64 VGPRs, eight waves per
// SIMD (it otherwise takes 69 and runs seven).  Measured at 2M gates
// (same-box alternating runs): 107.3-109.4 -> 103.5-103.8 us per level
// launch, with a 16-byte-per-lane register spill.
template <bool SEQ, bool DO_SIM, bool DO_MARK, bool ROBUST, bool WIDE, int NW>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(NW == 2 && ROBUST ? 8 : 1))) void k_fwd(
    DevNet net, const GateRec *__restrict__ recs, Planes pl, const uint64_t *__restrict__ va,
    const uint64_t *__restrict__ vb, uint64_t vstride, uint64_t wbase, LevelMap mp, uint32_t *__restrict__ clear32,
    uint32_t clear32_n) {
  fwd_body<SEQ, DO_SIM, DO_MARK, ROBUST, WIDE, NW>(net, recs, pl, va, vb, vstride, wbase, mp, clear32, clear32_n,
                                                   blockIdx.x, gridDim.x);
}

// Reverse pass over one level.  DO_REMOVE: m &= po | OR_fanouts m
// (markkernel.cuh:279-291), gathered by the fanouts: every node whose final
// mark is non-zero ORs it into ACC of each fanin marked at the same word
// (levels run in descending order, so all fanouts of a node have pushed
// before its level runs); the reader clears ACC again, so it is all zero
// between passes.  DO_MERGE: earliest marked T0/T1 pair per node
// (mergekernel.cuh:141-197) and the per-word list of marked nodes.
template <bool SEQ, bool DO_REMOVE, bool DO_MERGE, bool WALK = false>
__device__ __forceinline__ void rev_marked(const DevNet &net, const GateRec &r, const Planes &pl, uint32_t g,
                                           uint32_t w, uint32_t start, uint32_t *__restrict__ first_t0,
                                           uint32_t *__restrict__ first_t1, uint32_t *__restrict__ list,
                                           uint32_t *__restrict__ list_cnt, uint64_t list_stride,
                                           const uint64_t *fa_src = nullptr, const uint64_t *fb_src = nullptr,
                                           lds_u32 *walk_cnt = nullptr, uint32_t w0 = 0,
                                           const uint64_t *bits_src = nullptr, uint32_t bits_stride = 0);

// One (gate, word) item of the reverse pass.
template <bool SEQ, bool DO_REMOVE, bool DO_MERGE>
__device__ __forceinline__ void rev_item(const DevNet &net, const GateRec *__restrict__ recs, const Planes &pl,
                                         uint32_t g, uint32_t w, uint32_t start, uint32_t *__restrict__ first_t0,
                                         uint32_t *__restrict__ first_t1, uint32_t *__restrict__ list,
                                         uint32_t *__restrict__ list_cnt, uint64_t list_stride) {
  // first memory trip: this node's mark-present word and record
  const uint64_t nzw = pl.NZ[(uint64_t)g * pl.ntiles + (w >> 6)];
  const GateRec r = recs[g];
  if (DO_REMOVE) {
    pin(r.fin[0]);
    pin((uint32_t)r.nfi);
    pin((uint32_t)r.po);
  }
  if (!((nzw >> (w & 63)) & 1)) return;
  rev_marked<SEQ, DO_REMOVE, DO_MERGE>(net, r, pl, g, w, start, first_t0, first_t1, list, list_cnt, list_stride);
}

// The rest of rev_item for a (gate, word) whose mark-present bit is set; r is
// the gate's record, already loaded.  WALK (the level walk): every push into
// a word's accumulator comes from the workgroup that owns the word, earlier
// in the same launch, so the pushes are workgroup-scope atomics and the
// barrier between levels orders them before the read; the node's frame-1
// word and frame 2's source word are read from fa_src and fb_src (the walk's
// tile-major value plane) instead of V / VB; the word's list length is
// counted in LDS (walk_cnt[w - w0], written out by the caller at the end);
// and with bits_src a fanin f's mark-present bits of the tile are
// bits_src[f * bits_stride] (the walk's own records, bit u: word w0 + u)
// instead of its NZ word, which other workgroups' forward walks update with
// atomics in the same launch.
template <bool SEQ, bool DO_REMOVE, bool DO_MERGE, bool WALK>
__device__ __forceinline__ void rev_marked(const DevNet &net, const GateRec &r, const Planes &pl, uint32_t g,
                                           uint32_t w, uint32_t start, uint32_t *__restrict__ first_t0,
                                           uint32_t *__restrict__ first_t1, uint32_t *__restrict__ list,
                                           uint32_t *__restrict__ list_cnt, uint64_t list_stride,
                                           const uint64_t *fa_src, const uint64_t *fb_src, lds_u32 *walk_cnt,
                                           uint32_t w0, const uint64_t *bits_src, uint32_t bits_stride) {
  const uint64_t row = (uint64_t)g * pl.pitch + w;
  auto nz_word = [&](uint32_t f) -> uint64_t {  // bit (w & 63): f has a forward mark at word w
    if (WALK && bits_src) return bits_src[(uint64_t)f * bits_stride] << (w0 & 63);
    return pl.NZ[(uint64_t)f * pl.ntiles + (w >> 6)];
  };
  if (WALK && DO_REMOVE && !r.po) {
    // most forward marks die in the removal (no marked fanout pushed into
    // them): the accumulator alone decides, and a dying mark only needs its
    // word cleared, with none of the loads below
    if (!pl.ACC[row]) {
      pl.M[row] = 0;
      return;
    }
  }
  // second trip, all loads issued before any is used: this node's mark and
  // accumulator; for the merge, its frames as raw words (frame 1 and frame
  // 2's source) and its first-detection pairs so far; for the removal, the
  // register fanins' mark-present words and mark words (a mark word whose
  // bit is clear is read anyway and ignored below, instead of costing one
  // more round trip per fanin)
  uint64_t m = pl.M[row];
  const uint64_t acc = DO_REMOVE ? pl.ACC[row] : 0;
  uint64_t a = 0, b = 0;
  uint32_t cur0 = 0, cur1 = 0;
  if (DO_MERGE) {
    a = WALK ? *fa_src : pl.V[row];
    b = WALK ? *fb_src : SEQ ? pl.V[row + 1] : pl.VB[row];
    cur0 = first_t0[g];
    cur1 = first_t1[g];
  }
  uint64_t mf[kRegFanin], nzf[kRegFanin];
  if (DO_REMOVE) {
#pragma unroll
    for (int j = 0; j < kRegFanin; ++j) {
      nzf[j] = mf[j] = 0;
      if ((uint32_t)j < r.nfi) {
        nzf[j] = nz_word(r.fin[j]);
        mf[j] = pl.M[(uint64_t)r.fin[j] * pl.pitch + w];
      }
    }
#pragma unroll
    for (int j = 0; j < kRegFanin; ++j) {
      hold(nzf[j]);
      hold(mf[j]);
    }
    if (DO_MERGE) {
      hold(a);
      hold(b);
    }
  }
  if (DO_MERGE && SEQ) b = shift_next(a, b);
  if (DO_REMOVE) {
    __builtin_amdgcn_sched_barrier(0);  // every load above is issued before any use
    pin(m);
    pin(acc);
  }
  if (!m) return;
  if (DO_REMOVE) {
    if (acc) pl.ACC[row] = 0;
    if (!r.po && (m & ~acc)) {
      m &= acc;
      pl.M[row] = m;
    }
    if (m) {
      // the pushes: this node's final mark into ACC of each fanin marked at
      // the same pairs (all-ones mask for a fanin whose bit is set)
#pragma unroll
      for (int j = 0; j < kRegFanin; ++j) mf[j] &= (uint64_t)0 - ((nzf[j] >> (w & 63)) & 1);
#pragma unroll
      for (int j = 0; j < kRegFanin; ++j)
        if (mf[j] & m) push_acc<WALK>(&pl.ACC[(uint64_t)r.fin[j] * pl.pitch + w], mf[j] & m);
      for (uint32_t j = kRegFanin; j < r.nfi; ++j) {
        const uint32_t f = net.fin[r.fin_off + j];
        if (!((nz_word(f) >> (w & 63)) & 1)) continue;
        const uint64_t x = pl.M[(uint64_t)f * pl.pitch + w] & m;
        if (x) push_acc<WALK>(&pl.ACC[(uint64_t)f * pl.pitch + w], x);
      }
    }
  }
  if (DO_MERGE && m) {
    const uint64_t mm = m & merge_mask(pl, w);
    const uint64_t t0 = mm & ~b, t1 = mm & b;  // marked => transition; b = 1 means T1
    const uint32_t base = start + 64u * w;
    // the first-detection pairs only decrease within a grade (reset by an
    // earlier kernel), so a value read above is at least the current one: a
    // candidate that is not below it cannot lower the minimum and is skipped
    const uint32_t c0 = t0 ? base + (uint32_t)__builtin_ctzll(t0) : PDF_NONE;
    const uint32_t c1 = t1 ? base + (uint32_t)__builtin_ctzll(t1) : PDF_NONE;
    if (c0 < cur0) atomicMin(&first_t0[g], c0);
    if (c1 < cur1) atomicMin(&first_t1[g], c1);
    const uint32_t pos = WALK ? __hip_atomic_fetch_add(&walk_cnt[w - w0], 1u, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP)
                              : atomicAdd(&list_cnt[w], 1u);
    list[(uint64_t)w * list_stride + pos] = g;
  }
}

template <bool SEQ, bool DO_REMOVE, bool DO_MERGE, bool WIDE>
__device__ __forceinline__ void rev_body(const DevNet &net, const GateRec *__restrict__ recs, const Planes &pl,
                                         const LevelMap &mp, uint32_t start, uint32_t *__restrict__ first_t0,
                                         uint32_t *__restrict__ first_t1, uint32_t *__restrict__ list,
                                         uint32_t *__restrict__ list_cnt, uint64_t list_stride, uint32_t blk) {
  uint32_t g, w;
  if (!map_item(mp, blk, g, w)) return;
  if (WIDE) g = __builtin_amdgcn_readfirstlane(g);
  rev_item<SEQ, DO_REMOVE, DO_MERGE>(net, recs, pl, g, w, start, first_t0, first_t1, list, list_cnt, list_stride);
}

template <bool SEQ, bool DO_REMOVE, bool DO_MERGE, bool WIDE>
__global__ __launch_bounds__(kBlock) void k_reverse(DevNet net, const GateRec *__restrict__ recs, Planes pl,
                                                    LevelMap mp, uint32_t start,
                                                    uint32_t *__restrict__ first_t0, uint32_t *__restrict__ first_t1,
                                                    uint32_t *__restrict__ list, uint32_t *__restrict__ list_cnt,
                                                    uint64_t list_stride) {
  rev_body<SEQ, DO_REMOVE, DO_MERGE, WIDE>(net, recs, pl, mp, start, first_t0, first_t1, list, list_cnt, list_stride,
                                           blockIdx.x);
}

// Reverse pass with RT consecutive 64-word tiles per wave (64-word tiles
// only): one round trip loads the gate's record and its RT mark-present
// words, and only the tiles holding a mark run rev_marked, one after another.
// Where few (gate, tile) items hold a mark (long chunks of large netlists),
// RT times fewer waves are dispatched.  Block b takes tile group b % TG and
// gate group b / TG, one gate per wave.
template <bool SEQ, bool DO_REMOVE, bool DO_MERGE, int RT>
__global__ __launch_bounds__(kBlock) void k_reverse_tiles(DevNet net, const GateRec *__restrict__ recs, Planes pl,
                                                          LevelMap mp, uint32_t start,
                                                          uint32_t *__restrict__ first_t0,
                                                          uint32_t *__restrict__ first_t1,
                                                          uint32_t *__restrict__ list, uint32_t *__restrict__ list_cnt,
                                                          uint64_t list_stride) {
  const uint32_t TG = (mp.T + RT - 1) / RT;
  const uint32_t grp = blockIdx.x / TG, tg = blockIdx.x - grp * TG;
  const uint32_t gi = grp * (kBlock / 64) + (threadIdx.x >> 6);
  if (gi >= mp.count) return;
  const uint32_t g = __builtin_amdgcn_readfirstlane(mp.lo + gi);
  const uint32_t lane = threadIdx.x & 63;
  uint64_t nz[RT];
#pragma unroll
  for (int k = 0; k < RT; ++k) {
    const uint32_t t = tg * RT + k;
    nz[k] = t < mp.T ? pl.NZ[(uint64_t)g * pl.ntiles + t] : 0;
  }
  const GateRec r = recs[g];
  if (DO_REMOVE) {
    pin(r.fin[0]);
    pin((uint32_t)r.nfi);
    pin((uint32_t)r.po);
  }
#pragma unroll
  for (int k = 0; k < RT; ++k) hold(nz[k]);
#pragma unroll
  for (int k = 0; k < RT; ++k) {
    const uint32_t w = (tg * RT + k) * 64 + lane;
    if (((nz[k] >> lane) & 1) && w < mp.words)
      rev_marked<SEQ, DO_REMOVE, DO_MERGE>(net, r, pl, g, w, start, first_t0, first_t1, list, list_cnt, list_stride);
  }
}

// A level's sub-lists of compacted reverse items, read by one wave: lane s
// loads sub-list s's count (zero past the level's sub-lists); pre is its
// exclusive prefix sum, cnt the level's item count.
__device__ __forceinline__ void sublist_prefix(const uint32_t *__restrict__ items_cnt, uint32_t lane, uint32_t &pre,
                                               uint32_t &cnt) {
  static_assert(kRevSubLists == 64, "one sub-list count per lane");
  const uint32_t c = items_cnt[lane * kRevCntStride];
  uint32_t incl = c;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t o = __shfl_up(incl, off);
    if (lane >= (uint32_t)off) incl += o;
  }
  pre = incl - c;
  cnt = __shfl(incl, 63);
}

// Where item i (< cnt) of the level lies: in the last sub-list whose prefix
// is <= i (a later one starts past i; empty ones share their successor's
// prefix), at i minus that prefix.
__device__ __forceinline__ uint64_t sublist_slot(uint32_t pre, uint32_t i, uint32_t items_cap) {
  const uint32_t sub = 63 - (uint32_t)__builtin_clzll(__ballot(pre <= i));
  return (uint64_t)sub * items_cap + (i - (uint32_t)__shfl(pre, (int)sub));
}

// Compacted reverse pass over one level (PDF_TUNE_REV_COMPACT): the
// (gate, 64-word tile) items holding a forward mark, appended by the forward
// pass (RevTile, in no particular order: the items of a level are
// independent), one item per wave and iteration, striding over the count.
// Where almost every item is unmarked (2M gates: ~97%) this dispatches waves
// only for the marked ones, and their mark-present bits come with the item.
template <bool SEQ, bool DO_REMOVE, bool DO_MERGE>
__global__ __launch_bounds__(kBlock) void k_reverse_list(DevNet net, const GateRec *__restrict__ recs, Planes pl,
                                                         const RevTile *__restrict__ items,
                                                         const uint32_t *__restrict__ items_cnt, uint32_t items_cap,
                                                         uint32_t start, uint32_t *__restrict__ first_t0,
                                                         uint32_t *__restrict__ first_t1,
                                                         uint32_t *__restrict__ list, uint32_t *__restrict__ list_cnt,
                                                         uint64_t list_stride) {
  constexpr uint32_t wpb = kBlock / 64;
  uint32_t pre, cnt;
  sublist_prefix(items_cnt, threadIdx.x & 63, pre, cnt);
  const uint32_t lane = threadIdx.x & 63;
  for (uint32_t i = blockIdx.x * wpb + (threadIdx.x >> 6); i < cnt; i += gridDim.x * wpb) {
    const RevTile it = items[sublist_slot(pre, i, items_cap)];
    const uint32_t g = __builtin_amdgcn_readfirstlane(it.gate);
    const uint32_t t = __builtin_amdgcn_readfirstlane(it.tile);
    const GateRec r = recs[g];
    if (DO_REMOVE) {
      pin(r.fin[0]);
      pin((uint32_t)r.nfi);
      pin((uint32_t)r.po);
    }
    // bits are set only for words of the chunk
    if ((it.bits >> lane) & 1)
      rev_marked<SEQ, DO_REMOVE, DO_MERGE>(net, r, pl, g, t * 64 + lane, start, first_t0, first_t1, list, list_cnt,
                                           list_stride);
  }
}

// ------------------------------------------------------------- level walk
// The forward and reverse passes in one launch each instead of one launch per
// level (the reference's host level loops, simkernel.cuh:213-235 and
// markkernel.cuh:308-342).  Pairs are independent in both passes, so a
// workgroup that owns a tile of TW consecutive words of the chunk can walk
// every level on its own, with a workgroup barrier between levels and no
// wait on any other workgroup.  A lane takes one gate of the level and all
// TW words of the tile.  In sequence mode frame 2 of word u is frame 1 of
// words u and u + 1 shifted by one bit, so each node carries the word after
// the tile too (the halo, also word 0 of the next tile): NF = TW + 1 frame
// words a node; explicit pairs carry both frames, NF = 2 TW.
//
// Where the frames live.  A tile's frames of node g are one NV-word record of
// the tile-major plane VT[tile][n][NV]: a level's lanes take consecutive
// gates, so their stores are contiguous.  The workgroup also keeps the
// records of the most recent `cap` node ids in LDS (slot g % cap): fanins
// come 70% from the previous level in the synthetic netlists, and ids are
// sorted by level, so most fanin records are read from LDS and only older
// ones from VT.  While a level [lo, hi) runs, a record is taken from LDS iff
// its id is >= hi - cap (its slot is not rewritten by this level), and the
// level writes the slots of ids >= hi - cap only (one writer per slot).
// A record's last word holds the tile's mark-present bits of the node (bit u:
// a mark at word u), so a gate reads fanin 0's with its frames.  Marks also
// go to the node-major M plane of the level kernels for the reverse pass
// and the cover, and the bits to a compact byte plane BT[tile][n] that
// both read in place of the NZ plane (one writer per byte: nothing to zero).
// Workgroup b takes tile (b % 8) * per_xcd + b / 8: blocks are dealt to the
// 8 XCDs round-robin, so neighbouring tiles share an XCD and its L2.
constexpr int kWalkBlock = 1024;

// the node window: all 160 KiB of a CU's LDS (with 4-word tiles of
// consecutive pairs exactly 2,048 slots, so a node's slot is its id's low
// bits); the reverse walk reuses it, its list lengths in the last 64 bytes
constexpr uint32_t kWalkLdsBytes = 160 * 1024;
// records move as 16-byte words (ds_read_b128 / global_load_dwordx4)
typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) u64x2 lds_u64x2;

template <bool SEQ, int TW>
struct WalkShape {
  static constexpr int NF = SEQ ? TW + 1 : 2 * TW;           // frame words of a node record
  static constexpr int NV = (NF + 2) & ~1;                   // + the tile's mark-present bits (bit u: word u),
                                                             // padded to whole 16-byte words
  static constexpr int NMK = (TW + 1) & ~1;                  // mark words kept beside a record in LDS
  static constexpr int SV = NV + NMK;                        // words of an LDS slot: record, then marks
  static constexpr uint32_t cap = kWalkLdsBytes / (8 * SV);  // node slots in LDS
};

// Words of a node's record in the level walk (WalkShape::NV): the frame
// words and the mark-present bits, padded to an even count.
inline uint64_t walk_record_words(bool seq, uint32_t tw) { return ((seq ? tw + 1 : 2 * tw) + 2) & ~1ULL; }

// Words of one lane's walk planes: the tile-major records ([tiles][n][NV])
// followed by the tiles' mark-present bits, one byte per node
// ([tiles][n], bit u: word u of the tile).
inline uint64_t walk_plane_words(bool seq, uint32_t tw, uint64_t n, uint64_t words) {
  const uint64_t tiles = (words + tw - 1) / tw;
  return n * tiles * walk_record_words(seq, tw) + (n * tiles + 7) / 8;
}

template <int TW>
__device__ __forceinline__ bool walk_tile(const Planes &pl, uint32_t tiles, uint32_t &tile, uint32_t &w0,
                                          uint32_t &nw) {
  const uint32_t per_xcd = (tiles + 7) / 8;
  tile = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  if (tile >= tiles) return false;
  w0 = tile * TW;
  nw = min((uint32_t)TW, pl.words - w0);
  return true;
}

// Frame 2 of word u from a node record: the next bit of frame 1 in sequence
// mode, the second frame's word for explicit pairs.
template <bool SEQ, int TW, int NV>
__device__ __forceinline__ uint64_t frame2(const uint64_t (&rec)[NV], int u) {
  if constexpr (SEQ) return shift_next(rec[u], rec[u + 1]);
  else return rec[TW + u];
}

// The first N words of node f's NV-word record of the tile: from its LDS
// slot (SV words apart) when f is in the window, else from the tile-major
// plane.
template <int N, int NV, int SV, uint32_t CAP>
__device__ __forceinline__ void walk_fetch(uint64_t (&dst)[NV], uint32_t f, uint32_t win_lo, const lds_u64 *lds,
                                           const uint64_t *__restrict__ vt) {
  static_assert(NV % 2 == 0 && N <= NV && SV % 2 == 0, "records are whole 16-byte words");
  constexpr int Q = (N + 1) / 2;
  // one generic pointer a lane, no branch: an LDS read and an HBM read in
  // two branches of a wave would wait for each other's registers
  const uint64_t *src = f >= win_lo ? (const uint64_t *)(lds + (f % CAP) * SV) : vt + (uint64_t)f * NV;
  const u64x2 *s = reinterpret_cast<const u64x2 *>(src);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const u64x2 x = s[q];
    dst[2 * q] = x.x;
    dst[2 * q + 1] = x.y;
  }
}

// Node f's TW mark words of the tile: from its LDS slot (after the record)
// when f is in the window, else from the node-major mark plane (row mrow,
// from word w0: w0 is a multiple of TW and the pitch a multiple of 16 words
// at least the chunk's words rounded up, so all TW words are inside the
// row).  Plane words whose mark-present bit is clear are stale; the LDS
// copy holds every word.
template <int TW, int NV, int SV, uint32_t CAP>
__device__ __forceinline__ void walk_fetch_marks(uint64_t (&dst)[TW], uint32_t f, uint32_t win_lo,
                                                 const lds_u64 *lds, const uint64_t *__restrict__ mrow) {
  const uint64_t *src = f >= win_lo ? (const uint64_t *)(lds + (f % CAP) * SV + NV) : mrow;
  if constexpr (TW == 1) {
    dst[0] = src[0];
  } else {
#pragma unroll
    for (int q = 0; q < TW / 2; ++q) {
      const u64x2 y = reinterpret_cast<const u64x2 *>(src)[q];
      dst[2 * q] = y.x;
      dst[2 * q + 1] = y.y;
    }
  }
}

// Sim + forward mark of gate g (record r) over words w0 .. w0+nw-1 of the
// tile; the node's record (frames of every held word) goes to VT and, when
// `keep`, to its LDS slot.  PRE: read the record at pre_src into `pre`
// (the lane's gate of the next level) after this gate's loads.
template <bool SEQ, bool ROBUST, int TW, bool PRE>
__device__ __forceinline__ void walk_fwd_gate(const DevNet &net, const GateRec &r, const Planes &pl,
                                              const uint64_t *__restrict__ va, const uint64_t *__restrict__ vb,
                                              uint64_t vstride, uint64_t vbase, uint32_t g, uint32_t w0,
                                              uint32_t nw, uint64_t *__restrict__ vt, lds_u64 *lds,
                                              uint32_t win_lo, bool keep, const GateRec *pre_src, GateRec &pre,
                                              uint8_t *__restrict__ bt, bool all_records) {
  using S = WalkShape<SEQ, TW>;
  constexpr int NV = S::NV;
  constexpr int NU = SEQ ? TW + 1 : TW;  // frame-1 words held: the tile's and the halo
  const uint8_t ty = r.type;
  uint64_t a[NU], b[TW], m[TW];
#pragma unroll
  for (int u = 0; u < TW; ++u) { b[u] = 0; m[u] = 0; }
  if (ty == PDF_INPT) {
    const uint64_t *src = va + (uint64_t)g * vstride + vbase;
    const uint32_t nu = SEQ ? nw + 1 : nw;
#pragma unroll
    for (int u = 0; u < NU; ++u) a[u] = (uint32_t)u < nu ? src[u] : 0;
    if (!SEQ) {
      const uint64_t *srcb = vb + (uint64_t)g * vstride + vbase;
#pragma unroll
      for (int u = 0; u < TW; ++u) b[u] = (uint32_t)u < nw ? srcb[u] : 0;
    }
    if constexpr (PRE) pre = *pre_src;
#pragma unroll
    for (int u = 0; u < TW; ++u) {
      if ((uint32_t)u >= nw) continue;
      const uint64_t bu = SEQ ? shift_next(a[u], a[SEQ ? u + 1 : u]) : b[u];
      b[u] = bu;
      // pairs past the chunk's last one never mark (every other mark needs
      // its fanin 0 marked)
      m[u] = (a[u] ^ bu) & (w0 + u + 1 == pl.words ? pl.last_mask : ~0ULL);
    }
  } else {
    const uint32_t nfi = r.nfi;
    const uint32_t *more = net.fin + r.fin_off;
    const bool multi = ty >= PDF_AND && ty <= PDF_NOR;
    // (unknown types and XOR/XNOR with >= 2 fanins never mark)
    const bool passes = ty == PDF_BUFF || ty == PDF_DFF || ty == PDF_FROM || ty == PDF_NOT ||
                        (ty >= PDF_AND && ty <= PDF_XNOR && (nfi == 1 || multi));
    // one round trip: the fanin records (LDS or the tile-major plane), fanin
    // 0's mark words of the tile (read whatever its mark-present bits say;
    // words whose bit is clear are stale and masked off below) and the
    // lane's netlist record for the next level
    uint64_t x[kRegFanin][NV];
#pragma unroll
    for (int j = 0; j < kRegFanin; ++j) {
#pragma unroll
      for (int u = 0; u < NV; ++u) x[j][u] = 0;
      if (j == 0 || (multi && (uint32_t)j < nfi)) {
        // fanin 0's record with its mark-present bits, the others' frames
        if (j == 0) walk_fetch<NV, NV, S::SV, S::cap>(x[j], r.fin[j], win_lo, lds, vt);
        else walk_fetch<S::NF, NV, S::SV, S::cap>(x[j], r.fin[j], win_lo, lds, vt);
      }
    }
    uint64_t m0[TW];
#pragma unroll
    for (int u = 0; u < TW; ++u) m0[u] = 0;
    if (passes)
      walk_fetch_marks<TW, NV, S::SV, S::cap>(m0, r.fin[0], win_lo, lds,
                                              pl.M + (uint64_t)r.fin[0] * pl.pitch + w0);
    if constexpr (PRE) pre = *pre_src;
#pragma unroll
    for (int j = 0; j < kRegFanin; ++j)
#pragma unroll
      for (int u = 0; u < NV; ++u) hold(x[j][u]);
#pragma unroll
    for (int u = 0; u < TW; ++u) hold(m0[u]);
    const uint64_t nz0 = x[0][S::NF];  // fanin 0's mark-present bits
    // the chained 2-input fold (simkernel.cuh:131-148) of every held word of
    // every frame
    uint64_t v[S::NF];
#pragma unroll
    for (int u = 0; u < S::NF; ++u) v[u] = ty == PDF_NOT ? ~x[0][u] : x[0][u];
    if (nfi > 1) {
      if (multi) {
        uint64_t om, inv;
        fold_masks(ty, om, inv);
#pragma unroll
        for (int j = 1; j < kRegFanin; ++j) {
          if ((uint32_t)j >= nfi) continue;
#pragma unroll
          for (int u = 0; u < S::NF; ++u) fold_m(v[u], x[j][u], om, inv);
        }
        for (uint32_t j = kRegFanin; j < nfi; ++j) {
          uint64_t y[NV];
          walk_fetch<S::NF, NV, S::SV, S::cap>(y, more[j], win_lo, lds, vt);
#pragma unroll
          for (int u = 0; u < S::NF; ++u) fold_m(v[u], y[u], om, inv);
        }
      } else if (ty != PDF_BUFF && ty != PDF_DFF && ty != PDF_FROM) {
        // XOR/XNOR/NOT/others with >= 2 fanins give S0 (simkernel.cuh:89-92)
#pragma unroll
        for (int u = 0; u < S::NF; ++u) v[u] = 0;
      }
    }
#pragma unroll
    for (int u = 0; u < NU; ++u) a[u] = v[u];
#pragma unroll
    for (int u = 0; u < TW; ++u) b[u] = frame2<SEQ, TW>(v, u);
    // forward mark (markkernel.cuh:228-267): fanin 0's mark words where its
    // mark-present bit is set
#pragma unroll
    for (int u = 0; u < TW; ++u) m0[u] &= (uint64_t)0 - ((nz0 >> u) & 1);
#pragma unroll
    for (int u = 0; u < TW; ++u) {
      if (!passes || (uint32_t)u >= nw) continue;
      const uint64_t on = m0[u] & (a[u] ^ b[u]);
      if (!on) continue;
      if (nfi == 1 || !multi) {
        m[u] = on;  // the empty AND over off-path inputs
      } else {
        const bool andlike = ty == PDF_AND || ty == PDF_NAND;
        uint64_t rr;
        if (nfi <= kRegFanin) {
          uint64_t fa[kRegFanin], fb[kRegFanin];
#pragma unroll
          for (int j = 0; j < kRegFanin; ++j) {
            fa[j] = x[j][u];
            fb[j] = frame2<SEQ, TW>(x[j], u);
          }
          rr = side_ok_any<ROBUST>(fa, fb, nfi, andlike ? 0ULL : ~0ULL);
        } else {
          // more fanins than registers: each off-path pair re-read from LDS / VT
          rr = 0;
          for (uint32_t i = 0; i < nfi; ++i) {
            uint64_t si[NV];
            walk_fetch<S::NF, NV, S::SV, S::cap>(si, more[i], win_lo, lds, vt);
            const uint64_t ai = si[u], bi = frame2<SEQ, TW>(si, u);
            uint64_t acc = ~0ULL;
            for (uint32_t j = 0; j < nfi && acc; ++j) {
              if (j == i) continue;
              uint64_t sj[NV];
              walk_fetch<S::NF, NV, S::SV, S::cap>(sj, more[j], win_lo, lds, vt);
              const uint64_t aj = sj[u], bj = frame2<SEQ, TW>(sj, u);
              acc &= andlike ? sens<ROBUST, true>(ai, bi, aj, bj) : sens<ROBUST, false>(ai, bi, aj, bj);
            }
            rr |= acc;
          }
        }
        // markkernel.cuh:253-261: the on-path mark is always fanin 0's
        m[u] = rr & on;
      }
    }
  }
  // the node's record: frame 1 of the held words (and frame 2 for explicit
  // pairs), then its mark-present bits
  uint64_t rec[NV];
#pragma unroll
  for (int u = S::NF; u < NV; ++u) rec[u] = 0;
  if constexpr (SEQ) {
#pragma unroll
    for (int u = 0; u < NU; ++u) rec[u] = a[u];
  } else {
#pragma unroll
    for (int u = 0; u < TW; ++u) { rec[u] = a[u]; rec[TW + u] = b[u]; }
  }
  // marks to the node-major mark plane, their T1 pairs (mark & frame 2) to
  // the value plane's words (the walk keeps values in its own records; the
  // reverse walk splits its final marks into T0 / T1 by these), both only
  // where the word holds a mark; the mark-present bits to the tile's byte
  const uint64_t row = (uint64_t)g * pl.pitch + w0;
  uint64_t bits = 0;
#pragma unroll
  for (int u = 0; u < TW; ++u)
    if (m[u]) {
      pl.M[row + u] = m[u];
      pl.V[row + u] = m[u] & b[u];
      bits |= 1ULL << u;
    }
  rec[S::NF] = bits;
  bt[g] = (uint8_t)bits;
  // the record to the tile-major plane only where a later level reads it
  // back from there (or every record for the sim-plane getters)
  if (all_records || r.reach > S::cap) {
    u64x2 *dst = reinterpret_cast<u64x2 *>(vt + (uint64_t)g * NV);
#pragma unroll
    for (int q = 0; q < NV / 2; ++q) dst[q] = u64x2{rec[2 * q], rec[2 * q + 1]};
  }
  if (keep) {
    lds_u64x2 *slot = reinterpret_cast<lds_u64x2 *>(lds + (g % S::cap) * S::SV);
#pragma unroll
    for (int q = 0; q < NV / 2; ++q) slot[q] = u64x2{rec[2 * q], rec[2 * q + 1]};
    // every mark word, zeros included (the slot held another node before)
#pragma unroll
    for (int q = 0; q < S::NMK / 2; ++q)
      slot[NV / 2 + q] = u64x2{2 * q < TW ? m[2 * q] : 0, 2 * q + 1 < TW ? m[2 * q + 1] : 0};
  }
}

// Forward walk: levels 0..levels-1 of one tile, sim + forward mark.  recs
// are the records in the walk order (walk_order: each level's gates sorted
// by kind, fout_off holding the id).  VT is the tile-major value plane
// ([tiles][n][NV] words), BT the mark-present bytes ([tiles][n]).
template <bool SEQ, bool ROBUST, int TW>
__device__ __forceinline__ void walk_fwd_body(const DevNet &net, const GateRec *__restrict__ recs, const Planes &pl,
                                              const uint64_t *__restrict__ va, const uint64_t *__restrict__ vb,
                                              uint64_t vstride, uint64_t wbase, uint64_t *__restrict__ VT,
                                              uint8_t *__restrict__ BT, bool all_records,
                                              const uint32_t *__restrict__ level_off, uint32_t levels, uint32_t tiles,
                                              lds_u64 *lds) {
  using S = WalkShape<SEQ, TW>;
  uint32_t tile, w0, nw;
  if (!walk_tile<TW>(pl, tiles, tile, w0, nw)) return;
  uint64_t *vt = VT + (uint64_t)tile * net.n * S::NV;
  uint8_t *bt = BT + (uint64_t)tile * net.n;
  // each lane's first gate of the next level: its netlist record is loaded
  // one level ahead (records do not change), off the level's critical path
  GateRec rn{};
  if (level_off[0] + threadIdx.x < level_off[1]) rn = recs[level_off[0] + threadIdx.x];
  for (uint32_t k = 0; k < levels; ++k) {
    const uint32_t lo = level_off[k], hi = level_off[k + 1];
    const uint32_t win_lo = hi > S::cap ? hi - S::cap : 0;
    {
      const GateRec r0 = rn;
      // the lane's gate of the next level, read unconditionally (lanes past
      // that level's end read a record of this level, never used) into a
      // register that nothing reads before the next level: no wait of this
      // level covers it
      const bool pre = k + 1 < levels && hi + threadIdx.x < level_off[k + 2];
      const GateRec *pre_src = recs + (pre ? hi + threadIdx.x : lo);
      // recs is the walk order: lane i of the level takes the i-th record,
      // whose fout_off is its gate's id
      uint32_t i = lo + threadIdx.x;
      if (i < hi) {
        // the lane's first gate: its record came with the previous level,
        // and the next level's is issued after this gate's loads
        const uint32_t g = r0.fout_off;
        walk_fwd_gate<SEQ, ROBUST, TW, true>(net, r0, pl, va, vb, vstride, wbase + w0, g, w0, nw, vt, lds, win_lo,
                                             g >= win_lo, pre_src, rn, bt, all_records);
        for (i += kWalkBlock; i < hi; i += kWalkBlock) {
          GateRec unused;
          const GateRec r = recs[i];
          walk_fwd_gate<SEQ, ROBUST, TW, false>(net, r, pl, va, vb, vstride, wbase + w0, r.fout_off, w0, nw, vt,
                                                lds, win_lo, r.fout_off >= win_lo, nullptr, unused, bt,
                                                all_records);
        }
      } else {
        rn = *pre_src;
      }
    }
    __syncthreads();  // this level's records are visible to the next level's lanes
  }
}

// Removal and merge of gate g over the marked words of its tile (bit u of
// nzw: a forward mark at word w0 + u), the walk's form of rev_marked: every
// load of the gate goes out in one round trip -- its mark, accumulator and
// T1 words (the forward walk left mark & frame 2 in the value plane), its
// first-detection pairs and its register fanins' mark-present bytes --
// whatever the words turn out to need; a mark that dies (no marked fanout
// pushed into it) then only has its word cleared.  One atomicMin per
// polarity covers the whole tile: words ascend, so the first word with a T0
// (T1) mark holds the tile's earliest pair.
template <bool SEQ, int TW>
__device__ __forceinline__ void walk_rev_gate(const DevNet &net, const GateRec &r, const Planes &pl, uint32_t g,
                                              uint32_t w0, uint64_t nzw, const uint8_t *__restrict__ bt,
                                              uint32_t start, uint32_t *__restrict__ first_t0,
                                              uint32_t *__restrict__ first_t1, uint32_t *__restrict__ list,
                                              uint64_t list_stride, lds_u32 *cnt) {
  const uint64_t row = (uint64_t)g * pl.pitch + w0;
  // TW words from w0: inside the row (w0 is a multiple of TW and the pitch
  // a multiple of 16 words at least the chunk's words rounded up); words
  // without a mark-present bit are stale and never used
  auto load_tile = [&](uint64_t (&dst)[TW], const uint64_t *src) {
    if constexpr (TW == 1) {
      dst[0] = src[0];
    } else {
#pragma unroll
      for (int q = 0; q < TW / 2; ++q) {
        const u64x2 y = reinterpret_cast<const u64x2 *>(src)[q];
        dst[2 * q] = y.x;
        dst[2 * q + 1] = y.y;
      }
    }
  };
  uint64_t m[TW], acc[TW], mb[TW];
  uint32_t fb[kRegFanin];
  load_tile(m, pl.M + row);
  load_tile(acc, pl.ACC + row);
  load_tile(mb, pl.V + row);
  const uint32_t cur0 = first_t0[g], cur1 = first_t1[g];
#pragma unroll
  for (int j = 0; j < kRegFanin; ++j) fb[j] = (uint32_t)j < r.nfi ? bt[r.fin[j]] : 0;
  __builtin_amdgcn_sched_barrier(0);  // every load above is issued before any use
  uint32_t c0 = PDF_NONE, c1 = PDF_NONE;
#pragma unroll
  for (int u = 0; u < TW; ++u) {
    if (!((nzw >> u) & 1)) continue;
    const uint32_t w = w0 + u;
    uint64_t mu = m[u];
    const uint64_t au = acc[u];
    if (!mu) continue;
    if (au) pl.ACC[row + u] = 0;
    if (!r.po && (mu & ~au)) {
      mu &= au;
      pl.M[row + u] = mu;
    }
    if (!mu) continue;
    // the pushes: this node's final mark into ACC of each fanin with a
    // forward mark at this word.  Unlike rev_marked the push is not masked
    // by the fanin's mark word: the fanin keeps (its forward mark) & ACC, so
    // bits outside its mark change nothing, and it clears every ACC word it
    // has a mark-present bit for, which are the only words pushed to.
#pragma unroll
    for (int j = 0; j < kRegFanin; ++j)
      if ((fb[j] >> u) & 1) push_acc<true>(&pl.ACC[(uint64_t)r.fin[j] * pl.pitch + w], mu);
    for (uint32_t j = kRegFanin; j < r.nfi; ++j) {
      const uint32_t f = net.fin[r.fin_off + j];
      if ((bt[f] >> u) & 1) push_acc<true>(&pl.ACC[(uint64_t)f * pl.pitch + w], mu);
    }
    // the merge (mergekernel.cuh:141-197): marked implies a transition, and
    // T1 means frame 2 set, so mu & mb[u] (mb[u] = forward mark & frame 2)
    // are the T1 pairs of the final mark
    const uint64_t mm = mu & merge_mask(pl, w);
    const uint64_t t0 = mm & ~mb[u], t1 = mm & mb[u];
    const uint32_t base = start + 64u * w;
    if (t0 && c0 == PDF_NONE) c0 = base + (uint32_t)__builtin_ctzll(t0);
    if (t1 && c1 == PDF_NONE) c1 = base + (uint32_t)__builtin_ctzll(t1);
    const uint32_t pos = __hip_atomic_fetch_add(&cnt[u], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    list[(uint64_t)w * list_stride + pos] = g;
  }
  // the first-detection pairs only decrease within a grade, so a value read
  // above is at least the current one: a candidate not below it is skipped
  if (c0 < cur0) atomicMin(&first_t0[g], c0);
  if (c1 < cur1) atomicMin(&first_t1[g], c1);
}

// TW words from w0 of a node-major plane row: inside the row (w0 is a
// multiple of TW and the pitch a multiple of 16 words at least the chunk's
// words rounded up).
template <int TW>
__device__ __forceinline__ void load_tile(uint64_t (&dst)[TW], const uint64_t *src) {
  if constexpr (TW == 1) {
    dst[0] = src[0];
  } else {
#pragma unroll
    for (int q = 0; q < TW / 2; ++q) {
      const u64x2 y = reinterpret_cast<const u64x2 *>(src)[q];
      dst[2 * q] = y.x;
      dst[2 * q + 1] = y.y;
    }
  }
}

// What the staged reverse walk reads of one gate a level before the gate's
// level runs: its mark, T1 and accumulator words of the tile, its
// first-detection pairs and its register fanins' mark-present bytes.
// Nothing of it changes before the gate's level except the accumulator
// words that the level right above pushes, and those go to LDS.
template <int TW>
struct StagedGate {
  uint64_t m[TW], mb[TW], acc[TW];
  uint32_t cur0, cur1;
  uint32_t fb[kRegFanin];
};

template <int TW>
__device__ __forceinline__ void rev_item_load(StagedGate<TW> &it, const Planes &pl, const GateRec &r, uint32_t g,
                                              uint32_t w0, const uint8_t *__restrict__ bt,
                                              const uint32_t *__restrict__ first_t0,
                                              const uint32_t *__restrict__ first_t1) {
  const uint64_t row = (uint64_t)g * pl.pitch + w0;
  load_tile<TW>(it.m, pl.M + row);
  load_tile<TW>(it.mb, pl.V + row);
  load_tile<TW>(it.acc, pl.ACC + row);
  it.cur0 = first_t0[g];
  it.cur1 = first_t1[g];
#pragma unroll
  for (int j = 0; j < kRegFanin; ++j) it.fb[j] = (uint32_t)j < r.nfi ? bt[r.fin[j]] : 0;
}

// Removal and merge of gate g (level lo_cur .. ) over the marked words of
// its tile (bit u of nzw), from its staged item and the accumulator words
// the level above pushed into LDS (accl).  Pushes into fanins of the next
// level (ids nx_lo .. nx_lo + nx_n) go to that level's LDS accumulator
// (nx_acc, TW words a node, unmasked: the fanin keeps its forward mark &
// ACC, so bits outside its mark change nothing, and the buffer is zeroed
// by its readers); pushes into older fanins go to the ACC plane, and only
// to words where the fanin has a forward mark (the plane is all zero
// between passes: each reader clears the words it has a mark-present bit
// for).  Otherwise as walk_rev_gate.
template <bool SEQ, int TW>
__device__ __forceinline__ void walk_rev_apply(const DevNet &net, const GateRec &r, const Planes &pl, uint32_t g,
                                               uint32_t w0, uint64_t nzw, const StagedGate<TW> &it,
                                               const uint64_t (&accl)[TW], lds_u64 *nx_acc, uint32_t nx_lo,
                                               uint32_t nx_n, const uint8_t *__restrict__ bt, uint32_t start,
                                               uint32_t *__restrict__ first_t0, uint32_t *__restrict__ first_t1,
                                               uint32_t *__restrict__ list, uint64_t list_stride, lds_u32 *cnt) {
  const uint64_t row = (uint64_t)g * pl.pitch + w0;
  uint32_t c0 = PDF_NONE, c1 = PDF_NONE;
#pragma unroll
  for (int u = 0; u < TW; ++u) {
    if (!((nzw >> u) & 1)) continue;
    const uint32_t w = w0 + u;
    uint64_t mu = it.m[u];
    const uint64_t au = it.acc[u] | accl[u];
    if (!mu) continue;
    if (it.acc[u]) pl.ACC[row + u] = 0;
    if (!r.po && (mu & ~au)) {
      mu &= au;
      pl.M[row + u] = mu;
    }
    if (!mu) continue;
#pragma unroll
    for (int j = 0; j < kRegFanin; ++j) {
      if ((uint32_t)j >= r.nfi) continue;
      const uint32_t f = r.fin[j];
      if (f - nx_lo < nx_n)
        __hip_atomic_fetch_or(&nx_acc[(f - nx_lo) * TW + u], mu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      else if ((it.fb[j] >> u) & 1)
        push_acc<true>(&pl.ACC[(uint64_t)f * pl.pitch + w], mu);
    }
    for (uint32_t j = kRegFanin; j < r.nfi; ++j) {
      const uint32_t f = net.fin[r.fin_off + j];
      if (f - nx_lo < nx_n)
        __hip_atomic_fetch_or(&nx_acc[(f - nx_lo) * TW + u], mu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      else if ((bt[f] >> u) & 1)
        push_acc<true>(&pl.ACC[(uint64_t)f * pl.pitch + w], mu);
    }
    const uint64_t mm = mu & merge_mask(pl, w);
    const uint64_t t0 = mm & ~it.mb[u], t1 = mm & it.mb[u];
    const uint32_t base = start + 64u * w;
    if (t0 && c0 == PDF_NONE) c0 = base + (uint32_t)__builtin_ctzll(t0);
    if (t1 && c1 == PDF_NONE) c1 = base + (uint32_t)__builtin_ctzll(t1);
    const uint32_t pos = __hip_atomic_fetch_add(&cnt[u], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    list[(uint64_t)w * list_stride + pos] = g;
  }
  if (c0 < it.cur0) atomicMin(&first_t0[g], c0);
  if (c1 < it.cur1) atomicMin(&first_t1[g], c1);
}

// The reverse walk staged through registers and LDS (every level of the
// netlist has at most nx_cap nodes): what a gate reads from HBM is loaded
// while the level above it runs (StagedGate; its record and bits two levels
// ahead), and the level above pushes into the gate's accumulator in LDS
// (two buffers of nx_cap nodes x TW words, by level parity), so no level
// waits for a load from HBM.  Lane i takes node lo + i of each level.
template <bool SEQ, int TW>
__device__ __forceinline__ void walk_rev_staged(const DevNet &net, const GateRec *__restrict__ recs,
                                                const Planes &pl, const uint8_t *__restrict__ BT,
                                                const uint32_t *__restrict__ level_off, uint32_t levels,
                                                uint32_t tiles, uint32_t start, uint32_t *__restrict__ first_t0,
                                                uint32_t *__restrict__ first_t1, uint32_t *__restrict__ list,
                                                uint32_t *__restrict__ list_cnt, uint64_t list_stride,
                                                lds_u64 *accbuf, uint32_t nx_cap, lds_u32 *cnt) {
  uint32_t tile, w0, nw;
  if (!walk_tile<TW>(pl, tiles, tile, w0, nw)) return;
  if (threadIdx.x < TW) cnt[threadIdx.x] = 0;
  const uint8_t *bt = BT + (uint64_t)tile * net.n;
  for (uint32_t i = threadIdx.x; i < 2 * nx_cap * TW; i += kWalkBlock) accbuf[i] = 0;
  const uint32_t tid = threadIdx.x;
  // A: the lane's gate of the current level (record, bits, item); B: its
  // gate of the next level (record, bits)
  GateRec rA{}, rB{};
  uint32_t bA = 0, bB = 0;
  bool vB = false;  // the lane has a gate in the next level (bB is its bits)
  StagedGate<TW> itA{};
  {
    const uint32_t lo = level_off[levels - 1], hi = level_off[levels];
    if (lo + tid < hi) {
      rA = recs[lo + tid];
      bA = bt[lo + tid];  // (waited for here, once)
      if (bA) rev_item_load<TW>(itA, pl, rA, lo + tid, w0, bt, first_t0, first_t1);
    }
    if (levels >= 2) {
      const uint32_t lo1 = level_off[levels - 2];
      const uint32_t gs = lo1 + tid < lo ? lo1 + tid : lo1;
      rB = recs[gs];
      bB = bt[gs];
      vB = lo1 + tid < lo;
    }
  }
  __syncthreads();  // the buffers are zero
  for (uint32_t k = levels; k-- > 0;) {
    const uint32_t lo = level_off[k], hi = level_off[k + 1];
    const uint32_t nx_lo = k > 0 ? level_off[k - 1] : 0, nx_n = k > 0 ? lo - nx_lo : 0;
    lds_u64 *own = accbuf + (k & 1) * nx_cap * TW;
    lds_u64 *nx_acc = accbuf + ((k + 1) & 1) * nx_cap * TW;
    // the next level's items (its bits came with the previous level) and
    // the level after's records and bits: issued now, used after the barrier
    // (validity flags are applied when the values are used, after the
    // barrier, so no wait of this level covers the loads)
    StagedGate<TW> itC{};
    if (k > 0 && vB && bB) rev_item_load<TW>(itC, pl, rB, nx_lo + tid, w0, bt, first_t0, first_t1);
    GateRec rD{};
    uint32_t bD = 0;
    bool vD = false;
    if (k > 1) {
      const uint32_t lo2 = level_off[k - 2];
      const uint32_t gs = lo2 + tid < nx_lo ? lo2 + tid : lo2;
      rD = recs[gs];
      bD = bt[gs];
      vD = lo2 + tid < nx_lo;
    }
    uint32_t i = lo + tid;
    if (i < hi) {
      uint64_t accl[TW];
#pragma unroll
      for (int u = 0; u < TW; ++u) {
        accl[u] = own[(i - lo) * TW + u];
        own[(i - lo) * TW + u] = 0;
      }
      if (bA)
        walk_rev_apply<SEQ, TW>(net, rA, pl, i, w0, bA, itA, accl, nx_acc, nx_lo, nx_n, bt, start, first_t0,
                                first_t1, list, list_stride, cnt);
      for (i += kWalkBlock; i < hi; i += kWalkBlock) {
        // levels wider than the block: loaded here, not ahead
#pragma unroll
        for (int u = 0; u < TW; ++u) {
          accl[u] = own[(i - lo) * TW + u];
          own[(i - lo) * TW + u] = 0;
        }
        const uint32_t b = bt[i];
        if (!b) continue;
        const GateRec r = recs[i];
        StagedGate<TW> it;
        rev_item_load<TW>(it, pl, r, i, w0, bt, first_t0, first_t1);
        walk_rev_apply<SEQ, TW>(net, r, pl, i, w0, b, it, accl, nx_acc, nx_lo, nx_n, bt, start, first_t0,
                                first_t1, list, list_stride, cnt);
      }
    }
    __syncthreads();  // pushes into the next levels' accumulators are visible, own buffer zero
    rA = rB;
    bA = vB ? bB : 0;
    itA = itC;
    rB = rD;
    bB = bD;
    vB = vD;
  }
  if (threadIdx.x < nw) list_cnt[w0 + threadIdx.x] = cnt[threadIdx.x];
}

// Reverse walk: levels levels-1..0 of one tile, removal + merge + the
// per-word lists of marked nodes (rev_marked on every marked (gate, word) of
// the tile, frames from VT).  A node's fanouts are on higher levels, so they
// have pushed into its accumulator before its level runs; its fanins are on
// lower levels, processed after the barrier.
// NM: after per-level forward launches instead of the forward walk, so the
// mark-present bits come from NZ and the frames from V / VB (node-major).
template <bool SEQ, int TW, bool NM>
__device__ __forceinline__ void walk_rev_body(const DevNet &net, const GateRec *__restrict__ recs, const Planes &pl,
                                              const uint8_t *__restrict__ BT, const uint32_t *__restrict__ level_off,
                                              uint32_t levels, uint32_t tiles, uint32_t start,
                                              uint32_t *__restrict__ first_t0, uint32_t *__restrict__ first_t1,
                                              uint32_t *__restrict__ list, uint32_t *__restrict__ list_cnt,
                                              uint64_t list_stride, lds_u32 *cnt) {
  using S = WalkShape<SEQ, TW>;
  uint32_t tile, w0, nw;
  if (!walk_tile<TW>(pl, tiles, tile, w0, nw)) return;
  if (threadIdx.x < TW) cnt[threadIdx.x] = 0;
  const uint8_t *bt = NM ? nullptr : BT + (uint64_t)tile * net.n;
  const uint64_t tmask = (1ULL << nw) - 1;  // nw <= 8
  auto bits_of = [&](uint32_t g) -> uint64_t {  // mark-present bits of the tile, bit u: word w0 + u
    if (NM) return (pl.NZ[(uint64_t)g * pl.ntiles + (w0 >> 6)] >> (w0 & 63)) & tmask;
    return bt[g];
  };
  // each lane's first gate of a level: its mark-present bits are loaded one
  // level ahead, and its netlist record too when a bit is set (neither
  // changes in this pass)
  // (the walk reads a gate's words in one round trip; NM reads node-major
  // planes word by word)
  uint64_t bn = 0;
  bool bn_ok = false;
  GateRec rn{};
  {
    const uint32_t g = level_off[levels - 1] + threadIdx.x;
    if (g < level_off[levels]) {
      bn = bits_of(g);
      bn_ok = true;
      if (NM) {
        if (bn) rn = recs[g];
      } else {
        rn = recs[g];
      }
    }
  }
  __syncthreads();
  for (uint32_t k = levels; k-- > 0;) {
    const uint32_t lo = level_off[k], hi = level_off[k + 1];
    const uint64_t b0 = bn_ok ? bn : 0;
    const GateRec r0 = rn;
    const uint32_t gn = k > 0 ? level_off[k - 1] + threadIdx.x : 0;
    const bool pre = k > 0 && gn < lo;
    if (!NM) {
      // the next level's bits and record, read unconditionally from a valid
      // node (lanes past that level's end read one of this level, never
      // used) into registers nothing reads before the next level: no wait
      // of this level covers them
      const uint32_t gs = pre ? gn : lo;
      bn = bits_of(gs);
      rn = recs[gs];
      bn_ok = pre;
    } else {
      bn = pre ? bits_of(gn) : 0;
      bn_ok = true;
    }
    for (uint32_t g = lo + threadIdx.x; g < hi; g += kWalkBlock) {
      const bool first = g == lo + threadIdx.x;
      uint64_t nzw = first ? b0 : bits_of(g);
      if (!nzw) continue;
      const GateRec r = first ? r0 : recs[g];
      if constexpr (!NM) {
        walk_rev_gate<SEQ, TW>(net, r, pl, g, w0, nzw, bt, start, first_t0, first_t1, list, list_stride, cnt);
      } else {
        while (nzw) {
          const uint32_t w = w0 + (uint32_t)__builtin_ctzll(nzw);
          nzw &= nzw - 1;
          const uint64_t *fa = pl.V + (uint64_t)g * pl.pitch + w;
          const uint64_t *fb = SEQ ? fa + 1 : pl.VB + (uint64_t)g * pl.pitch + w;
          rev_marked<SEQ, true, true, true>(net, r, pl, g, w, start, first_t0, first_t1, list, list_cnt,
                                            list_stride, fa, fb, cnt, w0, nullptr, S::NV);
        }
      }
    }
    if (NM && bn) rn = recs[gn];
    __syncthreads();  // pushes into lower levels' accumulators are visible
  }
  if (threadIdx.x < nw) list_cnt[w0 + threadIdx.x] = cnt[threadIdx.x];
}

// The reverse walk alone (after per-level forward launches: NM).
template <bool SEQ, int TW, bool NM>
__global__ __launch_bounds__(kWalkBlock) void k_walk_rev(DevNet net, const GateRec *__restrict__ recs, Planes pl,
                                                          const uint32_t *__restrict__ level_off, uint32_t levels,
                                                          uint32_t tiles, uint32_t start,
                                                          uint32_t *__restrict__ first_t0,
                                                          uint32_t *__restrict__ first_t1,
                                                          uint32_t *__restrict__ list, uint32_t *__restrict__ list_cnt,
                                                          uint64_t list_stride) {
  __shared__ uint32_t cnt_words[TW];  // the tile's per-word list lengths
  walk_rev_body<SEQ, TW, NM>(net, recs, pl, nullptr, level_off, levels, tiles, start, first_t0, first_t1, list,
                             list_cnt, list_stride, (lds_u32 *)cnt_words);
}

// Forward and reverse walk of a tile in one launch: a tile's reverse pass
// needs only its own forward pass, so no workgroup waits for another.
template <bool SEQ, bool ROBUST, int TW>
__global__ __launch_bounds__(kWalkBlock) void k_walk(DevNet net, const GateRec *__restrict__ recs,
                                                      const GateRec *__restrict__ wrecs, Planes pl,
                                                      const uint64_t *__restrict__ va, const uint64_t *__restrict__ vb,
                                                      uint64_t vstride, uint64_t wbase, uint64_t *__restrict__ VT,
                                                      uint8_t *__restrict__ BT, uint32_t all_records,
                                                      const uint32_t *__restrict__ level_off, uint32_t levels,
                                                      uint32_t widest, uint32_t tiles,
                                                      uint64_t *__restrict__ actmap,
                                                      uint32_t *__restrict__ act_count,
                                                      uint64_t *__restrict__ chunk_total, uint32_t start,
                                                      uint32_t *__restrict__ first_t0, uint32_t *__restrict__ first_t1,
                                                      uint32_t *__restrict__ list, uint32_t *__restrict__ list_cnt,
                                                      uint64_t list_stride) {
  using S = WalkShape<SEQ, TW>;
  __shared__ uint64_t lds_words[S::cap * S::SV];
  static_assert(S::cap * S::SV >= 8 + 2 * TW, "room for the reverse walk's list lengths");
  {
    // the chunk's event and cover counters, each cleared by the block that
    // owns it (blocks are not ordered within the launch, so no block may
    // touch another tile's words): the pair bitmap's words of this tile,
    // and the pair count and chunk total by the block of tile 0; the list
    // lengths are written by their tiles at the end
    uint32_t t, w0, nw;
    if (walk_tile<TW>(pl, tiles, t, w0, nw)) {
      if (threadIdx.x < nw) actmap[w0 + threadIdx.x] = 0;
      if (t == 0 && threadIdx.x == 0) {
        *act_count = 0;
        *chunk_total = 0;
      }
    }
  }
  walk_fwd_body<SEQ, ROBUST, TW>(net, wrecs, pl, va, vb, vstride, wbase, VT, BT, all_records != 0, level_off, levels,
                                 tiles, (lds_u64 *)lds_words);
  __syncthreads();  // the forward walk's marks and bits are visible to the reverse walk
  // the reverse walk stages through the forward walk's LDS when two levels'
  // accumulators fit (the tile's per-word list lengths in the last 8 words)
  constexpr uint32_t nx_cap = (S::cap * S::SV - 8) / (2 * TW);
  lds_u32 *cnt = (lds_u32 *)((lds_u64 *)lds_words + S::cap * S::SV - 8);
  if (widest <= nx_cap)
    walk_rev_staged<SEQ, TW>(net, recs, pl, BT, level_off, levels, tiles, start, first_t0, first_t1, list, list_cnt,
                             list_stride, (lds_u64 *)lds_words, nx_cap, cnt);
  else
    walk_rev_body<SEQ, TW, false>(net, recs, pl, BT, level_off, levels, tiles, start, first_t0, first_t1, list,
                                  list_cnt, list_stride, cnt);
}

// Cross-chunk state to its initial value (one launch per grade).
__global__ void k_reset_grade(uint32_t n, uint32_t *first_t0, uint32_t *first_t1, uint64_t *gate_counts,
                              uint64_t *grand_total, uint32_t *err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    first_t0[i] = PDF_NONE;
    first_t1[i] = PDF_NONE;
    gate_counts[i] = 0;
  }
  if (i == 0) {
    *grand_total = 0;
    *err = 0;
  }
}

// Per-chunk merge/cover counters to zero (when no forward launch clears them).
__global__ void k_reset_chunk(uint32_t *state, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) state[i] = 0;
}

// Distinct first-detection patterns of this chunk (relative index): one
// node's first-detection pairs of the chunk into the pair bitmap, each
// pair appended to act_list once (by the lane that set its bit).
// WAVE (every lane of the wave reads the same bitmap and list: k_events):
// one atomic per distinct pair of the wave and one list reservation per
// wave; otherwise (a batch wave spans circuits) one of each per node.
template <bool WAVE>
__device__ __forceinline__ void events_one(uint32_t g, const uint32_t *__restrict__ first_t0,
                                           const uint32_t *__restrict__ first_t1, uint32_t start, uint32_t width,
                                           uint64_t *__restrict__ actmap, uint32_t *__restrict__ act_list,
                                           uint32_t *__restrict__ act_count) {
  const uint32_t x[2] = {first_t0[g], first_t1[g]};
  if (!WAVE) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (x[k] == PDF_NONE || x[k] < start || x[k] - start >= width) continue;
      const uint32_t rel = x[k] - start;
      const uint64_t bit = 1ULL << (rel & 63);
      const uint64_t old = atomicOr((unsigned long long *)&actmap[rel >> 6], (unsigned long long)bit);
      if (!(old & bit)) act_list[atomicAdd(act_count, 1u)] = rel;
    }
    return;
  }
  const uint32_t lane = threadIdx.x & 63;
  // one atomic per distinct pair of the wave (most nodes share a few first
  // pairs, and atomics on one word queue): the lowest lane holding a pair
  // leads it, the leaders' atomics go out together
  bool lead[2] = {false, false};
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const bool want = x[k] != PDF_NONE && x[k] >= start && x[k] - start < width;
    const uint32_t rel = x[k] - start;
    uint64_t todo = __ballot(want);
    while (todo) {
      const int leader = (int)__builtin_ctzll(todo);
      const uint32_t v = __shfl(rel, leader);
      todo &= ~__ballot(want && rel == v);
      if ((int)lane == leader) lead[k] = true;
    }
  }
  // a bit some wave already set (most of them, for the common pairs) needs
  // no atomic: read the words first (a stale read only costs the atomic)
  uint64_t old[2] = {0, 0};
#pragma unroll
  for (int k = 0; k < 2; ++k)
    if (lead[k]) old[k] = __hip_atomic_load(&actmap[(x[k] - start) >> 6], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const uint64_t bit = 1ULL << ((x[k] - start) & 63);
    if (lead[k] && !(old[k] & bit))
      old[k] = atomicOr((unsigned long long *)&actmap[(x[k] - start) >> 6], (unsigned long long)bit);
  }
  bool fresh[2];  // this lane set the pair's bit: it appends the pair
#pragma unroll
  for (int k = 0; k < 2; ++k) fresh[k] = lead[k] && !((old[k] >> ((x[k] - start) & 63)) & 1);
  // the wave's new pairs appended with one reservation
  const uint64_t b0 = __ballot(fresh[0]), b1 = __ballot(fresh[1]);
  const uint32_t n0 = (uint32_t)__builtin_popcountll(b0), cnt = n0 + (uint32_t)__builtin_popcountll(b1);
  if (!cnt) return;
  const int first = (int)__builtin_ctzll(b0 | b1);
  uint32_t base = 0;
  if ((int)lane == first) base = atomicAdd(act_count, cnt);
  base = __shfl(base, first);
  const uint64_t below = (1ULL << lane) - 1;
  if (fresh[0]) act_list[base + (uint32_t)__builtin_popcountll(b0 & below)] = x[0] - start;
  if (fresh[1]) act_list[base + n0 + (uint32_t)__builtin_popcountll(b1 & below)] = x[1] - start;
}

__global__ void k_events(uint32_t n, const uint32_t *__restrict__ first_t0, const uint32_t *__restrict__ first_t1,
                         uint32_t start, uint32_t width, uint64_t *__restrict__ actmap,
                         uint32_t *__restrict__ act_list, uint32_t *__restrict__ act_count) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) events_one<true>(g, first_t0, first_t1, start, width, actmap, act_list, act_count);
}

__device__ __forceinline__ bool add_ovf(uint64_t &acc, uint64_t v) {
  const uint64_t s = acc + v;
  const bool o = s < acc;
  acc = s;
  return o;
}

// 1,280 nodes (73.5 KB of LDS a wave with the hash and the pool: two waves
// per CU) hold the ~970 marked nodes of a pair of the cover-heavy input
constexpr uint32_t kCoverCap = 1280;   // nodes marked at one pair kept in LDS
constexpr uint32_t kHashSlots = 2048;  // open-addressing node -> LDS index
constexpr uint32_t kPool = 4096;       // fanout ids of those nodes kept in LDS
constexpr uint32_t kEmpty = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t hash_slot(uint32_t g) { return (g * 2654435761u) >> 21; }  // 11 bits

// How much of the cover DP stays in LDS: at most `nodes` marked nodes of a
// pair (else the DP keeps c/h in global scratch) and `pool` fanout ids (the
// rest are read from the netlist in the DP).  The defaults are the LDS
// sizes; smaller values (a tuning knob for tests) send more pairs down the
// global-memory paths.
struct CoverCaps {
  uint32_t nodes = kCoverCap, pool = kPool;
};

// Reverse path-count DP (coverkernel.cuh:194-228) for each distinct
// first-detection pair p of the chunk, over the nodes marked at p.  One wave
// per pair.  The word's node list (descending level order) is filtered to the
// nodes marked at p; their records, fanout ids and c/h live in LDS with a
// hash from node id to LDS index.  Runs of equal level are then processed
// lane-parallel, each node summing the c/h of its fanouts found in the hash
// (a fanout marked at p has a higher level, so it was finished in an earlier
// run).  If more than kCoverCap nodes are marked at p the DP keeps c/h in
// this wave's slice of cs/hs and reads everything from global memory.
// ALL_NEW: every pair of the chunk, every marked node counted as new (an
// empty detection history), so a pair's PI total is its number of marked
// PI->PO paths; written to pair_gain instead of the running totals.  In the
// normal mode a non-null pair_gain receives each processed pair's total
// (saturated to uint32; the caller zeroes it for the pairs not processed).
struct CoverLds {
  uint32_t s_node[kCoverCap], s_level[kCoverCap], s_pool_off[kCoverCap], s_fo_off[kCoverCap];
  uint64_t s_c[kCoverCap], s_h[kCoverCap];
  uint32_t h_key[kHashSlots];
  uint32_t s_pool[kPool];
  uint16_t s_nfo[kCoverCap];
  uint16_t h_val[kHashSlots];
  uint8_t s_flags[kCoverCap];  // bit0 po, bit1 INPT, bit2 new
};

// The DP of one pair (relative index rel in the chunk) by one wave.
// SEGNEW (segments of length K >= 2): a node is new at p when its bit in
// `newp` ([n][pl.pitch] words, built by k_seg_events) is set, instead of
// when its own first T0 / T1 is p.
template <bool ALL_NEW, bool SEGNEW = false>
__device__ void cover_one(CoverLds &S, const DevNet &net, const Planes &pl, uint32_t start,
                          const uint32_t *__restrict__ list, const uint32_t *__restrict__ list_cnt,
                          uint64_t list_stride, const uint32_t *__restrict__ first_t0,
                          const uint32_t *__restrict__ first_t1, uint64_t *c_s, uint64_t *h_s,
                          uint64_t *__restrict__ gate_counts, uint64_t *__restrict__ chunk_total,
                          uint64_t *__restrict__ grand_total, uint32_t *__restrict__ err,
                          uint32_t *__restrict__ pair_gain, uint32_t rel,
                          const uint64_t *__restrict__ newp = nullptr, uint64_t *__restrict__ pi_cover = nullptr,
                          uint64_t pi_ld = 0, CoverCaps caps = CoverCaps{}) {
  const uint32_t lane = threadIdx.x;
  const uint64_t below = (1ULL << lane) - 1;
  const uint32_t w = rel >> 6, j = rel & 63;
  const uint32_t p = start + rel;
  const uint32_t len = list_cnt[w];
  auto is_new = [&](uint32_t g) -> bool {
    if (ALL_NEW) return true;
    if (SEGNEW) return (newp[(uint64_t)g * pl.pitch + w] >> j) & 1;
    return first_t0[g] == p || first_t1[g] == p;
  };
  const uint32_t *L = list + (uint64_t)w * list_stride;
  // 1. the nodes marked at p, in list order: kScan entries a lane per round,
  // each round's list ids, then mark-present bits, then mark words loaded
  // together (three round trips per 64 kScan entries; a word's list holds
  // thousands of nodes where many paths are marked)
  constexpr int kScan = 8;
  for (uint32_t t = lane; t < kHashSlots; t += 64) S.h_key[t] = kEmpty;
  uint32_t K = 0;
  for (uint32_t base = 0; base < len; base += 64 * kScan) {
    uint32_t gs[kScan], pr[kScan];
    uint64_t mw[kScan];
#pragma unroll
    for (int u = 0; u < kScan; ++u) {
      const uint32_t idx = base + 64 * u + lane;
      gs[u] = idx < len ? L[idx] : kEmpty;
    }
#pragma unroll
    for (int u = 0; u < kScan; ++u) hold(gs[u]);
#pragma unroll
    for (int u = 0; u < kScan; ++u) pr[u] = gs[u] != kEmpty ? mark_present(pl, gs[u], w) : 0;
#pragma unroll
    for (int u = 0; u < kScan; ++u) hold(pr[u]);
#pragma unroll
    for (int u = 0; u < kScan; ++u) mw[u] = (pr[u] & 1) ? pl.M[(uint64_t)gs[u] * pl.pitch + w] : 0;
#pragma unroll
    for (int u = 0; u < kScan; ++u) hold(mw[u]);
#pragma unroll
    for (int u = 0; u < kScan; ++u) {
      const bool marked = (mw[u] >> j) & 1;
      const uint64_t bal = __ballot(marked);
      const uint32_t pos = K + (uint32_t)__builtin_popcountll(bal & below);
      if (marked && pos < kCoverCap) S.s_node[pos] = gs[u];
      K += (uint32_t)__builtin_popcountll(bal);
    }
  }
  __syncthreads();
  const bool in_lds = K <= caps.nodes;
  uint64_t tot = 0;
  bool ovf = false, ref16 = false;
  if (in_lds) {
    // 2. records, hash, and the fanout-id pool offsets: kScan nodes a lane
    // per round, each round's record tails (fout_off, nfi/nfo, type/po),
    // levels and first-detection words loaded together
    uint32_t carry = 0;
    for (uint32_t base = 0; base < K; base += 64 * kScan) {
      uint32_t gs[kScan], lv[kScan], n0[kScan], n1[kScan];
      uint64_t nwd[kScan];
      uint4 tail[kScan];  // bytes 16..31 of the record: fin_off, fout_off, nfi | nfo << 16, type | po << 8
#pragma unroll
      for (int u = 0; u < kScan; ++u) {
        const uint32_t idx = base + 64 * u + lane;
        gs[u] = idx < K ? S.s_node[idx] : kEmpty;
        lv[u] = n0[u] = n1[u] = 0;
        nwd[u] = 0;
        tail[u] = uint4{0, 0, 0, 0};
        if (gs[u] != kEmpty) {
          tail[u] = reinterpret_cast<const uint4 *>(net.rec + gs[u])[1];
          lv[u] = net.level[gs[u]];
          if (SEGNEW) nwd[u] = newp[(uint64_t)gs[u] * pl.pitch + w];
          else if (!ALL_NEW) {
            n0[u] = first_t0[gs[u]];
            n1[u] = first_t1[gs[u]];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < kScan; ++u) {
        hold(lv[u]);
        hold(n0[u]);
        hold(n1[u]);
        hold(nwd[u]);
      }
#pragma unroll
      for (int u = 0; u < kScan; ++u) {
        const uint32_t idx = base + 64 * u + lane;
        uint32_t nfo = 0;
        if (gs[u] != kEmpty) {
          const uint32_t g = gs[u];
          nfo = tail[u].z >> 16;
          const uint32_t type = tail[u].w & 0xFF, po = (tail[u].w >> 8) & 0xFF;
          const bool fresh = ALL_NEW || (SEGNEW ? ((nwd[u] >> j) & 1) != 0 : (n0[u] == p || n1[u] == p));
          S.s_level[idx] = lv[u];
          S.s_nfo[idx] = (uint16_t)nfo;
          S.s_fo_off[idx] = tail[u].y;
          S.s_flags[idx] = (uint8_t)((po == 1 ? 1 : 0) | (type == PDF_INPT ? 2 : 0) | (fresh ? 4 : 0));
          uint32_t h = hash_slot(g);
          while (atomicCAS(&S.h_key[h], kEmpty, g) != kEmpty) h = (h + 1) & (kHashSlots - 1);
          S.h_val[h] = (uint16_t)idx;
        }
        uint32_t incl = nfo;  // inclusive wave scan of the fanout counts
        for (int off = 1; off < 64; off <<= 1) {
          const uint32_t o = __shfl_up(incl, off);
          if (lane >= (uint32_t)off) incl += o;
        }
        if (gs[u] != kEmpty) S.s_pool_off[idx] = carry + incl - nfo;
        carry += __shfl(incl, 63);
      }
    }
    __syncthreads();
    // 3. fanout ids into the pool, one pool slot per lane and eight loads in
    // flight (nodes past the pool's end read theirs in the DP)
    const uint32_t F = carry < caps.pool ? carry : caps.pool;
    for (uint32_t base = 0; base < F; base += 8 * 64) {
      uint32_t ids[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const uint32_t slot = base + u * 64 + lane;
        ids[u] = 0;
        if (slot < F) {
          uint32_t lo = 0, hi = K - 1;  // last node whose pool range starts at or before slot
          while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (S.s_pool_off[mid] <= slot) lo = mid;
            else hi = mid - 1;
          }
          ids[u] = net.fout[S.s_fo_off[lo] + (slot - S.s_pool_off[lo])];
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const uint32_t slot = base + u * 64 + lane;
        if (slot < F) S.s_pool[slot] = ids[u];
      }
    }
    __syncthreads();
    // 4. the DP over runs of equal level, in LDS.  A run's pooled nodes (a
    // prefix of it: pool offsets grow along the list) have their fanout ids
    // in one contiguous pool range; a lane takes one fanout edge of that
    // range at a time (its node by binary search over the run's offsets),
    // looks the fanout up in the hash and adds its c / h into the node's
    // with LDS atomics, so a run costs its edge count over 64 rounds instead
    // of its largest fanout count.  Nodes past the pool sum their fanouts
    // from the netlist, one lane each.  Unsigned sums of non-negative terms
    // wrap iff some partial sum does, so each atomic's old value tells.
    for (uint32_t i = 0; i < K;) {
      const uint32_t idx = i + lane;
      const bool valid = idx < K;
      const uint32_t lv = valid ? S.s_level[idx] : 0xFFFFFFFFu;
      const uint32_t lv0 = __shfl(lv, 0);
      const uint64_t same = __ballot(valid && lv == lv0);
      const uint32_t run = ~same ? (uint32_t)__builtin_ctzll(~same) : 64u;  // contiguous prefix
      uint32_t off = 0, nfo = 0;
      uint8_t fl = 0;
      if (lane < run) {
        fl = S.s_flags[idx];
        off = S.s_pool_off[idx];
        nfo = S.s_nfo[idx];
        S.s_c[idx] = 0;       // fanout sums C and H (H from the PO's own path)
        S.s_h[idx] = fl & 1;
      }
      const bool pooled = lane < run && off + nfo <= caps.pool;
      const uint32_t rp = (uint32_t)__builtin_popcountll(__ballot(pooled));  // pooled prefix of the run
      const uint32_t e0 = __shfl(off, 0);
      const uint32_t e1 = rp ? __shfl(off + nfo, (int)rp - 1) : e0;
      __syncthreads();
      for (uint32_t e = e0 + lane; e < e1; e += 64) {
        uint32_t lo = i, hi = i + rp - 1;  // the last pooled node whose range starts at or before e
        while (lo < hi) {
          const uint32_t mid = (lo + hi + 1) >> 1;
          if (S.s_pool_off[mid] <= e) lo = mid;
          else hi = mid - 1;
        }
        const uint32_t fo = S.s_pool[e];
        uint32_t h = hash_slot(fo);
        uint32_t key;
        while ((key = S.h_key[h]) != kEmpty && key != fo) h = (h + 1) & (kHashSlots - 1);
        if (key == fo) {
          const uint32_t at = S.h_val[h];
          const uint64_t vc = S.s_c[at], vh = S.s_h[at];
          if (vc) {
            const uint64_t o = atomicAdd((unsigned long long *)&S.s_c[lo], (unsigned long long)vc);
            ovf |= o + vc < o;
          }
          if (vh) {
            const uint64_t o = atomicAdd((unsigned long long *)&S.s_h[lo], (unsigned long long)vh);
            ovf |= o + vh < o;
          }
        }
      }
      if (lane < run && !pooled) {
        const uint32_t fo_off = S.s_fo_off[idx];
        uint64_t C = 0, H = fl & 1;
        for (uint32_t e = 0; e < nfo; ++e) {
          const uint32_t fo = net.fout[fo_off + e];
          uint32_t h = hash_slot(fo);
          uint32_t key;
          while ((key = S.h_key[h]) != kEmpty && key != fo) h = (h + 1) & (kHashSlots - 1);
          if (key == fo) {
            ovf |= add_ovf(C, S.s_c[S.h_val[h]]);
            ovf |= add_ovf(H, S.s_h[S.h_val[h]]);
          }
        }
        S.s_c[idx] = C;
        S.s_h[idx] = H;
      }
      __syncthreads();
      if (lane < run) {
        const uint64_t C = S.s_c[idx], H = S.s_h[idx];
        uint64_t c = C, h = H;
        if (fl & 4) { ovf |= add_ovf(c, H); h = 0; }
        if (fl & 2) h = 0;
        // the reference's 16-bit fields: its fanout sums and final c / h
        // (coverkernel.cuh:210-211, 224-225); H includes the PO's own path
        ref16 |= C > 0xFFFF || H - (fl & 1) > 0xFFFF || c > 0xFFFF || h > 0xFFFF;
        S.s_c[idx] = c;
        S.s_h[idx] = h;
      }
      __syncthreads();
      i += run;
    }
    // 5. per-gate counts and the PI total, all nodes at once (kScan a lane
    // per round, their atomics issued together)
    for (uint32_t base = 0; base < K; base += 64 * kScan) {
      uint64_t cv[kScan], old[kScan];
#pragma unroll
      for (int u = 0; u < kScan; ++u) {
        const uint32_t idx = base + 64 * u + lane;
        cv[u] = idx < K ? S.s_c[idx] : 0;
        old[u] = 0;
        if (!ALL_NEW && cv[u])
          old[u] = atomicAdd((unsigned long long *)&gate_counts[S.s_node[idx]], (unsigned long long)cv[u]);
      }
#pragma unroll
      for (int u = 0; u < kScan; ++u) {
        const uint32_t idx = base + 64 * u + lane;
        if (!cv[u]) continue;
        if (!ALL_NEW) ovf |= old[u] + cv[u] < old[u];
        if (S.s_flags[idx] & 2) {
          ovf |= add_ovf(tot, cv[u]);
          if (pi_cover) pi_cover[(uint64_t)S.s_node[idx] * pi_ld + rel] = cv[u];
        }
      }
    }
  } else {
    // too many marked nodes for LDS: c/h in global scratch, the whole list
    for (uint32_t i = 0; i < len;) {
      const uint32_t idx = i + lane;
      const bool valid = idx < len;
      const uint32_t g = valid ? L[idx] : 0;
      const uint32_t lv = valid ? net.level[g] : 0xFFFFFFFFu;
      const uint32_t lv0 = __shfl(lv, 0);
      const uint64_t same = __ballot(valid && lv == lv0);
      const uint32_t run = ~same ? (uint32_t)__builtin_ctzll(~same) : 64u;
      if (lane < run && ((mark_of(pl, g, w) >> j) & 1)) {
        const GateRec r = net.rec[g];
        uint64_t C = 0, H = r.po == 1 ? 1 : 0;
        for (uint32_t e = 0; e < r.nfo; ++e) {
          const uint32_t fo = net.fout[r.fout_off + e];
          if ((mark_of(pl, fo, w) >> j) & 1) {
            ovf |= add_ovf(C, c_s[fo]);
            ovf |= add_ovf(H, h_s[fo]);
          }
        }
        uint64_t c = C, h = H;
        if (is_new(g)) { ovf |= add_ovf(c, H); h = 0; }
        if (r.type == PDF_INPT) h = 0;
        ref16 |= C > 0xFFFF || H - (r.po == 1 ? 1 : 0) > 0xFFFF || c > 0xFFFF || h > 0xFFFF;
        c_s[g] = c;
        h_s[g] = h;
        if (c) {
          if (!ALL_NEW) {
            const uint64_t old = atomicAdd((unsigned long long *)&gate_counts[g], (unsigned long long)c);
            ovf |= old + c < old;
          }
          if (r.type == PDF_INPT) {
            ovf |= add_ovf(tot, c);
            if (pi_cover) pi_cover[(uint64_t)g * pi_ld + rel] = c;
          }
        }
      }
      __threadfence_block();
      __syncthreads();
      i += run;
    }
  }
  // wave sum of the PI counts of this pattern (kernSumSingle)
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t o = __shfl_down(tot, off);
    ovf |= add_ovf(tot, o);
  }
  if (pair_gain && lane == 0) pair_gain[p] = tot > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)tot;
  bool ref32 = false;
  if (!ALL_NEW && lane == 0 && tot) {
    const uint64_t o0 = atomicAdd((unsigned long long *)chunk_total, (unsigned long long)tot);
    const uint64_t o1 = atomicAdd((unsigned long long *)grand_total, (unsigned long long)tot);
    ovf |= (o0 + tot < o0) || (o1 + tot < o1);
    // the chunk's running total passing 2^32 - 1: the reference's uint32
    // partial sums would wrap
    ref32 = o0 + tot > 0xFFFFFFFFull;
  }
  const uint32_t e = (__ballot(ovf) ? kErrOverflow : 0u) | (!ALL_NEW && __ballot(ref16) ? kErrRefU16 : 0u) |
                     (__ballot(ref32) ? kErrRefU32 : 0u);
  if (e && lane == 0) atomicOr(err, e);
  __syncthreads();
}

template <bool ALL_NEW, bool SEGNEW = false>
__global__ __launch_bounds__(64) void k_cover(DevNet net, Planes pl, uint32_t start, const uint32_t *__restrict__ list,
                                              const uint32_t *__restrict__ list_cnt, uint64_t list_stride,
                                              const uint32_t *__restrict__ act_list,
                                              const uint32_t *__restrict__ act_count,
                                              const uint32_t *__restrict__ first_t0,
                                              const uint32_t *__restrict__ first_t1, uint64_t *cs, uint64_t *hs,
                                              uint64_t *__restrict__ gate_counts, uint64_t *__restrict__ chunk_total,
                                              uint64_t *__restrict__ grand_total, uint32_t *__restrict__ err,
                                              uint32_t width, uint32_t *__restrict__ pair_gain,
                                              const uint64_t *__restrict__ newp, uint64_t *__restrict__ pi_cover,
                                              uint64_t pi_ld, CoverCaps caps) {
  __shared__ CoverLds S;
  uint64_t *c_s = cs + (uint64_t)blockIdx.x * net.n;
  uint64_t *h_s = hs + (uint64_t)blockIdx.x * net.n;
  const uint32_t nact = ALL_NEW ? width : *act_count;
  for (uint32_t k = blockIdx.x; k < nact; k += gridDim.x)
    cover_one<ALL_NEW, SEGNEW>(S, net, pl, start, list, list_cnt, list_stride, first_t0, first_t1, c_s, h_s,
                               gate_counts, chunk_total, grand_total, err, pair_gain, ALL_NEW ? k : act_list[k], newp,
                               pi_cover, pi_ld, caps);
}

// gpuMergeSegments<K> for K >= 2 (kernSegmentReduce2 with has_marked_segment,
// mergekernel.cuh:91-198): segment s ([K] node ids, PDF_NONE entries
// skipped) is detected at a pair when every node of it is marked there; its
// earliest detection with node 0 falling (T0) and rising (T1) is min-combined
// into sf0 / sf1.
// A wave holds 64 >> lps_log segments, 1 << lps_log lanes each (at least the
// chunk's number of 64-word blocks, or 64).  Phase 1: lane `sub` of a
// segment ANDs the nodes' mark-present words of block `sub` (one round of
// independent loads).  Phase 2: the whole wave visits each (segment, block)
// whose nodes all have marks, in lane order (so a segment's blocks in
// ascending order), lanes over the block's 64 words, and drops a segment's
// remaining blocks once both polarities are found.
template <bool SEQ>
__global__ __launch_bounds__(256) void k_merge_seg(Planes pl, const uint32_t *__restrict__ seg, uint32_t K,
                                                   uint64_t S, uint32_t lps_log, uint32_t start,
                                                   uint32_t *__restrict__ sf0, uint32_t *__restrict__ sf1) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t lps = 1u << lps_log;
  const uint64_t s_base = ((uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) << (6 - lps_log);
  if (s_base >= S) return;  // whole waves
  const uint32_t slot = lane >> lps_log, sub = lane & (lps - 1);
  const uint64_t s = s_base + slot;
  const bool valid = s < S;
  uint32_t ids[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) ids[i] = (valid && (uint32_t)i < K) ? seg[s * K + i] : PDF_NONE;
  const uint64_t slot_mask = lps == 64 ? ~0ULL : ((1ULL << lps) - 1);
  const uint32_t nblk = (pl.words + 63) >> 6;
  uint32_t best0 = PDF_NONE, best1 = PDF_NONE;  // this lane's segment
  for (uint32_t g0 = 0; g0 < nblk; g0 += lps) {
    const uint32_t blk = g0 + sub;
    const bool want = valid && blk < nblk && (best0 == PDF_NONE || best1 == PDF_NONE);
    // the nodes' mark-present words, all issued before they are combined
    uint64_t z[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      z[i] = ~0ULL;
      if (want && ids[i] != PDF_NONE) z[i] = pl.NZ[(uint64_t)ids[i] * pl.ntiles + blk];
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) hold(z[i]);
    uint64_t nzl = 0;
    if (want) {
      nzl = ~0ULL;
#pragma unroll
      for (int i = 0; i < 8; ++i) nzl &= z[i];
    }
    uint64_t todo = __ballot(nzl != 0);
    while (todo) {
      const int l = (int)__builtin_ctzll(todo);
      const uint32_t tslot = (uint32_t)l >> lps_log;
      const uint32_t f0 = __shfl(best0, l), f1 = __shfl(best1, l);
      if (f0 != PDF_NONE && f1 != PDF_NONE) {  // this segment is done
        todo &= ~(slot_mask << (tslot << lps_log));
        continue;
      }
      todo &= todo - 1;
      const uint64_t nz = ((uint64_t)(uint32_t)__shfl((uint32_t)(nzl >> 32), l) << 32) |
                          (uint32_t)__shfl((uint32_t)nzl, l);
      uint32_t gid[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) gid[i] = __builtin_amdgcn_readfirstlane(__shfl(ids[i], l));
      const uint32_t w = (g0 + ((uint32_t)l & (lps - 1))) * 64 + lane;
      // the nodes' mark words and node 0's frames (raw: frame 1 and frame
      // 2's source), all issued before any is used: one round trip a visit
      const bool on = w < pl.words && ((nz >> lane) & 1);
      uint64_t x[8], a = 0, b = 0;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        x[i] = ~0ULL;
        if (on && gid[i] != PDF_NONE) x[i] = pl.M[(uint64_t)gid[i] * pl.pitch + w];
      }
      if (on && gid[0] != PDF_NONE) {
        const uint64_t at = (uint64_t)gid[0] * pl.pitch + w;
        a = pl.V[at];
        b = SEQ ? pl.V[at + 1] : pl.VB[at];
      }
#pragma unroll
      for (int i = 0; i < 8; ++i) hold(x[i]);
      hold(a);
      hold(b);
      if (SEQ) b = shift_next(a, b);
      uint64_t m = 0;
      if (on) {
        m = merge_mask(pl, w);
#pragma unroll
        for (int i = 0; i < 8; ++i) m &= x[i];
      }
      const uint64_t t0 = m & ~b, t1 = m & b;  // marked => transition; b = 1 means T1
      const uint32_t base = start + 64u * w;
      const uint32_t c0 = t0 ? base + (uint32_t)__builtin_ctzll(t0) : PDF_NONE;
      const uint32_t c1 = t1 ? base + (uint32_t)__builtin_ctzll(t1) : PDF_NONE;
      const uint64_t b0 = __ballot(t0 != 0), b1 = __ballot(t1 != 0);
      // the lowest lane holds the lowest word, hence the earliest pair
      const uint32_t e0 = __shfl(c0, b0 ? (int)__builtin_ctzll(b0) : 0);
      const uint32_t e1 = __shfl(c1, b1 ? (int)__builtin_ctzll(b1) : 0);
      if (slot == tslot) {
        if (best0 == PDF_NONE && b0) best0 = e0;
        if (best1 == PDF_NONE && b1) best1 = e1;
      }
    }
  }
  if (valid && sub == 0) {  // atomic: chunks on different lanes merge concurrently
    if (best0 != PDF_NONE) atomicMin(&sf0[s], best0);
    if (best1 != PDF_NONE) atomicMin(&sf1[s], best1);
  }
}

// findActiveThreads<K> (coverkernel.cuh:127-154) for K >= 2: every segment
// whose first T0 or T1 pair lies in this chunk sets that pair's bit for each
// of its nodes in the new-node plane, and the pair joins the chunk's list of
// distinct first-detection pairs (as k_events does for K = 1).
__global__ void k_seg_events(const uint32_t *__restrict__ seg, uint32_t K, uint64_t S,
                             const uint32_t *__restrict__ sf0, const uint32_t *__restrict__ sf1, uint32_t start,
                             uint32_t width, uint64_t *__restrict__ newp, uint64_t pitch,
                             uint64_t *__restrict__ actmap, uint32_t *__restrict__ act_list,
                             uint32_t *__restrict__ act_count) {
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const uint32_t x[2] = {sf0[s], sf1[s]};
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    if (x[k] == PDF_NONE || x[k] < start || x[k] - start >= width) continue;
    const uint32_t rel = x[k] - start;
    const uint64_t bit = 1ULL << (rel & 63);
    for (uint32_t i = 0; i < K; ++i) {
      const uint32_t g = seg[s * K + i];
      if (g != PDF_NONE) atomicOr((unsigned long long *)&newp[(uint64_t)g * pitch + (rel >> 6)], (unsigned long long)bit);
    }
    const uint64_t old = atomicOr((unsigned long long *)&actmap[rel >> 6], (unsigned long long)bit);
    if (!(old & bit)) act_list[atomicAdd(act_count, 1u)] = rel;
  }
}

// The new-node plane back to zero after the cover: the words k_seg_events
// set for this chunk (the segment firsts have not changed since).
__global__ void k_seg_clear(const uint32_t *__restrict__ seg, uint32_t K, uint64_t S,
                            const uint32_t *__restrict__ sf0, const uint32_t *__restrict__ sf1, uint32_t start,
                            uint32_t width, uint64_t *__restrict__ newp, uint64_t pitch) {
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const uint32_t x[2] = {sf0[s], sf1[s]};
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    if (x[k] == PDF_NONE || x[k] < start || x[k] - start >= width) continue;
    const uint32_t rel = x[k] - start;
    for (uint32_t i = 0; i < K; ++i) {
      const uint32_t g = seg[s * K + i];
      if (g != PDF_NONE) newp[(uint64_t)g * pitch + (rel >> 6)] = 0;
    }
  }
}

// n words of p to v (grid-stride).
__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ------------------------------------------------------------ pair features
// 512 features per pair, BF16 (no reference counterpart: inputs of the
// learned pair ranking), stored feature-major: feat[f * ld + pair]:
//   [0, 128)    fraction of the nodes of level f with a transition
//   [128, 256)  fraction of the nodes of level f - 128 marked by the forward
//               pass (robust or non-robust sensitization from a PI, before
//               the reverse removal); both for levels up to the caller's
//               max_level, zero beyond
//   [256, 384)  fraction of the PIs of bucket b = f - 256 that toggle
//               (bucket b = PIs [floor(b I / 128), floor((b + 1) I / 128)))
//   [384, 512)  fanout-weighted share of bucket b = f - 384 that toggles:
//               sum of the fanouts of toggling PIs / sum of the bucket's fanouts
// Each value is float(count) * (1.0f / float(denominator)), rounded to BF16.
constexpr int kFeat = 512;
constexpr int kCountBits = 24;  // wide per-lane counters, up to 2^24 - 1
constexpr int kSmallBits = 6;   // narrow counters flushed every 63 additions

__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (uint16_t)(u >> 16);
}

// 64 per-lane counters held bit-sliced: bit j of c[k] is bit k of counter j.
// vc_inc adds x (one bit per lane) at bit k0; vc_add_sliced adds a
// bit-sliced kSmallBits-wide operand.
template <int B>
__device__ __forceinline__ void vc_inc(uint64_t (&c)[B], uint64_t x, int k0) {
#pragma unroll
  for (int k = 0; k < B; ++k) {
    if (k < k0) continue;
    const uint64_t t = c[k] & x;
    c[k] ^= x;
    x = t;
  }
}
template <int B>
__device__ __forceinline__ void vc_add_sliced(uint64_t (&c)[kCountBits], const uint64_t (&s)[B]) {
  uint64_t carry = 0;
#pragma unroll
  for (int k = 0; k < kCountBits; ++k) {
    const uint64_t y = k < B ? s[k] : 0;
    const uint64_t sum = c[k] ^ y ^ carry;
    carry = (c[k] & y) | (carry & (c[k] ^ y));
    c[k] = sum;
  }
}
__device__ __forceinline__ uint32_t vc_get(const uint64_t (&c)[kCountBits], uint32_t j) {
  uint32_t v = 0;
#pragma unroll
  for (int k = 0; k < kCountBits; ++k) v |= (uint32_t)((c[k] >> j) & 1) << k;
  return v;
}

// Level features: block = kFeatWords words x kFeatSlices slices of one
// feature f, blockIdx.y mapped to the simulated levels' transition features
// and then their forward-mark features.  Each slice counts its share of the
// level's gates (a level of ~800 gates is ~26 per slice: the slices are what
// gives the launch enough waves), consecutive lanes reading consecutive words
// of each row; the slices' bit-sliced counters are added in a tree in LDS,
// and each slice converts and writes 2 of its word's 64 pairs.  Levels past
// the simulated depth are zero-filled by the host.
constexpr uint32_t kFeatWords = 8, kFeatSlices = 32;
constexpr uint32_t kFeatPairsPerSlice = 64 / kFeatSlices;

template <bool SEQ>
__global__ __launch_bounds__(256) void k_features(const Planes pl, const uint32_t *__restrict__ level_off,
                                                  uint32_t levels_used, uint64_t first_pair, uint64_t num_pairs,
                                                  uint16_t *__restrict__ feat, uint64_t ld) {
  const uint32_t wl = threadIdx.x % kFeatWords, sl = threadIdx.x / kFeatWords;
  const uint32_t w = blockIdx.x * kFeatWords + wl;
  const uint32_t f = blockIdx.y < levels_used ? blockIdx.y : 128 + blockIdx.y - levels_used;
  const bool live = w < pl.words;
  uint64_t c[kCountBits];
#pragma unroll
  for (int k = 0; k < kCountBits; ++k) c[k] = 0;
  float denom = 0;
  {
    const uint32_t lv = f & 127;
    const bool marks = f >= 128;
    {
      const uint32_t lo = level_off[lv], hi = level_off[lv + 1];
      denom = (float)(hi - lo);
      const uint32_t gb = lo + (uint32_t)((uint64_t)(hi - lo) * sl / kFeatSlices);
      const uint32_t ge = lo + (uint32_t)((uint64_t)(hi - lo) * (sl + 1) / kFeatSlices);
      for (uint32_t g0 = gb; live && g0 < ge; g0 += 63) {
        uint64_t s[kSmallBits];
#pragma unroll
        for (int k = 0; k < kSmallBits; ++k) s[k] = 0;
        const uint32_t g1 = min(ge, g0 + 63);
        uint32_t g = g0;
        for (; g + 8 <= g1; g += 8) {  // eight rows' loads in flight
          uint64_t x[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            if (marks) {
              x[u] = mark_of(pl, g + u, w);
            } else {
              uint64_t a, b;
              frames<SEQ>(pl, g + u, w, a, b);
              x[u] = a ^ b;
            }
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) vc_inc<kSmallBits>(s, x[u], 0);
        }
        for (; g < g1; ++g) {
          uint64_t x;
          if (marks) {
            x = mark_of(pl, g, w);
          } else {
            uint64_t a, b;
            frames<SEQ>(pl, g, w, a, b);
            x = a ^ b;
          }
          vc_inc<kSmallBits>(s, x, 0);
        }
        vc_add_sliced(c, s);
      }
    }
  }
  {  // add the slices' counters: a tree, the upper half into the lower
    __shared__ uint64_t red[kFeatSlices / 2][kCountBits][kFeatWords];
    for (uint32_t half = kFeatSlices / 2; half >= 1; half >>= 1) {
      if (sl >= half && sl < 2 * half) {
#pragma unroll
        for (int k = 0; k < kCountBits; ++k) red[sl - half][k][wl] = c[k];
      }
      __syncthreads();
      if (sl < half) {
        uint64_t y[kCountBits];
#pragma unroll
        for (int k = 0; k < kCountBits; ++k) y[k] = red[sl][k][wl];
        vc_add_sliced(c, y);
      }
      __syncthreads();
    }
    if (sl == 0) {
#pragma unroll
      for (int k = 0; k < kCountBits; ++k) red[0][k][wl] = c[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kCountBits; ++k) c[k] = red[0][k][wl];
  }
  if (!live) return;
  const float scale = denom > 0 ? 1.0f / denom : 0.0f;
  const uint64_t base = first_pair + 64ull * w;
  static_assert(kFeatPairsPerSlice == 2, "one 4-byte store per slice");
  const uint32_t j0 = kFeatPairsPerSlice * sl;
  uint16_t *dst = feat + (uint64_t)f * ld + base;
  const uint32_t v = (uint32_t)to_bf16((float)vc_get(c, j0) * scale) |
                     ((uint32_t)to_bf16((float)vc_get(c, j0 + 1) * scale) << 16);
  if (base + j0 + 2 <= num_pairs) {
    *reinterpret_cast<uint32_t *>(dst + j0) = v;
  } else if (base + j0 < num_pairs) {
    dst[j0] = (uint16_t)v;
  }
}

// PI-bucket features f = 256 + blockIdx.y: thread = word.  A bucket holds a
// few PIs, so each pair's (weighted) count is accumulated directly, 64 float
// registers (integers below 2^24, exact), then scaled and written as 8
// 16-byte stores.
template <bool SEQ>
__global__ __launch_bounds__(256) void k_features_pi(const GateRec *__restrict__ recs, Planes pl, uint32_t num_pi,
                                                     uint64_t first_pair, uint64_t num_pairs,
                                                     uint16_t *__restrict__ feat, uint64_t ld) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x, f = 256 + blockIdx.y;
  if (w >= pl.words) return;
  const uint32_t bkt = (f - 256) & 127;
  const uint32_t lo = (uint32_t)((uint64_t)bkt * num_pi / 128), hi = (uint32_t)((uint64_t)(bkt + 1) * num_pi / 128);
  const bool weighted = f >= 384;
  float acc[64];
#pragma unroll
  for (int j = 0; j < 64; ++j) acc[j] = 0.f;
  float denom = 0;
  for (uint32_t g = lo; g < hi; ++g) {
    const float wt = weighted ? (float)recs[g].nfo : 1.f;
    denom += wt;
    uint64_t a, b;
    frames<SEQ>(pl, g, w, a, b);
    const uint64_t t = a ^ b;
#pragma unroll
    for (int j = 0; j < 64; ++j) acc[j] += ((t >> j) & 1) ? wt : 0.f;
  }
  const float scale = denom > 0 ? 1.0f / denom : 0.0f;
  const uint64_t base = first_pair + 64ull * w;
  uint16_t *dst = feat + (uint64_t)f * ld + base;
#pragma unroll
  for (uint32_t j0 = 0; j0 < 64; j0 += 8) {
    uint32_t v[4];
#pragma unroll
    for (int e = 0; e < 4; ++e)
      v[e] = (uint32_t)to_bf16(acc[j0 + 2 * e] * scale) | ((uint32_t)to_bf16(acc[j0 + 2 * e + 1] * scale) << 16);
    if (base + j0 + 8 <= num_pairs) {
      *reinterpret_cast<uint4 *>(dst + j0) = make_uint4(v[0], v[1], v[2], v[3]);
    } else {
      for (uint32_t e = 0; e < 8 && base + j0 + e < num_pairs; ++e)
        dst[j0 + e] = (uint16_t)(v[e >> 1] >> (16 * (e & 1)));
    }
  }
}

// Pair permutation of PI-major packed bits by 64 x 64 bit transposes.  A
// wave per (PI block ib, word w); lane l holds one 64-bit word and 64
// ballots transpose the tile.  Pair-major layout: pm[p * nb + ib] holds PIs
// 64 ib .. 64 ib + 63 of pair p.
__global__ __launch_bounds__(256) void k_pi_to_pair(const uint64_t *__restrict__ src, uint64_t stride,
                                                    uint32_t num_pi, uint64_t words, uint32_t nb,
                                                    uint64_t *__restrict__ pm) {
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t lane = threadIdx.x & 63;
  if (wave >= words * nb) return;  // wave-uniform
  const uint64_t w = wave / nb;
  const uint32_t ib = (uint32_t)(wave - w * nb);
  const uint32_t i = 64 * ib + lane;
  const uint64_t x = i < num_pi ? src[(uint64_t)i * stride + w] : 0;
  uint64_t mine = 0;
#pragma unroll
  for (int j = 0; j < 64; ++j) {
    const uint64_t m = __ballot((x >> j) & 1);
    if (lane == (uint32_t)j) mine = m;
  }
  pm[(64 * w + lane) * nb + ib] = mine;
}

// dst pair q = src pair perm[q] (q < num), zero beyond.
__global__ void k_gather_pairs(const uint64_t *__restrict__ src, uint64_t *__restrict__ dst,
                               const uint32_t *__restrict__ perm, uint64_t num, uint32_t nb, uint64_t total) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const uint64_t q = t / nb;
  const uint32_t ib = (uint32_t)(t - q * nb);
  dst[t] = q < num ? src[(uint64_t)perm[q] * nb + ib] : 0;
}

__global__ __launch_bounds__(256) void k_pair_to_pi(const uint64_t *__restrict__ pm, uint32_t nb, uint32_t num_pi,
                                                    uint64_t words, uint64_t *__restrict__ dst, uint64_t stride) {
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t lane = threadIdx.x & 63;
  if (wave >= words * nb) return;  // wave-uniform
  const uint64_t w = wave / nb;
  const uint32_t ib = (uint32_t)(wave - w * nb);
  const uint64_t y = pm[(64 * w + lane) * nb + ib];
  uint64_t mine = 0;
#pragma unroll
  for (int i = 0; i < 64; ++i) {
    const uint64_t m = __ballot((y >> i) & 1);
    if (lane == (uint32_t)i) mine = m;
  }
  if (64 * ib + lane < num_pi) dst[(uint64_t)(64 * ib + lane) * stride + w] = mine;
}

__global__ void k_iota(uint32_t *p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (uint32_t)i;
}

__global__ void k_copy_u64(const uint64_t *src, uint64_t *dst) { *dst = *src; }
__global__ void k_copy_result(const uint64_t *total, const uint32_t *err, uint64_t *total_out, uint32_t *err_out) {
  *total_out = *total;
  if (err_out) *err_out = *err & kErrOverflow;
}

// ------------------------------------------------------------ batched grading
// Many circuits, one chunk each, graded by level launches over the union of
// their levels: launch k runs level k of every circuit that has one.  A
// launch's blocks are split into segments, one per circuit, each with that
// circuit's level map; a block finds its segment by binary search.
struct BatchCircuit {
  DevNet net;
  const GateRec *recs;
  Planes pl;
  const uint64_t *va;
  uint64_t vstride;
  uint32_t *first0, *first1, *list, *list_cnt;
  uint64_t list_stride;
  uint32_t *cstate;  // per-chunk counters, cleared by the last forward level
  uint32_t n_cstate;
  uint64_t *actmap;
  uint32_t *act_list, *act_count;
  uint64_t *gate, *chunk_total, *grand;
  uint32_t *err;
  uint32_t width;  // pairs
};

struct BatchSeg {
  LevelMap mp;
  uint32_t circuit, block0, nblocks, clear;
};

__device__ __forceinline__ uint32_t find_seg(const BatchSeg *__restrict__ segs, uint32_t nseg, uint32_t b) {
  uint32_t lo = 0, hi = nseg - 1;  // last segment starting at or before b
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (segs[mid].block0 <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// NW = 2 (whole-wave tiles only): segments of 128-word tiles, each lane
// computing words w and w + 64.
template <bool ROBUST, bool WIDE, int NW>
__global__ __launch_bounds__(kBlock) void k_fwd_batch(const BatchCircuit *__restrict__ circ,
                                                      const BatchSeg *__restrict__ segs, uint32_t nseg) {
  const BatchSeg sg = segs[find_seg(segs, nseg, blockIdx.x)];
  const BatchCircuit &c = circ[sg.circuit];
  const DevNet net = c.net;
  const Planes pl = c.pl;
  fwd_body<true, true, true, ROBUST, WIDE, NW>(net, c.recs, pl, c.va, nullptr, c.vstride, 0, sg.mp,
                                               sg.clear ? c.cstate : nullptr, sg.clear ? c.n_cstate : 0,
                                               blockIdx.x - sg.block0, sg.nblocks);
}

template <bool WIDE>
__global__ __launch_bounds__(kBlock) void k_rev_batch(const BatchCircuit *__restrict__ circ,
                                                      const BatchSeg *__restrict__ segs, uint32_t nseg) {
  const BatchSeg sg = segs[find_seg(segs, nseg, blockIdx.x)];
  const BatchCircuit &c = circ[sg.circuit];
  const DevNet net = c.net;
  const Planes pl = c.pl;
  rev_body<true, true, true, WIDE>(net, c.recs, pl, sg.mp, 0, c.first0, c.first1, c.list, c.list_cnt,
                                   c.list_stride, blockIdx.x - sg.block0);
}

// Compacted reverse pass (PDF_TUNE_REV_COMPACT): most (gate, 64-word tile)
// items of a level hold no mark at all (their mark-present word is zero), yet
// a plain level launch spends a wave on each.  A scan launch, thread per
// item, appends the items with a non-zero mark-present word to a work list;
// a work launch, a fixed grid of waves striding over the list, runs the same
// per-(gate, word) reverse logic on them, lane = word of the tile.
struct BatchScanSeg {
  uint32_t circuit, lo, count, tiles;  // items = count gates x tiles 64-word tiles
  uint32_t block0, pad;
};
struct RevItem {
  uint32_t circuit, gate, tile;
};

__device__ __forceinline__ uint32_t find_scan_seg(const BatchScanSeg *__restrict__ segs, uint32_t nseg, uint32_t b) {
  uint32_t lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (segs[mid].block0 <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_rev_scan_batch(const BatchCircuit *__restrict__ circ,
                                                        const BatchScanSeg *__restrict__ segs, uint32_t nseg,
                                                        RevItem *__restrict__ work, uint32_t *__restrict__ work_cnt) {
  const BatchScanSeg sg = segs[find_scan_seg(segs, nseg, blockIdx.x)];
  const uint64_t item = (uint64_t)(blockIdx.x - sg.block0) * 256 + threadIdx.x;
  bool marked = false;
  uint32_t g = 0, t = 0;
  if (item < (uint64_t)sg.count * sg.tiles) {
    g = sg.lo + (uint32_t)(item / sg.tiles);
    t = (uint32_t)(item % sg.tiles);
    const Planes &pl = circ[sg.circuit].pl;
    marked = pl.NZ[(uint64_t)g * pl.ntiles + t] != 0;
  }
  const uint64_t bal = __ballot(marked);
  if (!bal) return;  // whole wave
  const uint32_t lane = threadIdx.x & 63;
  const int leader = (int)__builtin_ctzll(bal);
  uint32_t base = 0;
  if ((int)lane == leader) base = atomicAdd(work_cnt, (uint32_t)__builtin_popcountll(bal));
  base = __shfl(base, leader);
  if (marked) work[base + (uint32_t)__builtin_popcountll(bal & ((1ULL << lane) - 1))] = RevItem{sg.circuit, g, t};
}

__global__ __launch_bounds__(256) void k_rev_work_batch(const BatchCircuit *__restrict__ circ,
                                                        const RevItem *__restrict__ work,
                                                        const uint32_t *__restrict__ work_cnt) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t nwaves = gridDim.x * 4;
  const uint32_t cnt = *work_cnt;
  for (uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6); i < cnt; i += nwaves) {
    const RevItem it = work[i];
    const uint32_t ci = __builtin_amdgcn_readfirstlane(it.circuit);
    const uint32_t g = __builtin_amdgcn_readfirstlane(it.gate);
    const uint32_t t = __builtin_amdgcn_readfirstlane(it.tile);
    const BatchCircuit &c = circ[ci];
    const uint32_t w = t * 64 + lane;
    if (w < c.pl.words)
      rev_item<true, true, true>(c.net, c.recs, c.pl, g, w, 0, c.first0, c.first1, c.list, c.list_cnt,
                                 c.list_stride);
  }
}

// The batch's compacted reverse pass from lists its forward launches built
// (PDF_TUNE_REV_COMPACT = 2): union level k's marked (gate, 64-word tile)
// items of every circuit, in 64 sub-lists (by forward block), the circuit in
// the tile field's high bits; a fixed grid of waves strides over them, lane
// = word of the tile, and the item's mark-present bits come with it.
constexpr uint32_t kBatchTileBits = 20;  // tiles of a circuit < 2^20, circuits < 2^12
__global__ __launch_bounds__(256) void k_rev_list_batch(const BatchCircuit *__restrict__ circ,
                                                        const RevTile *__restrict__ items,
                                                        const uint32_t *__restrict__ items_cnt, uint32_t items_cap) {
  const uint32_t lane = threadIdx.x & 63;
  uint32_t pre, cnt;
  sublist_prefix(items_cnt, lane, pre, cnt);
  const uint32_t nwaves = gridDim.x * 4;
  for (uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6); i < cnt; i += nwaves) {
    const RevTile it = items[sublist_slot(pre, i, items_cap)];
    const uint32_t ci = __builtin_amdgcn_readfirstlane(it.tile >> kBatchTileBits);
    const uint32_t t = __builtin_amdgcn_readfirstlane(it.tile & ((1u << kBatchTileBits) - 1));
    const uint32_t g = __builtin_amdgcn_readfirstlane(it.gate);
    const BatchCircuit &c = circ[ci];
    const GateRec r = c.recs[g];
    pin(r.fin[0]);
    pin((uint32_t)r.nfi);
    pin((uint32_t)r.po);
    // bits are set only for words of the circuit's chunk
    if ((it.bits >> lane) & 1)
      rev_marked<true, true, true>(c.net, r, c.pl, g, t * 64 + lane, 0, c.first0, c.first1, c.list, c.list_cnt,
                                   c.list_stride);
  }
}

// Thread t = node t - node_base[c] of circuit c (node_base: prefix sums).
__device__ __forceinline__ uint32_t find_circuit(const uint32_t *__restrict__ node_base, uint32_t C, uint32_t t) {
  uint32_t lo = 0, hi = C - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (node_base[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void k_reset_batch(const BatchCircuit *__restrict__ circ, const uint32_t *__restrict__ node_base,
                              uint32_t C, uint32_t total) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const uint32_t ci = find_circuit(node_base, C, t);
  const BatchCircuit &c = circ[ci];
  const uint32_t g = t - node_base[ci];
  c.first0[g] = PDF_NONE;
  c.first1[g] = PDF_NONE;
  c.gate[g] = 0;
  if (g == 0) {
    *c.grand = 0;
    *c.err = 0;
  }
}

__global__ void k_events_batch(const BatchCircuit *__restrict__ circ, const uint32_t *__restrict__ node_base,
                               uint32_t C, uint32_t total) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const uint32_t ci = find_circuit(node_base, C, t);
  const BatchCircuit &c = circ[ci];
  events_one<false>(t - node_base[ci], c.first0, c.first1, 0, c.width, c.actmap, c.act_list, c.act_count);
}

// Every circuit's first-detection pairs, dealt round-robin to the waves
// across circuits; c/h fallback scratch is one scratch_n slice per wave.
__global__ __launch_bounds__(64) void k_cover_batch(const BatchCircuit *__restrict__ circ, uint32_t C, uint64_t *cs,
                                                    uint64_t *hs, uint64_t scratch_n) {
  __shared__ CoverLds S;
  uint64_t *c_s = cs + (uint64_t)blockIdx.x * scratch_n;
  uint64_t *h_s = hs + (uint64_t)blockIdx.x * scratch_n;
  uint64_t dealt = 0;  // pairs of the circuits before this one
  for (uint32_t ci = 0; ci < C; ++ci) {
    const BatchCircuit &c = circ[ci];
    const uint32_t nact = *c.act_count;
    const DevNet net = c.net;
    const Planes pl = c.pl;
    uint32_t k = (uint32_t)((blockIdx.x + gridDim.x - dealt % gridDim.x) % gridDim.x);
    for (; k < nact; k += gridDim.x)
      cover_one<false>(S, net, pl, 0, c.list, c.list_cnt, c.list_stride, c.first0, c.first1, c_s, h_s, c.gate,
                       c.chunk_total, c.grand, c.err, nullptr, c.act_list[k]);
    dealt += nact;
  }
}

__global__ void k_gather_totals(const BatchCircuit *__restrict__ circ, uint32_t C, uint64_t *__restrict__ totals,
                                uint32_t *__restrict__ errs) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  totals[c] = *circ[c].grand;
  errs[c] = *circ[c].err;
}

}  // namespace

// ----------------------------------------------------------------- host side

struct pdf_grader {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  Netlist nl;
  // netlist in HBM
  GateRec *d_rec = nullptr;
  uint8_t *d_type = nullptr, *d_po = nullptr;
  uint32_t *d_fin_off = nullptr, *d_fin = nullptr, *d_fout_off = nullptr, *d_fout = nullptr, *d_level = nullptr;
  uint32_t *d_level_off = nullptr;
  // the level walk's forward order: each level's records sorted by how much
  // work their gate takes (walk_order), fout_off holding the node's id
  GateRec *d_wrec = nullptr;
  GateRec *d_frec = nullptr;  // the per-level forward launches' records in fanin_order
  uint32_t tune_fwd_order = 0;  // PDF_TUNE_FWD_ORDER
  uint32_t widest = 0;  // nodes of the widest level

  // vectors: sequence mode keeps vwords + 1 words per PI row (positions
  // wrapped mod num_pairs, so the word after the last is available)
  uint64_t num_pairs = 0, vwords = 0, vstride = 0;
  bool seq = true;
  uint64_t *d_va = nullptr, *d_vb = nullptr;
  // chunk planes
  uint64_t chunk_pairs = 0, pitch = 0, ntiles = 0;
  uint64_t *d_V = nullptr, *d_VB = nullptr, *d_M = nullptr, *d_NZ = nullptr, *d_ACC = nullptr;
  // the level walk's tile-major value plane ([tiles][n][words a node]), when
  // PDF_TUNE_WALK is set; walk_planes: the last chunk graded left its value
  // planes there (its tile width), not in V / VB
  uint64_t *d_VT = nullptr;
  uint32_t walk_planes = 0;
  // merge / cover state
  uint32_t *d_first0 = nullptr, *d_first1 = nullptr;
  // segments of length seg_k >= 2 (stage API): node ids [seg_count][seg_k]
  // padded with PDF_NONE, each segment's first T0 / T1 pair, and the chunk's
  // plane of nodes new at a pair ([n][pitch] bits, allocated on first use)
  uint32_t seg_k = 0, merged_segs = 1;
  uint64_t seg_count = 0;
  uint32_t *d_seg = nullptr, *d_sfirst0 = nullptr, *d_sfirst1 = nullptr;
  uint64_t *d_NEW = nullptr;
  uint32_t grade_segs = 1;  // segment length of the fused grade being queued
  bool one_lane = false;    // queue the fused grade on one stream (pdf_grade_async)
  // stage-API debug copies (pdf_grader_set_debug): the forward marks before
  // removal (mark plane + present bits) and each PI's c per pair of the
  // chunk ([num_pi][chunk_pairs])
  uint32_t debug = 0;
  uint64_t *d_dbg_M = nullptr, *d_dbg_NZ = nullptr, *d_dbg_pic = nullptr;
  int64_t dbg_fwd_chunk = -1, dbg_pic_chunk = -1;
  uint64_t *pic_target = nullptr;  // d_dbg_pic while pdf_cover runs with the debug flag
  uint64_t *d_gate = nullptr;
  // exact total over all chunks and the error flags, one allocation
  // (d_err points at word 1) read back in one copy into h_result (pinned)
  uint64_t *d_grand = nullptr;
  uint32_t *d_err = nullptr;
  uint64_t *h_result = nullptr;
  uint32_t *d_list = nullptr, *d_act_list = nullptr;
  // per-chunk counters in one zeroed u32 block: actmap (u64 per word),
  // list_cnt (per word), act_count, chunk total (u64)
  uint32_t *d_cstate = nullptr, n_cstate = 0;
  uint64_t *d_actmap = nullptr, *d_chunk_total = nullptr;
  uint32_t *d_list_cnt = nullptr, *d_act_count = nullptr;
  uint64_t *d_cs = nullptr, *d_hs = nullptr;
  uint32_t cover_waves = 0;
  CoverCaps cover_caps;  // PDF_TUNE_COVER_NODES / PDF_TUNE_COVER_POOL
  uint32_t max_cover_waves = 512;  // 1 for batch members (the batch has its own scratch)
  // per-launch event timing (pdf_grade_timed)
  std::vector<hipEvent_t> ev_pool;
  std::vector<int> ev_family;
  std::vector<uint32_t> ev_count;  // launches in each timed run
  bool timing = false, tm_open = false;
  // tuning knobs (pdf_grader_set_tuning)
  uint32_t tune_words_per_lane = 1;
  uint32_t tune_rev_tiles = 0;  // PDF_TUNE_REV_TILES (0: by tile count)
  uint32_t tune_rev_compact = 0;  // PDF_TUNE_REV_COMPACT: reverse levels over the forward pass's item lists
  // the compacted reverse pass's items of the lane's chunk, level k >= 1 at
  // (level_off[k] - level_off[1]) * ntiles, and their per-level counts
  RevTile *d_rtl = nullptr;
  uint32_t *d_rtl_cnt = nullptr;
  // the lists the last forward pass queued filled (its lane's d_rtl, or
  // null) and its chunk start: a reverse pass of that chunk on that lane
  // runs over them; every other forward call clears this
  const RevTile *rev_lists_buf = nullptr;
  uint64_t rev_lists_start = 0;
  uint32_t tune_use_graph = 1;
  uint32_t tune_walk = 0;  // PDF_TUNE_WALK: 0 level launches, else words per tile of the level walk
  uint32_t walk_full = 0;  // the last walk wrote every node's record (debug flags were set)
  uint32_t merge_window = 0;  // pdf_grader_set_merge_window
  uint32_t tune_rev_walk = 0;  // PDF_TUNE_REV_WALK: words per tile of a reverse walk after per-level forward launches
  // chunk lanes (PDF_TUNE_LANES): the fused grade runs consecutive chunks on
  // `lanes` streams, each lane with its own chunk buffers; lane 0 is the set
  // of fields above, lanes 1.. are kept here
  struct Lane {
    hipStream_t stream = nullptr;
    uint64_t *V = nullptr, *VB = nullptr, *M = nullptr, *NZ = nullptr, *ACC = nullptr, *VT = nullptr;
    uint32_t *list = nullptr, *cstate = nullptr, *act_list = nullptr;
    uint64_t *actmap = nullptr, *chunk_total = nullptr;
    uint32_t *list_cnt = nullptr, *act_count = nullptr;
    uint64_t *cs = nullptr, *hs = nullptr;
    uint64_t *NEW = nullptr;    // new-node plane (segments of length K >= 2)
    RevTile *rtl = nullptr;
    uint32_t *rtl_cnt = nullptr;
    hipEvent_t done = nullptr;  // merges of this lane's latest chunk and all before it
  };
  std::vector<Lane> lanes;      // lanes 1..; empty for one lane
  // with lanes and graphs: each chunk's forward and reverse levels captured
  // on its lane's stream (a graph per chunk; covers are launched directly)
  std::vector<hipGraphExec_t> chunk_graphs;
  int chunk_graphs_robust = -1;
  hipEvent_t lane0_done = nullptr, lane_start = nullptr;
  uint32_t tune_lanes = 1;
  // ranking scratch (pdf_grader_rank_pairs), kept between calls
  float *d_rk_keys = nullptr;
  uint32_t *d_rk_idx = nullptr, *d_rk_perm = nullptr;
  uint64_t *d_rk_va = nullptr, *d_rk_vb = nullptr, *d_rk_pm0 = nullptr, *d_rk_pm1 = nullptr;
  void *d_rk_temp = nullptr;
  size_t rk_pairs = 0, rk_vsize = 0, rk_temp_bytes = 0;
  // captured fused grade (pdf_grade), rebuilt when its shape changes
  hipGraphExec_t graph = nullptr;
  int graph_robust = -1;  // robust | segs << 1 of the captured grade
  // call-order tracking for the stage API
  int64_t stage_chunk = -1;
  int stage = 0;  // 1 = simulated, 2 = marked, 3 = merged, 4 = covered
  int stage_robust = 0;

  DevNet net() const { return DevNet{d_rec, d_type, d_po, d_fin_off, d_fin, d_fout_off, d_fout, d_level, nl.n}; }
};

namespace {

bool hip_ok(hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return false;
}

#define HIPCHK(expr)                                \
  do {                                              \
    if (!hip_ok((expr), #expr)) return PDF_ERR_HIP; \
  } while (0)

template <typename T>
int dalloc(T **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  if (hipMalloc((void **)p, sizeof(T) * count) != hipSuccess) {
    (void)hipGetLastError();
    set_error("hipMalloc of " + std::to_string(sizeof(T) * count) + " bytes failed");
    return PDF_ERR_NOMEM;
  }
  return PDF_OK;
}

template <typename T>
void dfree(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

uint32_t grid_for(uint64_t items, uint32_t block) { return (uint32_t)((items + block - 1) / block); }

enum Family { kFwd = 0, kRev = 1, kEvents = 2, kCover = 3, kFill = 4, kSegMerge = 5 };

// With g->timing set, each run of consecutive launches of one kernel family
// (e.g. the forward pass's level launches) is bracketed by a pair of pooled
// events, so the per-launch average includes the short gaps between the
// launches of the run but no event records inside it.
void tm_close(pdf_grader *g) {
  if (!g->timing || !g->tm_open) return;
  (void)hipEventRecord(g->ev_pool[g->ev_family.size() * 2 - 1], g->stream);
  g->tm_open = false;
}
void tm_begin(pdf_grader *g, int family) {
  if (!g->timing) return;
  if (g->tm_open && g->ev_family.back() == family) {
    g->ev_count.back() += 1;
    return;
  }
  tm_close(g);
  const size_t i = g->ev_family.size() * 2;
  while (g->ev_pool.size() < i + 2) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) { g->timing = false; return; }
    g->ev_pool.push_back(e);
  }
  g->ev_family.push_back(family);
  g->ev_count.push_back(1);
  (void)hipEventRecord(g->ev_pool[i], g->stream);
  g->tm_open = true;
}
void tm_end(pdf_grader *) {}

int check_chunk(pdf_grader *g, uint64_t chunk, uint64_t start, uint64_t *width) {
  if (!g->num_pairs) { set_error("no vectors loaded"); return PDF_ERR_STATE; }
  if (!g->chunk_pairs) { set_error("chunk size not set"); return PDF_ERR_STATE; }
  const uint64_t nchunks = (g->num_pairs + g->chunk_pairs - 1) / g->chunk_pairs;
  if (chunk >= nchunks) { set_error("chunk index out of range"); return PDF_ERR_ARG; }
  if (start != chunk * g->chunk_pairs) { set_error("start_pattern must equal chunk * chunk_pairs"); return PDF_ERR_ARG; }
  *width = std::min(g->chunk_pairs, g->num_pairs - start);
  return PDF_OK;
}

Planes planes(pdf_grader *g, uint64_t width) {
  const uint32_t tail = (uint32_t)(width % 64);
  return Planes{g->d_V,  g->d_VB, g->d_M, g->d_NZ, g->d_ACC, g->pitch, (uint32_t)((width + 63) / 64),
                (uint32_t)g->ntiles, tail ? (1ULL << tail) - 1 : ~0ULL, g->merge_window};
}

LevelMap level_map(uint32_t words, uint32_t lo, uint32_t count, bool extra_column, uint32_t *grid) {
  uint32_t twlog = 0;
  while (twlog < 6 && (1u << twlog) < words) ++twlog;
  LevelMap mp{(words + (1u << twlog) - 1) >> twlog, twlog, words, lo, count, 0};
  const uint32_t per_block = kBlock >> twlog;
  mp.main_blocks = mp.T * ((count + per_block - 1) / per_block);
  *grid = mp.main_blocks + (extra_column ? grid_for(count, kBlock) : 0);
  return mp;
}

// Blocks of a compacted reverse launch: enough waves for a level's marked
// items in one or two strides (2M gates: ~10k marked items of ~320k).
constexpr uint64_t kRevListBlocks = 2048;

// The compacted reverse pass's sub-lists for chunks of `ntiles` 64-word
// tiles, the entries they hold per gate of a level (a level of `count`
// gates: sub-lists of count x ceil(ntiles / nsub) entries), and the
// placement of level k's lists and counters.
inline uint32_t rev_nsub(uint64_t ntiles) { return (uint32_t)std::min<uint64_t>(ntiles, kRevSubLists); }
inline uint64_t rev_slots_per_gate(uint64_t ntiles) {
  const uint64_t ns = rev_nsub(ntiles);
  return ns * ((ntiles + ns - 1) / ns);
}
void rev_lists_of(const pdf_grader *g, uint32_t k, LevelMap &mp) {
  const uint64_t ns = rev_nsub(g->ntiles);
  const uint32_t lo = g->nl.level_off[k], count = g->nl.level_off[k + 1] - lo;
  mp.items = g->d_rtl + rev_slots_per_gate(g->ntiles) * (lo - g->nl.level_off[1]);
  mp.items_cnt = g->d_rtl_cnt + (uint64_t)k * kRevSubLists * kRevCntStride;
  mp.items_cap = (uint32_t)(count * ((g->ntiles + ns - 1) / ns));
  mp.items_nsub = (uint32_t)ns;
}

// Words per lane for the fused sequence-mode pass (tuning knob, default 1):
// two only when the chunk splits into 128-word tiles whose count is a
// multiple of 8 (each tile keeps its XCD).
uint32_t words_per_lane(uint32_t want, bool seq, bool sim, bool mark, uint64_t width) {
  const uint64_t words = (width + 63) / 64;
  for (uint32_t nw = want; nw >= 2; nw /= 2)
    if (seq && sim && mark && width % 64 == 0 && words % (64 * nw) == 0 && (words / (64 * nw)) % 8 == 0) return nw;
  return 1;
}

template <bool SEQ, bool S, bool M, bool R, bool WIDE, int NW>
void fwd_one(pdf_grader *g, uint32_t grid, const DevNet &net, const Planes &pl, uint64_t wbase, const LevelMap &mp,
             uint32_t *clear, uint32_t nclear) {
  const GateRec *recs = mp.perm ? g->d_frec : g->d_rec;
  k_fwd<SEQ, S, M, R, WIDE, NW><<<grid, kBlock, 0, g->stream>>>(net, recs, pl, g->d_va, g->d_vb, g->vstride, wbase,
                                                                 mp, clear, nclear);
}

// clear_state: the last level's extra-column threads zero the per-chunk
// merge/cover counters (sequence-mode simulation only).
template <bool SEQ, bool S, bool M>
void launch_fwd_t(pdf_grader *g, uint64_t start, uint64_t width, int robust, bool clear_state, uint32_t last_level,
                  uint32_t k_lo = 0, uint32_t k_hi = 0xFFFFFFFFu) {
  const Planes pl = planes(g, width);
  const DevNet net = g->net();
  const uint32_t nw = words_per_lane(g->tune_words_per_lane, SEQ, S, M, width);
  // the compacted reverse pass's item lists: built by a marking pass over
  // every level whose waves hold one gate each (chunks of >= 64 words)
  const bool lists = M && g->d_rtl && k_lo == 0 && k_hi >= g->nl.L && last_level >= g->nl.L && pl.words >= 64 &&
                     g->nl.L >= 1;
  g->rev_lists_buf = lists ? g->d_rtl : nullptr;
  g->rev_lists_start = start;
  if (lists) {
    const uint64_t nc = ((uint64_t)g->nl.L + 1) * kRevSubLists * kRevCntStride;
    tm_begin(g, kFill);
    k_fill_u32<<<(uint32_t)std::min<uint64_t>(grid_for(nc, 256), 1024), 256, 0, g->stream>>>(g->d_rtl_cnt, nc, 0u);
    tm_end(g);
  }
  for (uint32_t k = k_lo; k <= std::min(std::min(g->nl.L, last_level), k_hi); ++k) {
    uint32_t grid;
    const uint32_t lo = g->nl.level_off[k], count = g->nl.level_off[k + 1] - lo;
    LevelMap mp = level_map(pl.words, lo, count, SEQ && S, &grid);
    if (nw >= 2) {  // 64 nw-word tiles, one node per wave
      mp.T = pl.words / (64 * nw);
      mp.main_blocks = mp.T * ((count + (kBlock >> 6) - 1) / (kBlock >> 6));
      grid = mp.main_blocks + grid_for(count, kBlock);
    }
    if (lists && k >= 1) rev_lists_of(g, k, mp);  // level 0's reverse pass stays dense
    mp.perm = g->tune_fwd_order && g->d_frec;
    const bool last = k == std::min(g->nl.L, last_level) && SEQ && S && clear_state;
    uint32_t *clear = last ? g->d_cstate : nullptr;
    const uint32_t nclear = last ? g->n_cstate : 0;
    const uint64_t wbase = start / 64;
    tm_begin(g, kFwd);
    if (nw == 2) {
      if (robust) fwd_one<SEQ, S, M, true, true, (SEQ && S && M) ? 2 : 1>(g, grid, net, pl, wbase, mp, clear, nclear);
      else fwd_one<SEQ, S, M, false, true, (SEQ && S && M) ? 2 : 1>(g, grid, net, pl, wbase, mp, clear, nclear);
    } else if (mp.twlog == 6) {
      if (robust) fwd_one<SEQ, S, M, true, true, 1>(g, grid, net, pl, wbase, mp, clear, nclear);
      else fwd_one<SEQ, S, M, false, true, 1>(g, grid, net, pl, wbase, mp, clear, nclear);
    } else {
      if (robust) fwd_one<SEQ, S, M, true, false, 1>(g, grid, net, pl, wbase, mp, clear, nclear);
      else fwd_one<SEQ, S, M, false, false, 1>(g, grid, net, pl, wbase, mp, clear, nclear);
    }
    tm_end(g);
  }
}

template <bool S, bool M>
int launch_fwd(pdf_grader *g, uint64_t start, uint64_t width, int robust, bool clear_state = false,
               uint32_t last_level = 0xFFFFFFFFu, uint32_t k_lo = 0, uint32_t k_hi = 0xFFFFFFFFu) {
  g->walk_planes = 0;  // M and NZ now hold this pass's marks
  if (g->seq) launch_fwd_t<true, S, M>(g, start, width, robust, clear_state, last_level, k_lo, k_hi);
  else launch_fwd_t<false, S, M>(g, start, width, robust, false, last_level, k_lo, k_hi);
  HIPCHK(hipGetLastError());
  return PDF_OK;
}

// 64-word tiles per wave of the reverse pass for a level of `count` gates
// and T tiles: the knob's value, or (0) the largest of 8, 4, 2 that leaves
// the launch at least 16,384 waves (64 per CU), else 1.  Measured on an
// MI355X: 2M gates x 32 tiles (320k waves) 8 tiles per wave, reverse pass
// 53.6 -> 33.4 ms a grade; 50k gates x 128 tiles (106k waves) 4 or 8,
// 8.3 -> 7.9 ms; 50k gates x 8, 16 or 32 tiles (6.6k-27k waves) 1: there
// most waves hold a mark and the extra tiles only lengthen each wave.
uint32_t rev_tiles(uint32_t want, uint32_t T, uint32_t count) {
  if (want) return want;
  uint32_t rt = 8;
  while (rt > 1 && (uint64_t)count * T < 16384ull * rt) rt >>= 1;
  return rt;
}

template <bool SEQ, bool R, bool M>
void launch_reverse_t(pdf_grader *g, uint64_t start, uint64_t width, uint32_t k_lo = 0,
                      uint32_t k_hi = 0xFFFFFFFFu) {
  const Planes pl = planes(g, width);
  const DevNet net = g->net();
  // the item lists of this chunk's forward pass on this lane, if it built them
  const bool lists = g->d_rtl && g->rev_lists_buf == g->d_rtl && g->rev_lists_start == start;
  for (uint32_t k = std::min(g->nl.L, k_hi) + 1; k-- > k_lo;) {
    uint32_t grid;
    const LevelMap mp =
        level_map(pl.words, g->nl.level_off[k], g->nl.level_off[k + 1] - g->nl.level_off[k], false, &grid);
    tm_begin(g, kRev);
    if (lists && k >= 1) {
      // at most every (gate, tile) item of the level; waves stride over the count
      const uint32_t lgrid = (uint32_t)std::min<uint64_t>(
          kRevListBlocks, std::max<uint64_t>(1, ((uint64_t)mp.count * mp.T + kBlock / 64 - 1) / (kBlock / 64)));
      LevelMap lm = mp;
      rev_lists_of(g, k, lm);
      k_reverse_list<SEQ, R, M><<<lgrid, kBlock, 0, g->stream>>>(net, g->d_rec, pl, lm.items, lm.items_cnt,
                                                                 lm.items_cap, (uint32_t)start, g->d_first0,
                                                                 g->d_first1, g->d_list, g->d_list_cnt, g->nl.n);
      tm_end(g);
      continue;
    }
    const uint32_t rt = mp.twlog == 6 ? rev_tiles(g->tune_rev_tiles, mp.T, mp.count) : 1;
    const uint32_t rgrid = ((mp.count + kBlock / 64 - 1) / (kBlock / 64)) * ((mp.T + rt - 1) / rt);
    if (rt == 8)
      k_reverse_tiles<SEQ, R, M, 8><<<rgrid, kBlock, 0, g->stream>>>(net, g->d_rec, pl, mp, (uint32_t)start,
                                                                      g->d_first0, g->d_first1, g->d_list,
                                                                      g->d_list_cnt, g->nl.n);
    else if (rt == 4)
      k_reverse_tiles<SEQ, R, M, 4><<<rgrid, kBlock, 0, g->stream>>>(net, g->d_rec, pl, mp, (uint32_t)start,
                                                                      g->d_first0, g->d_first1, g->d_list,
                                                                      g->d_list_cnt, g->nl.n);
    else if (rt == 2)
      k_reverse_tiles<SEQ, R, M, 2><<<rgrid, kBlock, 0, g->stream>>>(net, g->d_rec, pl, mp, (uint32_t)start,
                                                                      g->d_first0, g->d_first1, g->d_list,
                                                                      g->d_list_cnt, g->nl.n);
    else if (mp.twlog == 6)
      k_reverse<SEQ, R, M, true><<<grid, kBlock, 0, g->stream>>>(net, g->d_rec, pl, mp, (uint32_t)start, g->d_first0,
                                                                 g->d_first1, g->d_list, g->d_list_cnt, g->nl.n);
    else
      k_reverse<SEQ, R, M, false><<<grid, kBlock, 0, g->stream>>>(net, g->d_rec, pl, mp, (uint32_t)start, g->d_first0,
                                                                  g->d_first1, g->d_list, g->d_list_cnt, g->nl.n);
    tm_end(g);
  }
}

int reset_chunk(pdf_grader *g) {
  tm_begin(g, kFill);
  k_reset_chunk<<<grid_for(g->n_cstate, 256), 256, 0, g->stream>>>(g->d_cstate, g->n_cstate);
  tm_end(g);
  HIPCHK(hipGetLastError());
  return PDF_OK;
}

template <bool R, bool M>
int launch_reverse(pdf_grader *g, uint64_t start, uint64_t width, uint32_t k_lo = 0, uint32_t k_hi = 0xFFFFFFFFu) {
  if (g->seq) launch_reverse_t<true, R, M>(g, start, width, k_lo, k_hi);
  else launch_reverse_t<false, R, M>(g, start, width, k_lo, k_hi);
  HIPCHK(hipGetLastError());
  return PDF_OK;
}

template <bool SEQ, int TW, bool NM>
void launch_rev_walk_t(pdf_grader *g, uint64_t start, const Planes &pl, uint32_t tiles) {
  const uint32_t grid = 8 * ((tiles + 7) / 8);  // walk_tile skips the blocks past the last tile
  tm_begin(g, kRev);
  k_walk_rev<SEQ, TW, NM><<<grid, kWalkBlock, 0, g->stream>>>(g->net(), g->d_rec, pl, g->d_level_off,
                                                              g->nl.L + 1, tiles, (uint32_t)start, g->d_first0,
                                                              g->d_first1, g->d_list, g->d_list_cnt, g->nl.n);
  tm_end(g);
}

// The reverse pass alone as a level walk (tile width tw) over node-major
// planes, after per-level forward launches.
int launch_rev_walk(pdf_grader *g, uint64_t start, uint64_t width, uint32_t tw) {
  const Planes pl = planes(g, width);
  const uint32_t tiles = (pl.words + tw - 1) / tw;
  if (g->seq) {
    switch (tw) {
      case 1: launch_rev_walk_t<true, 1, true>(g, start, pl, tiles); break;
      case 2: launch_rev_walk_t<true, 2, true>(g, start, pl, tiles); break;
      case 4: launch_rev_walk_t<true, 4, true>(g, start, pl, tiles); break;
      default: launch_rev_walk_t<true, 8, true>(g, start, pl, tiles); break;
    }
  } else {
    switch (tw) {
      case 1: launch_rev_walk_t<false, 1, true>(g, start, pl, tiles); break;
      case 2: launch_rev_walk_t<false, 2, true>(g, start, pl, tiles); break;
      case 4: launch_rev_walk_t<false, 4, true>(g, start, pl, tiles); break;
      default: launch_rev_walk_t<false, 8, true>(g, start, pl, tiles); break;
    }
  }
  HIPCHK(hipGetLastError());
  return PDF_OK;
}

// The level walk's mark-present bytes BT[tile][n] in the lane's walk
// buffer: after the records of every tile of a full chunk.
uint8_t *walk_bits(const pdf_grader *g, uint32_t tw) {
  const uint64_t tiles_alloc = ((g->chunk_pairs + 63) / 64 + tw - 1) / tw;
  return reinterpret_cast<uint8_t *>(g->d_VT + (uint64_t)g->nl.n * tiles_alloc * walk_record_words(g->seq, tw));
}

template <bool SEQ, bool R, int TW>
void launch_walk_t(pdf_grader *g, uint64_t start, const Planes &pl, uint32_t tiles) {
  const DevNet net = g->net();
  const uint32_t grid = 8 * ((tiles + 7) / 8);  // walk_tile skips the blocks past the last tile
  uint8_t *bt = walk_bits(g, TW);
  tm_begin(g, kFwd);
  k_walk<SEQ, R, TW><<<grid, kWalkBlock, 0, g->stream>>>(net, g->d_rec, g->d_wrec, pl, g->d_va, g->d_vb, g->vstride, start / 64,
                                                          g->d_VT, bt, g->walk_full, g->d_level_off, g->nl.L + 1,
                                                          g->widest, tiles, g->d_actmap,
                                                          g->d_act_count, g->d_chunk_total, (uint32_t)start,
                                                          g->d_first0, g->d_first1, g->d_list, g->d_list_cnt,
                                                          g->nl.n);
  tm_end(g);
}

template <bool SEQ, bool R>
void launch_walk_tw(pdf_grader *g, uint64_t start, const Planes &pl, uint32_t tw) {
  const uint32_t tiles = (pl.words + tw - 1) / tw;
  switch (tw) {
    case 1: launch_walk_t<SEQ, R, 1>(g, start, pl, tiles); break;
    case 2: launch_walk_t<SEQ, R, 2>(g, start, pl, tiles); break;
    case 4: launch_walk_t<SEQ, R, 4>(g, start, pl, tiles); break;
    default: launch_walk_t<SEQ, R, 8>(g, start, pl, tiles); break;
  }
}

// The forward and reverse passes of one chunk in one level-walk launch
// (words per tile = g->tune_walk), which also zeroes the per-chunk counters.
int launch_walk(pdf_grader *g, uint64_t start, uint64_t width, int robust) {
  const Planes pl = planes(g, width);
  const uint32_t tw = g->tune_walk;
  g->walk_planes = tw;
  // every record is kept only for the sim-plane getters (debug flags set)
  g->walk_full = g->debug != 0;
  if (g->seq) {
    if (robust) launch_walk_tw<true, true>(g, start, pl, tw);
    else launch_walk_tw<true, false>(g, start, pl, tw);
  } else {
    if (robust) launch_walk_tw<false, true>(g, start, pl, tw);
    else launch_walk_tw<false, false>(g, start, pl, tw);
  }
  HIPCHK(hipGetLastError());
  return PDF_OK;
}

// Sim, mark, removal and merge of one chunk of a fused grade: the level walk
// when PDF_TUNE_WALK is set, else the per-level launches.
int launch_passes(pdf_grader *g, uint64_t start, uint64_t width, int robust) {
  // (segment merges for K >= 2 read node-major value planes)
  if (g->tune_walk && g->grade_segs == 1 && g->d_VT) return launch_walk(g, start, width, robust);
  g->walk_planes = 0;
  int rc = launch_fwd<true, true>(g, start, width, robust, /*clear_state=*/true);
  if (!rc && !g->seq) rc = reset_chunk(g);
  if (!rc) rc = g->tune_rev_walk ? launch_rev_walk(g, start, width, g->tune_rev_walk)
                                 : launch_reverse<true, true>(g, start, width);
  return rc;
}

int launch_cover(pdf_grader *g, uint64_t start, uint64_t width, uint32_t *pair_total = nullptr) {
  Planes pl = planes(g, width);
  if (g->walk_planes) {  // the walk's mark-present bytes, not NZ
    pl.bt = walk_bits(g, g->walk_planes);
    pl.bt_shift = (uint32_t)__builtin_ctz(g->walk_planes);
    pl.bt_n = g->nl.n;
  }
  tm_begin(g, kEvents);
  k_events<<<grid_for(g->nl.n, 256), 256, 0, g->stream>>>(g->nl.n, g->d_first0, g->d_first1, (uint32_t)start,
                                                         (uint32_t)width, g->d_actmap, g->d_act_list, g->d_act_count);
  tm_end(g);
  tm_begin(g, kCover);
  k_cover<false><<<g->cover_waves, 64, 0, g->stream>>>(g->net(), pl, (uint32_t)start, g->d_list, g->d_list_cnt,
                                                       g->nl.n, g->d_act_list, g->d_act_count, g->d_first0,
                                                       g->d_first1, g->d_cs, g->d_hs, g->d_gate, g->d_chunk_total,
                                                       g->d_grand, g->d_err, (uint32_t)width, pair_total,
                                                       nullptr, g->pic_target, g->chunk_pairs, g->cover_caps);
  tm_end(g);
  HIPCHK(hipGetLastError());
  return PDF_OK;
}

// Segments of length K >= 2: the chunk's new-node plane from the segment
// firsts, the cover DP reading it, and the plane's words back to zero (it is
// all zero between chunks).
int launch_cover_segs(pdf_grader *g, uint64_t start, uint64_t width) {
  const Planes pl = planes(g, width);
  tm_begin(g, kEvents);
  k_seg_events<<<grid_for(g->seg_count, 256), 256, 0, g->stream>>>(
      g->d_seg, g->seg_k, g->seg_count, g->d_sfirst0, g->d_sfirst1, (uint32_t)start, (uint32_t)width, g->d_NEW,
      g->pitch, g->d_actmap, g->d_act_list, g->d_act_count);
  tm_end(g);
  tm_begin(g, kCover);
  k_cover<false, true><<<g->cover_waves, 64, 0, g->stream>>>(
      g->net(), pl, (uint32_t)start, g->d_list, g->d_list_cnt, g->nl.n, g->d_act_list, g->d_act_count, g->d_first0,
      g->d_first1, g->d_cs, g->d_hs, g->d_gate, g->d_chunk_total, g->d_grand, g->d_err, (uint32_t)width, nullptr,
      g->d_NEW, g->pic_target, g->chunk_pairs, g->cover_caps);
  tm_end(g);
  tm_begin(g, kEvents);
  k_seg_clear<<<grid_for(g->seg_count, 256), 256, 0, g->stream>>>(g->d_seg, g->seg_k, g->seg_count, g->d_sfirst0,
                                                                  g->d_sfirst1, (uint32_t)start, (uint32_t)width,
                                                                  g->d_NEW, g->pitch);
  tm_end(g);
  HIPCHK(hipGetLastError());
  return PDF_OK;
}

void drop_graph(pdf_grader *g);
int queue_seg_reset(pdf_grader *g);

// After a failed launch the new-node plane may not be all zero: free it (and
// any captured grade holding its address); the next use allocates it zeroed.
void drop_new_plane(pdf_grader *g) {
  drop_graph(g);
  dfree(g->d_NEW);
  for (auto &l : g->lanes) dfree(l.NEW);
}

constexpr uint64_t kMaxSegments = 1ULL << 27;  // segment list entries kept in HBM: 4 GiB at K = 8

void free_segments(pdf_grader *g) {
  dfree(g->d_seg);
  dfree(g->d_sfirst0);
  dfree(g->d_sfirst1);
  g->seg_k = 0;
  g->seg_count = 0;
}

// The segment list of length K (built once, kept until K changes), its
// first-pair arrays and the new-node plane.
int ensure_segments(pdf_grader *g, uint32_t K) {
  if (g->seg_k != K) {
    drop_graph(g);  // a captured grade holds the old list's addresses
    free_segments(g);
    const uint64_t S = count_segments(g->nl, K);
    if (S > kMaxSegments) {
      set_error("pdf_merge: " + std::to_string(S) + " segments of length " + std::to_string(K) +
                " exceed the supported " + std::to_string(kMaxSegments));
      return PDF_ERR_UNSUPPORTED;
    }
    std::vector<uint32_t> h(S * K);
    list_segments(g->nl, K, h.data(), S);
    int rc = dalloc(&g->d_seg, S * K);
    if (!rc) rc = dalloc(&g->d_sfirst0, S);
    if (!rc) rc = dalloc(&g->d_sfirst1, S);
    if (!rc && !hip_ok(hipMemcpy(g->d_seg, h.data(), S * K * sizeof(uint32_t), hipMemcpyHostToDevice), "hipMemcpy"))
      rc = PDF_ERR_HIP;
    if (rc) { free_segments(g); return rc; }
    g->seg_k = K;
    g->seg_count = S;
    if ((rc = queue_seg_reset(g))) return rc;
    HIPCHK(hipStreamSynchronize(g->stream));
  }
  // one new-node plane per lane, zeroed once here; every cover leaves it zero again
  const uint64_t words = (uint64_t)g->nl.n * g->pitch;
  auto alloc_new = [&](uint64_t *&p) -> int {
    if (p) return PDF_OK;
    int rc = dalloc(&p, words);
    if (!rc && !hip_ok(hipMemset(p, 0, words * sizeof(uint64_t)), "hipMemset")) rc = PDF_ERR_HIP;
    if (rc) dfree(p);
    return rc;
  };
  int rc = alloc_new(g->d_NEW);
  for (auto &l : g->lanes)
    if (!rc) rc = alloc_new(l.NEW);
  return rc;
}

template <bool SEQ>
void launch_merge_seg_t(pdf_grader *g, uint64_t start, uint64_t width) {
  const Planes pl = planes(g, width);
  // lanes per segment: the chunk's 64-word blocks rounded up to a power of two, at most 64
  const uint32_t nblk = (pl.words + 63) / 64;
  uint32_t lps_log = 0;
  while (lps_log < 6 && (1u << lps_log) < nblk) ++lps_log;
  const uint64_t waves = (g->seg_count + (64u >> lps_log) - 1) >> (6 - lps_log);
  k_merge_seg<SEQ><<<grid_for(waves, 4), 256, 0, g->stream>>>(pl, g->d_seg, g->seg_k, g->seg_count, lps_log,
                                                            (uint32_t)start, g->d_sfirst0, g->d_sfirst1);
}

// Segment firsts to PDF_NONE (queued; a kernel, so it may be graph-captured).
int queue_seg_reset(pdf_grader *g) {
  const uint32_t blocks = (uint32_t)std::min<uint64_t>(grid_for(g->seg_count, 256), 65536);
  tm_begin(g, kFill);
  k_fill_u32<<<blocks, 256, 0, g->stream>>>(g->d_sfirst0, g->seg_count, PDF_NONE);
  k_fill_u32<<<blocks, 256, 0, g->stream>>>(g->d_sfirst1, g->seg_count, PDF_NONE);
  tm_end(g);
  HIPCHK(hipGetLastError());
  return PDF_OK;
}

int launch_merge_seg(pdf_grader *g, uint64_t start, uint64_t width) {
  tm_begin(g, kSegMerge);
  if (g->seq) launch_merge_seg_t<true>(g, start, width);
  else launch_merge_seg_t<false>(g, start, width);
  tm_end(g);
  HIPCHK(hipGetLastError());
  return PDF_OK;
}

void free_rank(pdf_grader *g) {
  dfree(g->d_rk_keys);
  dfree(g->d_rk_idx);
  dfree(g->d_rk_perm);
  dfree(g->d_rk_va);
  dfree(g->d_rk_vb);
  dfree(g->d_rk_pm0);
  dfree(g->d_rk_pm1);
  if (g->d_rk_temp) (void)hipFree(g->d_rk_temp);
  g->d_rk_temp = nullptr;
  g->rk_pairs = g->rk_vsize = g->rk_temp_bytes = 0;
}

// Waits for a stream's queued work unless the stream is being captured
// (then nothing of it runs yet).
void drain(hipStream_t s) {
  if (!s) return;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) == hipSuccess && st == hipStreamCaptureStatusNone) (void)hipStreamSynchronize(s);
}

// Destroys the captured grades, after every lane's stream has finished what
// it queued: an executable graph destroyed while a stream still runs it
// faults the host inside the runtime.
void drop_graph(pdf_grader *g) {
  drain(g->stream);
  for (auto &l : g->lanes) drain(l.stream);
  if (g->graph) (void)hipGraphExecDestroy(g->graph);
  g->graph = nullptr;
  g->graph_robust = -1;
  for (hipGraphExec_t e : g->chunk_graphs) (void)hipGraphExecDestroy(e);
  g->chunk_graphs.clear();
  g->chunk_graphs_robust = -1;
}

void lane_save(const pdf_grader *g, pdf_grader::Lane &l) {
  l.stream = g->stream;
  l.V = g->d_V;
  l.VB = g->d_VB;
  l.M = g->d_M;
  l.NZ = g->d_NZ;
  l.ACC = g->d_ACC;
  l.VT = g->d_VT;
  l.list = g->d_list;
  l.cstate = g->d_cstate;
  l.act_list = g->d_act_list;
  l.actmap = g->d_actmap;
  l.chunk_total = g->d_chunk_total;
  l.list_cnt = g->d_list_cnt;
  l.act_count = g->d_act_count;
  l.cs = g->d_cs;
  l.hs = g->d_hs;
  l.NEW = g->d_NEW;
  l.rtl = g->d_rtl;
  l.rtl_cnt = g->d_rtl_cnt;
}

void lane_load(pdf_grader *g, const pdf_grader::Lane &l) {
  g->stream = l.stream;
  g->d_V = l.V;
  g->d_VB = l.VB;
  g->d_M = l.M;
  g->d_NZ = l.NZ;
  g->d_ACC = l.ACC;
  g->d_VT = l.VT;
  g->d_list = l.list;
  g->d_cstate = l.cstate;
  g->d_act_list = l.act_list;
  g->d_actmap = l.actmap;
  g->d_chunk_total = l.chunk_total;
  g->d_list_cnt = l.list_cnt;
  g->d_act_count = l.act_count;
  g->d_cs = l.cs;
  g->d_hs = l.hs;
  g->d_NEW = l.NEW;
  g->d_rtl = l.rtl;
  g->d_rtl_cnt = l.rtl_cnt;
}

void free_chunk_buffers(pdf_grader *g);
int alloc_chunk_buffers(pdf_grader *g, uint64_t chunk_pairs);

// Frees every lane's chunk buffers (the extra lanes' streams and events too).
// Every lane's queued work, graph launches included, finishes before any of
// its graph executables or buffers is released.
void free_chunk(pdf_grader *g) {
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  for (auto &l : g->lanes)
    if (l.stream) (void)hipStreamSynchronize(l.stream);
  drop_graph(g);
  if (!g->lanes.empty()) {
    pdf_grader::Lane l0;
    lane_save(g, l0);
    for (auto &l : g->lanes) {
      if (l.stream) (void)hipStreamSynchronize(l.stream);
      lane_load(g, l);
      free_chunk_buffers(g);
      if (l.stream) (void)hipStreamDestroy(l.stream);
      if (l.done) (void)hipEventDestroy(l.done);
    }
    g->lanes.clear();
    lane_load(g, l0);
  }
  free_chunk_buffers(g);
  dfree(g->d_dbg_M);
  dfree(g->d_dbg_NZ);
  dfree(g->d_dbg_pic);
  g->dbg_fwd_chunk = g->dbg_pic_chunk = -1;
  g->chunk_pairs = 0;
}

int alloc_chunk_buffers(pdf_grader *g, uint64_t chunk_pairs) {
  const uint64_t n = g->nl.n;
  const uint64_t words = (chunk_pairs + 63) / 64;
  int rc = dalloc(&g->d_V, n * g->pitch);
  if (!rc && !g->seq) rc = dalloc(&g->d_VB, n * g->pitch);
  if (!rc) rc = dalloc(&g->d_M, n * g->pitch);
  if (!rc) rc = dalloc(&g->d_NZ, n * g->ntiles);
  if (!rc) rc = dalloc(&g->d_ACC, n * g->pitch);
  // the removal accumulator is zero between passes (each reader clears it)
  if (!rc && !hip_ok(hipMemset(g->d_ACC, 0, n * g->pitch * 8), "hipMemset")) rc = PDF_ERR_HIP;
  if (!rc && g->tune_walk) rc = dalloc(&g->d_VT, walk_plane_words(g->seq, g->tune_walk, n, words));
  if (!rc) rc = dalloc(&g->d_list, n * words);
  const uint32_t ct_off = (uint32_t)((3 * words + 2) & ~1ULL);
  g->n_cstate = ct_off + 2;
  if (!rc) rc = dalloc(&g->d_cstate, g->n_cstate);
  if (!rc) {
    g->d_actmap = reinterpret_cast<uint64_t *>(g->d_cstate);
    g->d_list_cnt = g->d_cstate + 2 * words;
    g->d_act_count = g->d_cstate + 3 * words;
    g->d_chunk_total = reinterpret_cast<uint64_t *>(g->d_cstate + ct_off);
  }
  if (!rc) rc = dalloc(&g->d_act_list, chunk_pairs);
  if (!rc) rc = dalloc(&g->d_cs, n * g->cover_waves);
  if (!rc) rc = dalloc(&g->d_hs, n * g->cover_waves);
  if (!rc && g->tune_rev_compact) {
    // room for every (gate, 64-word tile) item above level 0, and each
    // level's sub-list counters
    rc = dalloc(&g->d_rtl, (n - g->nl.level_off[1]) * rev_slots_per_gate(g->ntiles));
    if (!rc) rc = dalloc(&g->d_rtl_cnt, ((uint64_t)g->nl.L + 1) * kRevSubLists * kRevCntStride);
  }
  g->rev_lists_buf = nullptr;
  return rc;
}

void free_chunk_buffers(pdf_grader *g) {
  dfree(g->d_NEW);
  dfree(g->d_V);
  dfree(g->d_VB);
  dfree(g->d_M);
  dfree(g->d_NZ);
  dfree(g->d_ACC);
  dfree(g->d_VT);
  g->walk_planes = 0;
  dfree(g->d_list);
  dfree(g->d_cstate);
  g->d_actmap = g->d_chunk_total = nullptr;
  g->d_list_cnt = g->d_act_count = nullptr;
  g->n_cstate = 0;
  dfree(g->d_act_list);
  dfree(g->d_cs);
  dfree(g->d_hs);
  dfree(g->d_rtl);
  dfree(g->d_rtl_cnt);
  g->rev_lists_buf = nullptr;
}

// Queue the reset of the cross-chunk state (no host synchronization).
int queue_reset(pdf_grader *g) {
  tm_begin(g, kFill);
  k_reset_grade<<<grid_for(std::max(g->nl.n, 2u), 256), 256, 0, g->stream>>>(g->nl.n, g->d_first0, g->d_first1,
                                                                            g->d_gate, g->d_grand, g->d_err);
  tm_end(g);
  HIPCHK(hipGetLastError());
  g->stage = 0;
  g->stage_chunk = -1;
  return PDF_OK;
}

// A new vector set or an explicit reset: node state, and the segment firsts
// if a segment list exists (grades with segments also reset them first).
int reset_state(pdf_grader *g) {
  int rc = queue_reset(g);
  if (!rc && g->seg_count) rc = queue_seg_reset(g);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(g->stream));
  return PDF_OK;
}

struct EventTimer {
  pdf_grader *g;
  float *ms;
  explicit EventTimer(pdf_grader *g_, float *ms_) : g(g_), ms(ms_) {
    if (ms) (void)hipEventRecord(g->ev0, g->stream);
  }
  int finish() {
    if (ms) {
      HIPCHK(hipEventRecord(g->ev1, g->stream));
      HIPCHK(hipEventSynchronize(g->ev1));
      HIPCHK(hipEventElapsedTime(ms, g->ev0, g->ev1));
    } else {
      HIPCHK(hipStreamSynchronize(g->stream));
    }
    return PDF_OK;
  }
};

// Copy rows [n][cols] of a pitched device plane to a dense host array.
int copy_rows(pdf_grader *g, uint64_t *dst, const uint64_t *src, uint64_t cols) {
  HIPCHK(hipMemcpy2D(dst, cols * 8, src, g->pitch * 8, cols * 8, g->nl.n, hipMemcpyDeviceToHost));
  return PDF_OK;
}

// The level walk's forward order: within each level, gates sorted by the
// work their lane takes -- single-input gates that pass their fanin's mark
// on, then gates that never mark (XOR/XNOR/NOT and unknown types with two or
// more fanins simulate to S0, simkernel.cuh:89-92), then AND/NAND/OR/NOR by
// fan-in -- ties by id.  A wave then holds gates of one kind, and the
// folds and sensitization terms no lane of it needs are branched over.
// (Putting the gates that read a fanin from outside the LDS window last,
// so that waves without such reads compute while the others wait, measured
// slower: 0.605 against 0.591 ms per 50k grade.)  The copies keep every
// field the walk reads; fout_off (fanouts are not read) holds the node's
// id.
// The per-level forward launches' gate order (PDF_TUNE_FWD_ORDER): within
// each level, a breadth-first walk over "shares a fanin with" from each
// unvisited gate in id order, so gates reading the same fanin row are
// dispatched close together and the row is read from HBM once and then from
// L2 (a level's waves run a few hundred gates at a time).  Records are
// returned in that order with fout_off holding the node id (the forward pass
// reads no fanouts).
std::vector<GateRec> fanin_order(const Netlist &nl, const std::vector<GateRec> &rec) {
  std::vector<GateRec> out(nl.n);
  std::vector<std::pair<uint32_t, uint32_t>> readers;  // (fanin, gate) of the level, sorted
  std::vector<uint8_t> seen;
  std::vector<uint32_t> queue;
  for (uint32_t k = 0; k <= nl.L; ++k) {
    const uint32_t lo = nl.level_off[k], hi = nl.level_off[k + 1];
    readers.clear();
    for (uint32_t g = lo; g < hi; ++g)
      for (uint32_t e = nl.fin_off[g]; e < nl.fin_off[g + 1]; ++e) readers.push_back({nl.fin[e], g});
    std::sort(readers.begin(), readers.end());
    seen.assign(hi - lo, 0);
    queue.clear();
    for (uint32_t g0 = lo; g0 < hi; ++g0) {
      if (seen[g0 - lo]) continue;
      seen[g0 - lo] = 1;
      size_t head = queue.size();
      queue.push_back(g0);
      while (head < queue.size()) {
        const uint32_t h = queue[head++];
        for (uint32_t e = nl.fin_off[h]; e < nl.fin_off[h + 1]; ++e) {
          auto it = std::lower_bound(readers.begin(), readers.end(), std::make_pair(nl.fin[e], 0u));
          for (; it != readers.end() && it->first == nl.fin[e]; ++it)
            if (!seen[it->second - lo]) {
              seen[it->second - lo] = 1;
              queue.push_back(it->second);
            }
        }
      }
    }
    for (uint32_t i = lo; i < hi; ++i) {  // queue holds every gate of the level once
      out[i] = rec[queue[i - lo]];
      out[i].fout_off = queue[i - lo];
    }
  }
  return out;
}

std::vector<GateRec> walk_order(const Netlist &nl, const std::vector<GateRec> &rec) {
  std::vector<GateRec> out(nl.n);
  std::vector<uint32_t> ids;
  for (uint32_t k = 0; k <= nl.L; ++k) {
    const uint32_t lo = nl.level_off[k], hi = nl.level_off[k + 1];
    ids.resize(hi - lo);
    for (uint32_t i = lo; i < hi; ++i) ids[i - lo] = i;
    auto key = [&](uint32_t i) -> uint32_t {
      const uint8_t ty = nl.type[i];
      const uint32_t nfi = rec[i].nfi;
      if (ty == PDF_INPT) return 0;
      const bool multi = ty >= PDF_AND && ty <= PDF_NOR;
      if (multi && nfi >= 2) return 10 + std::min<uint32_t>(nfi, 5);
      const bool passes = ty == PDF_BUFF || ty == PDF_DFF || ty == PDF_FROM || ty == PDF_NOT ||
                          (ty >= PDF_AND && ty <= PDF_XNOR && nfi == 1);
      return passes ? 1 : 2;
    };
    std::stable_sort(ids.begin(), ids.end(), [&](uint32_t a, uint32_t b) { return key(a) < key(b); });
    for (uint32_t i = lo; i < hi; ++i) {
      out[i] = rec[ids[i - lo]];
      out[i].fout_off = ids[i - lo];
    }
  }
  return out;
}

}  // namespace

extern "C" {

int pdf_grader_create(int device, const pdf_netlist *nlh, pdf_grader **out) {
  if (!nlh || !out) { set_error("pdf_grader_create: null argument"); return PDF_ERR_ARG; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    set_error("pdf_grader_create: no HIP device available");
    return PDF_ERR_HIP;
  }
  if (device < 0 || device >= ndev) { set_error("pdf_grader_create: bad device index"); return PDF_ERR_ARG; }
  HIPCHK(hipSetDevice(device));
  auto *g = new (std::nothrow) pdf_grader;
  if (!g) { set_error("out of memory"); return PDF_ERR_NOMEM; }
  g->device = device;
  g->nl = nlh->nl;
  const Netlist &nl = g->nl;
  int rc = PDF_OK;
  auto up = [&](auto *&dptr, const auto &vec) {
    if (rc) return;
    using T = std::remove_reference_t<decltype(vec[0])>;
    rc = dalloc(&dptr, vec.size());
    if (!rc && !vec.empty() &&
        !hip_ok(hipMemcpy(dptr, vec.data(), sizeof(T) * vec.size(), hipMemcpyHostToDevice), "hipMemcpy"))
      rc = PDF_ERR_HIP;
  };
  if (!hip_ok(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking), "hipStreamCreate") ||
      !hip_ok(hipEventCreate(&g->ev0), "hipEventCreate") || !hip_ok(hipEventCreate(&g->ev1), "hipEventCreate")) {
    pdf_grader_destroy(g);
    return PDF_ERR_HIP;
  }
  std::vector<GateRec> rec(nl.n);
  for (uint32_t i = 0; i < nl.n; ++i) {
    GateRec &r = rec[i];
    std::memset(&r, 0, sizeof(r));
    r.fin_off = nl.fin_off[i];
    r.fout_off = nl.fout_off[i];
    r.nfi = (uint16_t)(nl.fin_off[i + 1] - nl.fin_off[i]);
    r.nfo = (uint16_t)(nl.fout_off[i + 1] - nl.fout_off[i]);
    for (uint32_t j = 0; j < r.nfi && j < (uint32_t)kRegFanin; ++j) r.fin[j] = nl.fin[nl.fin_off[i] + j];
    r.type = nl.type[i];
    r.po = nl.po[i];
  }
  for (uint32_t i = nl.num_pi; i < nl.n; ++i) {
    const uint8_t ty = nl.type[i];
    const bool multi = ty >= PDF_AND && ty <= PDF_NOR;
    const uint32_t end = nl.level_off[nl.level[i] + 1];  // end of the reader's level
    for (uint32_t e = nl.fin_off[i]; e < nl.fin_off[i + 1]; ++e) {
      if (e != nl.fin_off[i] && !multi) break;
      const uint32_t f = nl.fin[e];
      rec[f].reach = (uint16_t)std::max<uint32_t>(rec[f].reach, std::min<uint32_t>(end - f, 65535u));
    }
  }
  up(g->d_rec, rec);
  up(g->d_wrec, walk_order(nl, rec));
  up(g->d_frec, fanin_order(nl, rec));
  for (uint32_t k = 0; k <= nl.L; ++k) g->widest = std::max(g->widest, nl.level_off[k + 1] - nl.level_off[k]);
  up(g->d_type, nl.type);
  up(g->d_po, nl.po);
  up(g->d_fin_off, nl.fin_off);
  up(g->d_fin, nl.fin);
  up(g->d_fout_off, nl.fout_off);
  up(g->d_fout, nl.fout);
  up(g->d_level, nl.level);
  up(g->d_level_off, nl.level_off);
  if (!rc) rc = dalloc(&g->d_first0, nl.n);
  if (!rc) rc = dalloc(&g->d_first1, nl.n);
  if (!rc) rc = dalloc(&g->d_gate, nl.n);
  if (!rc) rc = dalloc(&g->d_grand, 2);
  if (!rc) g->d_err = reinterpret_cast<uint32_t *>(g->d_grand + 1);
  if (!rc && hipHostMalloc((void **)&g->h_result, 2 * sizeof(uint64_t), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    g->h_result = nullptr;
    set_error("hipHostMalloc of the result words failed");
    rc = PDF_ERR_NOMEM;
  }
  if (!rc) rc = reset_state(g);
  if (rc) { pdf_grader_destroy(g); return rc; }
  *out = g;
  return PDF_OK;
}

void pdf_grader_destroy(pdf_grader *g) {
  if (!g) return;
  (void)hipSetDevice(g->device);
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  free_chunk(g);
  dfree(g->d_rec);
  dfree(g->d_type);
  dfree(g->d_po);
  dfree(g->d_fin_off);
  dfree(g->d_fin);
  dfree(g->d_fout_off);
  dfree(g->d_fout);
  dfree(g->d_level);
  dfree(g->d_level_off);
  dfree(g->d_wrec);
  dfree(g->d_frec);
  dfree(g->d_va);
  dfree(g->d_vb);
  free_rank(g);
  free_segments(g);
  dfree(g->d_first0);
  dfree(g->d_first1);
  dfree(g->d_gate);
  dfree(g->d_grand);
  g->d_err = nullptr;  // word 1 of d_grand
  if (g->h_result) (void)hipHostFree(g->h_result);
  for (hipEvent_t e : g->ev_pool) (void)hipEventDestroy(e);
  if (g->lane0_done) (void)hipEventDestroy(g->lane0_done);
  if (g->lane_start) (void)hipEventDestroy(g->lane_start);
  if (g->ev0) (void)hipEventDestroy(g->ev0);
  if (g->ev1) (void)hipEventDestroy(g->ev1);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
}

// Pairs per chunk when the caller gives none: all pairs if the chunk buffers
// fit in half of the free HBM, else the largest multiple of 64 that does (the
// reference sizes its chunk from free GPU memory, gpuCalculateSimulPatterns,
// main.cu:150).
static uint64_t default_chunk(pdf_grader *g) {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  // V, M, ACC (and VB) planes plus the marked-node lists, per pair
  const double per_pair = g->nl.n * ((g->seq ? 3.0 : 4.0) / 8.0 + 4.0 / 64.0) + 16.0;
  const double budget = 0.5 * (double)free_b - (double)(1ull << 30);
  if (budget <= 0) return 64;
  const uint64_t fit = (uint64_t)(budget / per_pair) / 64 * 64;
  return fit >= g->num_pairs ? 0 : std::max<uint64_t>(fit, 64);
}

static int load_vectors(pdf_grader *g, const uint64_t *a, const uint64_t *b, uint64_t num) {
  if (!g || !a) { set_error("null argument"); return PDF_ERR_ARG; }
  if (num == 0) { set_error("no pairs"); return PDF_ERR_ARG; }
  if (num > 0xFFFFFFFFull - 64) { set_error("at most 2^32-65 pairs (pattern ids are uint32)"); return PDF_ERR_ARG; }
  HIPCHK(hipSetDevice(g->device));
  HIPCHK(hipStreamSynchronize(g->stream));
  dfree(g->d_va);
  dfree(g->d_vb);
  free_chunk(g);
  g->num_pairs = 0;
  const uint64_t I = g->nl.num_pi;
  const uint64_t words = (num + 63) / 64;
  const uint32_t tail = (uint32_t)(num % 64);
  g->seq = b == nullptr;
  g->vstride = g->seq ? words + 1 : words;
  std::vector<uint64_t> host(I * g->vstride);
  auto bit = [&](const uint64_t *src, uint64_t i, uint64_t pos) { return (src[i * words + pos / 64] >> (pos % 64)) & 1; };
  auto stage = [&](const uint64_t *src) {
    for (uint64_t i = 0; i < I; ++i) {
      uint64_t *row = host.data() + i * g->vstride;
      std::copy(src + i * words, src + (i + 1) * words, row);
      if (tail) row[words - 1] &= (1ULL << tail) - 1;
      if (g->seq) {
        // positions past the last vector wrap to vector 0, 1, ...
        // (simkernel.cuh:116-120 pairs the last vector with vector 0)
        for (uint64_t pos = num; pos < 64 * (words + 1); ++pos)
          if (bit(src, i, (pos - num) % num)) row[pos / 64] |= 1ULL << (pos % 64);
      }
    }
  };
  int rc = dalloc(&g->d_va, host.size());
  if (!rc && b) rc = dalloc(&g->d_vb, host.size());
  if (rc) return rc;
  stage(a);
  HIPCHK(hipMemcpy(g->d_va, host.data(), host.size() * 8, hipMemcpyHostToDevice));
  if (b) {
    std::fill(host.begin(), host.end(), 0);
    stage(b);
    HIPCHK(hipMemcpy(g->d_vb, host.data(), host.size() * 8, hipMemcpyHostToDevice));
  }
  g->num_pairs = num;
  g->vwords = words;
  return pdf_grader_set_chunk(g, default_chunk(g));
}

int pdf_grader_set_vectors(pdf_grader *g, const uint64_t *bits, uint64_t num_vectors) {
  return load_vectors(g, bits, nullptr, num_vectors);
}

int pdf_grader_set_pairs(pdf_grader *g, const uint64_t *a_bits, const uint64_t *b_bits, uint64_t num_pairs) {
  if (!b_bits) { set_error("pdf_grader_set_pairs: null b_bits"); return PDF_ERR_ARG; }
  return load_vectors(g, a_bits, b_bits, num_pairs);
}

int pdf_grader_set_chunk(pdf_grader *g, uint64_t chunk_pairs) {
  if (!g) { set_error("null grader"); return PDF_ERR_ARG; }
  if (!g->num_pairs) { set_error("pdf_grader_set_chunk: load vectors first"); return PDF_ERR_STATE; }
  if (chunk_pairs == 0 || chunk_pairs >= g->num_pairs) chunk_pairs = g->num_pairs;
  else if (chunk_pairs % 64) { set_error("pdf_grader_set_chunk: chunk must be a multiple of 64 pairs"); return PDF_ERR_ARG; }
  HIPCHK(hipSetDevice(g->device));
  HIPCHK(hipStreamSynchronize(g->stream));
  free_chunk(g);
  const uint64_t words = (chunk_pairs + 63) / 64;
  g->pitch = (words + 1 + 15) / 16 * 16;  // + the sequence-mode extra column
  g->ntiles = (words + 63) / 64;
  const uint64_t n = g->nl.n;
  // cover scratch: one n-long c/h slice per wave, about 1 GiB at most
  const uint64_t per_wave = 2 * n * sizeof(uint64_t);
  g->cover_waves =
      (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g->max_cover_waves, (1ull << 30) / per_wave));
  int rc = alloc_chunk_buffers(g, chunk_pairs);
  // extra lanes: the same buffers again, with a stream and an event each
  const uint64_t nchunks = (g->num_pairs + chunk_pairs - 1) / chunk_pairs;
  const uint32_t nlanes = (uint32_t)std::min<uint64_t>(g->tune_lanes, nchunks);
  if (!rc && nlanes > 1) {
    pdf_grader::Lane l0;
    lane_save(g, l0);
    for (uint32_t i = 1; i < nlanes && !rc; ++i) {
      pdf_grader::Lane l;
      lane_load(g, pdf_grader::Lane{});
      rc = alloc_chunk_buffers(g, chunk_pairs);
      lane_save(g, l);
      l.stream = nullptr;
      if (!rc && !hip_ok(hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking), "hipStreamCreate")) rc = PDF_ERR_HIP;
      if (!rc && !hip_ok(hipEventCreateWithFlags(&l.done, hipEventDisableTiming), "hipEventCreate")) rc = PDF_ERR_HIP;
      g->lanes.push_back(l);  // freed by free_chunk even when incomplete
    }
    lane_load(g, l0);
  }
  if (!rc && nlanes > 1 && !g->lane0_done &&
      (!hip_ok(hipEventCreateWithFlags(&g->lane0_done, hipEventDisableTiming), "hipEventCreate") ||
       !hip_ok(hipEventCreateWithFlags(&g->lane_start, hipEventDisableTiming), "hipEventCreate")))
    rc = PDF_ERR_HIP;
  if (rc) { free_chunk(g); return rc; }
  g->chunk_pairs = chunk_pairs;
  return reset_state(g);
}


int pdf_grader_num_chunks(const pdf_grader *g, uint64_t *num_chunks) {
  if (!g || !num_chunks) { set_error("null argument"); return PDF_ERR_ARG; }
  if (!g->chunk_pairs) { set_error("no vectors / chunk configured"); return PDF_ERR_STATE; }
  *num_chunks = (g->num_pairs + g->chunk_pairs - 1) / g->chunk_pairs;
  return PDF_OK;
}

int pdf_grader_chunk_width(const pdf_grader *g, uint64_t chunk, uint64_t *pairs) {
  if (!g || !pairs) { set_error("null argument"); return PDF_ERR_ARG; }
  if (!g->chunk_pairs) { set_error("no vectors / chunk configured"); return PDF_ERR_STATE; }
  const uint64_t start = chunk * g->chunk_pairs;
  if (start >= g->num_pairs) { set_error("chunk index out of range"); return PDF_ERR_ARG; }
  *pairs = std::min(g->chunk_pairs, g->num_pairs - start);
  return PDF_OK;
}

int pdf_grader_reset(pdf_grader *g) {
  if (!g) { set_error("null grader"); return PDF_ERR_ARG; }
  HIPCHK(hipSetDevice(g->device));
  return reset_state(g);
}

int pdf_sim(pdf_grader *g, uint64_t chunk, uint64_t start, float *ms) {
  if (!g) { set_error("null grader"); return PDF_ERR_ARG; }
  uint64_t width;
  int rc = check_chunk(g, chunk, start, &width);
  if (rc) return rc;
  HIPCHK(hipSetDevice(g->device));
  EventTimer tm(g, ms);
  g->walk_planes = 0;
  if ((rc = launch_fwd<true, false>(g, start, width, 0))) return rc;
  if ((rc = tm.finish())) return rc;
  g->stage_chunk = (int64_t)chunk;
  g->stage = 1;
  return PDF_OK;
}

int pdf_mark(pdf_grader *g, uint64_t chunk, uint64_t start, int robust, float *ms) {
  if (!g) { set_error("null grader"); return PDF_ERR_ARG; }
  uint64_t width;
  int rc = check_chunk(g, chunk, start, &width);
  if (rc) return rc;
  if (g->stage_chunk != (int64_t)chunk || g->stage < 1) { set_error("pdf_mark: simulate this chunk first"); return PDF_ERR_STATE; }
  HIPCHK(hipSetDevice(g->device));
  if ((g->debug & PDF_DEBUG_FORWARD_MARKS) && !g->d_dbg_M) {
    if ((rc = dalloc(&g->d_dbg_M, (uint64_t)g->nl.n * g->pitch))) return rc;
    if ((rc = dalloc(&g->d_dbg_NZ, (uint64_t)g->nl.n * g->ntiles))) return rc;
  }
  EventTimer tm(g, ms);
  if ((rc = launch_fwd<false, true>(g, start, width, robust))) return rc;
  if (g->debug & PDF_DEBUG_FORWARD_MARKS) {  // markkernel.cuh:323-324 (gpumark-full.log)
    HIPCHK(hipMemcpyAsync(g->d_dbg_M, g->d_M, (uint64_t)g->nl.n * g->pitch * 8, hipMemcpyDeviceToDevice, g->stream));
    HIPCHK(hipMemcpyAsync(g->d_dbg_NZ, g->d_NZ, (uint64_t)g->nl.n * g->ntiles * 8, hipMemcpyDeviceToDevice,
                          g->stream));
    g->dbg_fwd_chunk = (int64_t)chunk;
  }
  if ((rc = launch_reverse<true, false>(g, start, width))) return rc;
  if ((rc = tm.finish())) return rc;
  g->stage = 2;
  g->stage_robust = robust;
  return PDF_OK;
}

int pdf_merge(pdf_grader *g, uint64_t chunk, uint64_t start, int segs, float *ms) {
  if (!g) { set_error("null grader"); return PDF_ERR_ARG; }
  if (segs < 1 || segs > 8) { set_error("pdf_merge: segment length must be 1..8"); return PDF_ERR_ARG; }
  uint64_t width;
  int rc = check_chunk(g, chunk, start, &width);
  if (rc) return rc;
  if (g->stage_chunk != (int64_t)chunk || g->stage < 2) { set_error("pdf_merge: mark this chunk first"); return PDF_ERR_STATE; }
  HIPCHK(hipSetDevice(g->device));
  if (segs >= 2 && (rc = ensure_segments(g, (uint32_t)segs))) return rc;
  EventTimer tm(g, ms);
  if ((rc = reset_chunk(g))) return rc;
  // per-node merge and the per-word lists of marked nodes (the cover's input)
  if ((rc = launch_reverse<false, true>(g, start, width))) return rc;
  if (segs >= 2 && (rc = launch_merge_seg(g, start, width))) return rc;
  if ((rc = tm.finish())) return rc;
  g->stage = 3;
  g->merged_segs = (uint32_t)segs;
  return PDF_OK;
}

int pdf_cover(pdf_grader *g, uint64_t chunk, uint64_t start, int segs, uint64_t *coverage, float *ms) {
  if (!g || !coverage) { set_error("null argument"); return PDF_ERR_ARG; }
  if (segs < 1 || segs > 8) { set_error("pdf_cover: segment length must be 1..8"); return PDF_ERR_ARG; }
  uint64_t width;
  int rc = check_chunk(g, chunk, start, &width);
  if (rc) return rc;
  // exactly once per merged chunk: a second call would count its pairs again
  if (g->stage_chunk != (int64_t)chunk || g->stage != 3) { set_error("pdf_cover: merge this chunk first"); return PDF_ERR_STATE; }
  if ((uint32_t)segs != g->merged_segs) { set_error("pdf_cover: segment length differs from the merge's"); return PDF_ERR_STATE; }
  HIPCHK(hipSetDevice(g->device));
  if (segs >= 2 && !g->d_NEW && (rc = ensure_segments(g, (uint32_t)segs))) return rc;
  const bool pic = g->debug & PDF_DEBUG_PI_COVER;
  if (pic && !g->d_dbg_pic && (rc = dalloc(&g->d_dbg_pic, (uint64_t)g->nl.num_pi * g->chunk_pairs))) return rc;
  EventTimer tm(g, ms);
  if (pic) {  // coverkernel.cuh:309-312 (gpucover.log): c of every PI per pair, zero where not covered
    const uint64_t nw = (uint64_t)g->nl.num_pi * g->chunk_pairs * 2;
    k_fill_u32<<<(uint32_t)std::min<uint64_t>(grid_for(nw, 256), 65536), 256, 0, g->stream>>>(
        reinterpret_cast<uint32_t *>(g->d_dbg_pic), nw, 0u);
    g->pic_target = g->d_dbg_pic;
  }
  rc = segs >= 2 ? launch_cover_segs(g, start, width) : launch_cover(g, start, width);
  g->pic_target = nullptr;
  if (rc) {
    if (segs >= 2) drop_new_plane(g);
    return rc;
  }
  if (pic) g->dbg_pic_chunk = (int64_t)chunk;
  if ((rc = tm.finish())) return rc;
  uint64_t tot = 0;
  uint32_t err = 0;
  HIPCHK(hipMemcpy(&tot, g->d_chunk_total, sizeof(tot), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&err, g->d_err, sizeof(err), hipMemcpyDeviceToHost));
  g->stage = 4;
  if (err & kErrOverflow) { set_error("pdf_cover: uint64 path count overflow"); return PDF_ERR_OVERFLOW; }
  *coverage = tot;
  return PDF_OK;
}

// Every launch of one fused grade (reset, then per chunk: forward levels,
// reverse levels, events, cover), in stream order.
// The fused grade on several lanes: chunk c runs on lane c % lanes, a group
// of `lanes` consecutive chunks at a time, their level launches issued
// interleaved (level k of every chunk of the group, then level k + 1), so
// the lanes' kernels overlap whether launched directly or replayed from a
// graph.  A chunk's cover needs the merges of every earlier chunk, so after
// its reverse pass the lane waits for the previous chunk's `done` and
// records its own (done of chunk c then covers all merges up to c).
// Results equal the one-lane order.
static int queue_grade_lanes(pdf_grader *g, int robust) {
  const uint32_t L = 1 + (uint32_t)g->lanes.size();
  const uint32_t K = g->grade_segs;
  std::vector<pdf_grader::Lane> ln(L);
  lane_save(g, ln[0]);
  ln[0].done = g->lane0_done;
  for (uint32_t i = 1; i < L; ++i) ln[i] = g->lanes[i - 1];
  int rc = queue_reset(g);
  if (!rc && K >= 2) rc = queue_seg_reset(g);
  if (rc) return rc;
  HIPCHK(hipEventRecord(g->lane_start, ln[0].stream));
  for (uint32_t i = 1; i < L; ++i) HIPCHK(hipStreamWaitEvent(ln[i].stream, g->lane_start, 0));
  const uint64_t nchunks = (g->num_pairs + g->chunk_pairs - 1) / g->chunk_pairs;
  for (uint64_t c0 = 0; c0 < nchunks && !rc; c0 += L) {
    const uint64_t c1 = std::min<uint64_t>(nchunks, c0 + L);
    auto chunk = [&](uint64_t c, uint64_t &start, uint64_t &width) {
      lane_load(g, ln[c % L]);
      start = c * g->chunk_pairs;
      width = std::min(g->chunk_pairs, g->num_pairs - start);
    };
    uint64_t start, width;
    if (g->tune_walk && K == 1 && g->d_VT) {
      // one forward and one reverse walk per chunk, each on its lane
      for (uint64_t c = c0; c < c1 && !rc; ++c) {
        chunk(c, start, width);
        rc = launch_walk(g, start, width, robust);
      }
    } else {
      if (!g->seq)
        for (uint64_t c = c0; c < c1 && !rc; ++c) {
          chunk(c, start, width);
          rc = reset_chunk(g);
        }
      for (uint32_t k = 0; k <= g->nl.L && !rc; ++k)
        for (uint64_t c = c0; c < c1 && !rc; ++c) {
          chunk(c, start, width);
          rc = launch_fwd<true, true>(g, start, width, robust, /*clear_state=*/true, 0xFFFFFFFFu, k, k);
        }
      if (g->tune_rev_walk) {
        for (uint64_t c = c0; c < c1 && !rc; ++c) {
          chunk(c, start, width);
          rc = launch_rev_walk(g, start, width, g->tune_rev_walk);
        }
      } else {
        for (uint32_t k = g->nl.L + 1; k-- > 0 && !rc;)
          for (uint64_t c = c0; c < c1 && !rc; ++c) {
            chunk(c, start, width);
            rc = launch_reverse<true, true>(g, start, width, k, k);
          }
      }
    }
    if (K >= 2)
      for (uint64_t c = c0; c < c1 && !rc; ++c) {
        chunk(c, start, width);
        rc = launch_merge_seg(g, start, width);
      }
    for (uint64_t c = c0; c < c1 && !rc; ++c) {
      chunk(c, start, width);
      const pdf_grader::Lane &l = ln[c % L];
      if (c > 0 && !hip_ok(hipStreamWaitEvent(l.stream, ln[(c - 1) % L].done, 0), "hipStreamWaitEvent"))
        rc = PDF_ERR_HIP;
      if (!rc && !hip_ok(hipEventRecord(l.done, l.stream), "hipEventRecord")) rc = PDF_ERR_HIP;
      if (!rc) rc = K >= 2 ? launch_cover_segs(g, start, width) : launch_cover(g, start, width);
    }
  }
  lane_load(g, ln[0]);
  // lane 0 waits for the other lanes' last chunks
  for (uint32_t i = 1; i < L && !rc; ++i) {
    if (!hip_ok(hipEventRecord(ln[i].done, ln[i].stream), "hipEventRecord") ||
        !hip_ok(hipStreamWaitEvent(ln[0].stream, ln[i].done, 0), "hipStreamWaitEvent"))
      rc = PDF_ERR_HIP;
  }
  return rc;
}

// Lanes with graphs: capture (once per robust flag) each chunk's forward and
// reverse levels on its lane's stream.
static int build_chunk_graphs(pdf_grader *g, int robust) {
  drop_graph(g);
  const uint32_t L = 1 + (uint32_t)g->lanes.size();
  std::vector<pdf_grader::Lane> ln(L);
  lane_save(g, ln[0]);
  for (uint32_t i = 1; i < L; ++i) ln[i] = g->lanes[i - 1];
  const uint64_t nchunks = (g->num_pairs + g->chunk_pairs - 1) / g->chunk_pairs;
  int rc = PDF_OK;
  for (uint64_t c = 0; c < nchunks && !rc; ++c) {
    lane_load(g, ln[c % L]);
    const uint64_t start = c * g->chunk_pairs;
    const uint64_t width = std::min(g->chunk_pairs, g->num_pairs - start);
    hipGraph_t graph = nullptr;
    if (!hip_ok(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture")) {
      rc = PDF_ERR_HIP;
      break;
    }
    rc = launch_passes(g, start, width, robust);
    const hipError_t e = hipStreamEndCapture(g->stream, &graph);
    if (!rc && !hip_ok(e, "hipStreamEndCapture")) rc = PDF_ERR_HIP;
    hipGraphExec_t exec = nullptr;
    if (!rc && !hip_ok(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate")) rc = PDF_ERR_HIP;
    if (graph) (void)hipGraphDestroy(graph);
    if (exec) g->chunk_graphs.push_back(exec);
  }
  lane_load(g, ln[0]);
  if (rc) {
    drop_graph(g);
    return rc;
  }
  g->chunk_graphs_robust = robust != 0;
  return PDF_OK;
}

// The fused grade on lanes from the per-chunk graphs: a group of `lanes`
// chunk graphs launched on their streams, then their covers in chunk order
// behind the `done` chain (as in queue_grade_lanes).
static int queue_grade_lane_graphs(pdf_grader *g, int robust) {
  if (g->chunk_graphs_robust != (robust != 0)) {
    int rc = build_chunk_graphs(g, robust);
    if (rc) return rc;
  }
  const uint32_t L = 1 + (uint32_t)g->lanes.size();
  const uint32_t K = g->grade_segs;
  std::vector<pdf_grader::Lane> ln(L);
  lane_save(g, ln[0]);
  ln[0].done = g->lane0_done;
  for (uint32_t i = 1; i < L; ++i) ln[i] = g->lanes[i - 1];
  int rc = queue_reset(g);
  if (!rc && K >= 2) rc = queue_seg_reset(g);
  if (rc) return rc;
  HIPCHK(hipEventRecord(g->lane_start, ln[0].stream));
  for (uint32_t i = 1; i < L; ++i) HIPCHK(hipStreamWaitEvent(ln[i].stream, g->lane_start, 0));
  const uint64_t nchunks = g->chunk_graphs.size();
  for (uint64_t c0 = 0; c0 < nchunks && !rc; c0 += L) {
    const uint64_t c1 = std::min<uint64_t>(nchunks, c0 + L);
    for (uint64_t c = c0; c < c1 && !rc; ++c)
      if (!hip_ok(hipGraphLaunch(g->chunk_graphs[c], ln[c % L].stream), "hipGraphLaunch")) rc = PDF_ERR_HIP;
    // segment merges behind each chunk's graph, on its lane (atomic, so the
    // lanes' merges may overlap)
    for (uint64_t c = c0; c < c1 && !rc && K >= 2; ++c) {
      lane_load(g, ln[c % L]);
      const uint64_t start = c * g->chunk_pairs;
      rc = launch_merge_seg(g, start, std::min(g->chunk_pairs, g->num_pairs - start));
    }
    for (uint64_t c = c0; c < c1 && !rc; ++c) {
      const pdf_grader::Lane &l = ln[c % L];
      lane_load(g, l);
      const uint64_t start = c * g->chunk_pairs;
      const uint64_t width = std::min(g->chunk_pairs, g->num_pairs - start);
      if (c > 0 && !hip_ok(hipStreamWaitEvent(l.stream, ln[(c - 1) % L].done, 0), "hipStreamWaitEvent"))
        rc = PDF_ERR_HIP;
      if (!rc && !hip_ok(hipEventRecord(l.done, l.stream), "hipEventRecord")) rc = PDF_ERR_HIP;
      if (!rc) rc = K >= 2 ? launch_cover_segs(g, start, width) : launch_cover(g, start, width);
    }
  }
  lane_load(g, ln[0]);
  for (uint32_t i = 1; i < L && !rc; ++i) {
    if (!hip_ok(hipEventRecord(ln[i].done, ln[i].stream), "hipEventRecord") ||
        !hip_ok(hipStreamWaitEvent(ln[0].stream, ln[i].done, 0), "hipStreamWaitEvent"))
      rc = PDF_ERR_HIP;
  }
  return rc;
}

// With segments of length K >= 2 (g->grade_segs) each chunk's reverse pass
// is followed by the segment merge and the cover reads the new-node plane
// built from the segment firsts.
static int queue_grade(pdf_grader *g, int robust) {
  const uint32_t K = g->grade_segs;
  if (!g->lanes.empty() && !g->timing && !g->one_lane) return queue_grade_lanes(g, robust);
  int rc = queue_reset(g);
  if (rc) return rc;
  if (K >= 2 && (rc = queue_seg_reset(g))) return rc;
  const uint64_t nchunks = (g->num_pairs + g->chunk_pairs - 1) / g->chunk_pairs;
  for (uint64_t c = 0; c < nchunks; ++c) {
    const uint64_t start = c * g->chunk_pairs;
    const uint64_t width = std::min(g->chunk_pairs, g->num_pairs - start);
    if ((rc = launch_passes(g, start, width, robust))) return rc;
    if (K >= 2) {
      if ((rc = launch_merge_seg(g, start, width))) return rc;
      if ((rc = launch_cover_segs(g, start, width))) return rc;
    } else if ((rc = launch_cover(g, start, width))) {
      return rc;
    }
  }
  return PDF_OK;
}

static int grade_impl(pdf_grader *g, int robust, uint64_t *total) {
  if (!g || !total) { set_error("null argument"); return PDF_ERR_ARG; }
  if (!g->chunk_pairs) { set_error("pdf_grade: load vectors first"); return PDF_ERR_STATE; }
  HIPCHK(hipSetDevice(g->device));
  int rc;
  const uint64_t nchunks = (g->num_pairs + g->chunk_pairs - 1) / g->chunk_pairs;
  const int key = (robust != 0) | (int)(g->grade_segs << 1);
  if (g->tune_use_graph && !g->timing && !g->lanes.empty()) {
    if ((rc = queue_grade_lane_graphs(g, robust))) return rc;
  } else if (g->tune_use_graph && !g->timing) {
    // the launch sequence depends only on the chunk shape, the robust flag
    // and the segment length, so it is captured once and replayed
    if (!g->graph || g->graph_robust != key) {
      drop_graph(g);
      hipGraph_t graph = nullptr;
      HIPCHK(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
      rc = queue_grade(g, robust);
      const hipError_t e = hipStreamEndCapture(g->stream, &graph);
      if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      HIPCHK(e);
      const hipError_t ie = hipGraphInstantiate(&g->graph, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ie != hipSuccess) { g->graph = nullptr; HIPCHK(ie); }
      g->graph_robust = key;
    }
    HIPCHK(hipGraphLaunch(g->graph, g->stream));
  } else {
    if ((rc = queue_grade(g, robust))) return rc;
    tm_close(g);
  }
  HIPCHK(hipMemcpyAsync(g->h_result, g->d_grand, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  const uint64_t tot = g->h_result[0];
  const uint32_t err = (uint32_t)g->h_result[1];  // d_err is the low half of word 1
  g->stage = 4;
  g->stage_chunk = (int64_t)nchunks - 1;
  g->stage_robust = robust;
  g->merged_segs = g->grade_segs;
  if (err & kErrOverflow) { set_error("pdf_grade: uint64 path count overflow"); return PDF_ERR_OVERFLOW; }
  *total = tot;
  return PDF_OK;
}

int pdf_grade(pdf_grader *g, int robust, uint64_t *total) { return grade_impl(g, robust, total); }

int pdf_grader_set_merge_window(pdf_grader *g, uint32_t pairs) {
  if (!g) { set_error("null grader"); return PDF_ERR_ARG; }
  HIPCHK(hipSetDevice(g->device));
  HIPCHK(hipStreamSynchronize(g->stream));
  g->merge_window = pairs;
  drop_graph(g);  // captured launches hold the old window
  return PDF_OK;
}

uint64_t pdf_ref_unmerged_segments(uint64_t S) {
  // gpuMergeSegments (mergekernel.cuh:229-250): ceil(S / 20) blocks of 20
  // segments; launch i runs min(65535, blocks left) of them starting at
  // segment 65535 i (count += BLOCK_STEP), so it reaches segments
  // [65535 i, 65535 i + 20 b_i); the blocks left drop by 65535 per launch
  const uint64_t step = 65535, per = 20;
  uint64_t left = (S + per - 1) / per, covered = 0, reach = 0;
  for (uint64_t i = 0; left > 0; ++i) {
    const uint64_t b = std::min(left, step);
    const uint64_t lo = std::min(S, i * step), hi = std::min(S, i * step + per * b);
    if (hi > reach) {  // launches start in increasing order: count what extends past the reach so far
      covered += hi - std::max(lo, reach);
      reach = hi;
    }
    left = left >= step ? left - step : 0;
  }
  return S - covered;
}

int pdf_grader_ref_flags(const pdf_grader *g, uint32_t *flags, uint64_t *unmerged_segments) {
  if (!g || !flags) { set_error("null argument"); return PDF_ERR_ARG; }
  HIPCHK(hipSetDevice(g->device));
  uint32_t err = 0;
  HIPCHK(hipMemcpy(&err, g->d_err, sizeof(err), hipMemcpyDeviceToHost));
  const uint64_t S = g->merged_segs >= 2 ? g->seg_count : g->nl.n;
  const uint64_t un = pdf_ref_unmerged_segments(S);
  *flags = ((err & kErrRefU16) ? PDF_REF_EXCEEDS_U16 : 0u) | ((err & kErrRefU32) ? PDF_REF_EXCEEDS_U32 : 0u) |
           (un ? PDF_REF_SEGMENT_STRIDE : 0u) |
           (g->merge_window == 0 && g->chunk_pairs > 4224 ? PDF_REF_MERGE_WINDOW : 0u);
  if (unmerged_segments) *unmerged_segments = un;
  return PDF_OK;
}

int pdf_grader_set_tuning(pdf_grader *g, int key, int64_t value) {
  if (!g) { set_error("null grader"); return PDF_ERR_ARG; }
  switch (key) {
    case PDF_TUNE_LANES: {
      if (value < 1 || value > 4) { set_error("lanes must be 1..4"); return PDF_ERR_ARG; }
      g->tune_lanes = (uint32_t)value;
      return g->chunk_pairs ? pdf_grader_set_chunk(g, g->chunk_pairs) : PDF_OK;
    }
    case PDF_TUNE_WORDS_PER_LANE:
      if (value != 1 && value != 2) { set_error("words per lane must be 1 or 2"); return PDF_ERR_ARG; }
      g->tune_words_per_lane = (uint32_t)value;
      break;
    case PDF_TUNE_USE_GRAPH:
      g->tune_use_graph = value != 0;
      break;
    case PDF_TUNE_REV_WALK:
      if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8) {
        set_error("words per reverse-walk tile must be 0 (per-level launches), 1, 2, 4 or 8");
        return PDF_ERR_ARG;
      }
      g->tune_rev_walk = (uint32_t)value;
      break;
    case PDF_TUNE_COVER_NODES:
      if (value < 1 || value > (int64_t)kCoverCap) { set_error("cover LDS nodes must be 1..1280"); return PDF_ERR_ARG; }
      g->cover_caps.nodes = (uint32_t)value;
      break;
    case PDF_TUNE_COVER_POOL:
      if (value < 0 || value > (int64_t)kPool) { set_error("cover fanout pool must be 0..4096"); return PDF_ERR_ARG; }
      g->cover_caps.pool = (uint32_t)value;
      break;
    case PDF_TUNE_WALK:
      if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8) {
        set_error("words per walk tile must be 0 (per-level launches), 1, 2, 4 or 8");
        return PDF_ERR_ARG;
      }
      if (g->tune_walk == (uint32_t)value) break;
      g->tune_walk = (uint32_t)value;
      // the walk's tile-major plane is one of the chunk buffers
      return g->chunk_pairs ? pdf_grader_set_chunk(g, g->chunk_pairs) : PDF_OK;
    case PDF_TUNE_REV_TILES:
      if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8) {
        set_error("reverse tiles per wave must be 0 (automatic), 1, 2, 4 or 8");
        return PDF_ERR_ARG;
      }
      g->tune_rev_tiles = (uint32_t)value;
      break;
    case PDF_TUNE_FWD_ORDER:
      if (value != 0 && value != 1) { set_error("forward gate order must be 0 (ids) or 1 (fanin locality)"); return PDF_ERR_ARG; }
      g->tune_fwd_order = (uint32_t)value;
      break;
    case PDF_TUNE_REV_COMPACT:
      if (value != 0 && value != 1) { set_error("compacted reverse pass must be 0 or 1"); return PDF_ERR_ARG; }
      if (g->tune_rev_compact == (uint32_t)value) break;
      g->tune_rev_compact = (uint32_t)value;
      // the item lists are chunk buffers
      return g->chunk_pairs ? pdf_grader_set_chunk(g, g->chunk_pairs) : PDF_OK;
    default:
      set_error("unknown tuning key");
      return PDF_ERR_ARG;
  }
  HIPCHK(hipSetDevice(g->device));
  HIPCHK(hipStreamSynchronize(g->stream));
  drop_graph(g);
  return PDF_OK;
}

static int grade_timed_impl(pdf_grader *g, int robust, uint64_t *total, pdf_kernel_times *times) {
  g->ev_family.clear();
  g->ev_count.clear();
  g->tm_open = false;
  g->timing = true;
  const int rc = grade_impl(g, robust, total);
  tm_close(g);
  g->timing = false;
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(g->stream));
  std::memset(times, 0, sizeof(*times));
  double *ms[6] = {&times->fwd_ms, &times->reverse_ms, &times->events_ms, &times->cover_ms, &times->fill_ms,
                   &times->merge_ms};
  uint32_t *cnt[6] = {&times->fwd_launches, &times->reverse_launches, &times->events_launches,
                      &times->cover_launches, &times->fill_launches, &times->merge_launches};
  for (size_t i = 0; i < g->ev_family.size(); ++i) {
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, g->ev_pool[2 * i], g->ev_pool[2 * i + 1]));
    *ms[g->ev_family[i]] += t;
    *cnt[g->ev_family[i]] += g->ev_count[i];
  }
  if (!g->ev_family.empty()) {
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, g->ev_pool[0], g->ev_pool[2 * g->ev_family.size() - 1]));
    times->wall_ms = t;
  }
  return PDF_OK;
}

int pdf_grade_timed(pdf_grader *g, int robust, uint64_t *total, pdf_kernel_times *times) {
  if (!g || !total || !times) { set_error("null argument"); return PDF_ERR_ARG; }
  HIPCHK(hipSetDevice(g->device));
  return grade_timed_impl(g, robust, total, times);
}

int pdf_grade_async(pdf_grader *g, int robust, int segs, void *stream, uint64_t *total_dev, uint32_t *err_dev) {
  if (!g || !stream || !total_dev) { set_error("null argument"); return PDF_ERR_ARG; }
  if (segs < 1 || segs > 8) { set_error("pdf_grade_async: segment length must be 1..8"); return PDF_ERR_ARG; }
  if (!g->chunk_pairs) { set_error("pdf_grade_async: load vectors first"); return PDF_ERR_STATE; }
  if (segs >= 2 && (g->seg_k != (uint32_t)segs || !g->d_NEW)) {
    set_error("pdf_grade_async: no segment list of this length yet (one pdf_grade_segs call sets it up)");
    return PDF_ERR_STATE;
  }
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev != g->device) HIPCHK(hipSetDevice(g->device));
  // everything below only queues kernels on the caller's stream: no
  // allocation, no synchronous copy, no wait, no memset node
  const hipStream_t own = g->stream;
  g->stream = static_cast<hipStream_t>(stream);
  g->grade_segs = (uint32_t)segs;
  g->one_lane = true;
  int rc = queue_grade(g, robust);
  if (!rc) {
    k_copy_result<<<1, 1, 0, g->stream>>>(g->d_grand, g->d_err, total_dev, err_dev);
    if (!hip_ok(hipGetLastError(), "k_copy_result")) rc = PDF_ERR_HIP;
  }
  g->stream = own;
  g->grade_segs = 1;
  g->one_lane = false;
  if (rc) return rc;
  const uint64_t nchunks = (g->num_pairs + g->chunk_pairs - 1) / g->chunk_pairs;
  g->stage = 4;
  g->stage_chunk = (int64_t)nchunks - 1;
  g->stage_robust = robust;
  g->merged_segs = (uint32_t)segs;
  return PDF_OK;
}

// The four stages and the reset queued on a caller's stream: the same
// launches as the synchronous calls with the grader's stream swapped for the
// caller's, and nothing else (no allocation, no synchronous copy, no wait, no
// memset node), so a caller can capture them into its own HIP graph.
namespace {
struct OnStream {
  pdf_grader *g;
  hipStream_t own;
  OnStream(pdf_grader *g_, void *stream) : g(g_), own(g_->stream) { g->stream = static_cast<hipStream_t>(stream); }
  ~OnStream() { g->stream = own; }
};

int async_begin(pdf_grader *g, void *stream, const char *what) {
  if (!g || !stream) { set_error(std::string(what) + ": null argument"); return PDF_ERR_ARG; }
  if (g->debug) {
    set_error(std::string(what) + ": debug dumps need the synchronous stage calls");
    return PDF_ERR_STATE;
  }
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev != g->device) HIPCHK(hipSetDevice(g->device));
  return PDF_OK;
}
}  // namespace

int pdf_grader_reset_async(pdf_grader *g, void *stream) {
  int rc = async_begin(g, stream, "pdf_grader_reset_async");
  if (rc) return rc;
  OnStream on(g, stream);
  rc = queue_reset(g);
  if (!rc && g->seg_count) rc = queue_seg_reset(g);
  return rc;
}

int pdf_sim_async(pdf_grader *g, uint64_t chunk, uint64_t start, void *stream) {
  int rc = async_begin(g, stream, "pdf_sim_async");
  if (rc) return rc;
  uint64_t width;
  if ((rc = check_chunk(g, chunk, start, &width))) return rc;
  OnStream on(g, stream);
  g->walk_planes = 0;
  if ((rc = launch_fwd<true, false>(g, start, width, 0))) return rc;
  g->stage_chunk = (int64_t)chunk;
  g->stage = 1;
  return PDF_OK;
}

int pdf_mark_async(pdf_grader *g, uint64_t chunk, uint64_t start, int robust, void *stream) {
  int rc = async_begin(g, stream, "pdf_mark_async");
  if (rc) return rc;
  uint64_t width;
  if ((rc = check_chunk(g, chunk, start, &width))) return rc;
  if (g->stage_chunk != (int64_t)chunk || g->stage < 1) {
    set_error("pdf_mark_async: simulate this chunk first");
    return PDF_ERR_STATE;
  }
  OnStream on(g, stream);
  if ((rc = launch_fwd<false, true>(g, start, width, robust))) return rc;
  if ((rc = launch_reverse<true, false>(g, start, width))) return rc;
  g->stage = 2;
  g->stage_robust = robust;
  return PDF_OK;
}

int pdf_merge_async(pdf_grader *g, uint64_t chunk, uint64_t start, int segs, void *stream) {
  int rc = async_begin(g, stream, "pdf_merge_async");
  if (rc) return rc;
  if (segs < 1 || segs > 8) { set_error("pdf_merge_async: segment length must be 1..8"); return PDF_ERR_ARG; }
  uint64_t width;
  if ((rc = check_chunk(g, chunk, start, &width))) return rc;
  if (g->stage_chunk != (int64_t)chunk || g->stage < 2) {
    set_error("pdf_merge_async: mark this chunk first");
    return PDF_ERR_STATE;
  }
  if (segs >= 2 && (g->seg_k != (uint32_t)segs || !g->d_NEW)) {
    set_error("pdf_merge_async: no segment list of this length yet (one pdf_merge or pdf_grade_segs call sets it "
              "up)");
    return PDF_ERR_STATE;
  }
  OnStream on(g, stream);
  if ((rc = reset_chunk(g))) return rc;
  if ((rc = launch_reverse<false, true>(g, start, width))) return rc;
  if (segs >= 2 && (rc = launch_merge_seg(g, start, width))) return rc;
  g->stage = 3;
  g->merged_segs = (uint32_t)segs;
  return PDF_OK;
}

int pdf_cover_async(pdf_grader *g, uint64_t chunk, uint64_t start, int segs, void *stream,
                    uint64_t *chunk_total_dev, uint64_t *per_gate_accum_dev, uint32_t *err_dev) {
  int rc = async_begin(g, stream, "pdf_cover_async");
  if (rc) return rc;
  if (!chunk_total_dev) { set_error("pdf_cover_async: null chunk_total_dev"); return PDF_ERR_ARG; }
  if (segs < 1 || segs > 8) { set_error("pdf_cover_async: segment length must be 1..8"); return PDF_ERR_ARG; }
  uint64_t width;
  if ((rc = check_chunk(g, chunk, start, &width))) return rc;
  if (g->stage_chunk != (int64_t)chunk || g->stage != 3) {
    set_error("pdf_cover_async: merge this chunk first");
    return PDF_ERR_STATE;
  }
  if ((uint32_t)segs != g->merged_segs) {
    set_error("pdf_cover_async: segment length differs from the merge's");
    return PDF_ERR_STATE;
  }
  OnStream on(g, stream);
  rc = segs >= 2 ? launch_cover_segs(g, start, width) : launch_cover(g, start, width);
  if (rc) return rc;
  k_copy_result<<<1, 1, 0, g->stream>>>(g->d_chunk_total, g->d_err, chunk_total_dev, err_dev);
  HIPCHK(hipGetLastError());
  if (per_gate_accum_dev)
    HIPCHK(hipMemcpyAsync(per_gate_accum_dev, g->d_gate, (uint64_t)g->nl.n * sizeof(uint64_t),
                          hipMemcpyDeviceToDevice, g->stream));
  g->stage = 4;
  return PDF_OK;
}

}  // extern "C"

// ------------------------------------------- grading from a static scratch
// pdf_grade_noalloc carves every chunk buffer (value, mark, present and
// accumulator planes, node lists, counters, the cover's c/h fallback) out
// of one file-scope device array instead of the grader's own allocations.
// Each device holds its own copy of the array for the life of the process;
// host code reaches it only through the address hipGetSymbolAddress gives on
// that device (the array's name is a host-side stand-in).  All calls on one
// device share the array, so calls that overlap on one device race.
namespace {

static __device__ __align__(256) uint64_t scratch[PDF_NOALLOC_SCRATCH_BYTES / 8];

constexpr int kScratchDevices = 64;
std::mutex scratch_mu;
uint64_t *scratch_dev[kScratchDevices] = {};

// The scratch array's address on the current device `dev` (looked up once
// per device, with that device current).
int scratch_address(int dev, uint64_t **out) {
  if (dev < 0 || dev >= kScratchDevices) {
    set_error("pdf_grade_noalloc: device index " + std::to_string(dev) + " beyond the supported " +
              std::to_string(kScratchDevices));
    return PDF_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      set_error(std::string("pdf_grade_noalloc: hipGetSymbolAddress of the scratch array failed: ") +
                hipGetErrorString(e));
      return PDF_ERR_HIP;
    }
    scratch_dev[dev] = static_cast<uint64_t *>(p);
  }
  *out = scratch_dev[dev];
  return PDF_OK;
}

// Where one chunk's buffers sit in the scratch array (byte offsets, each
// 256-byte aligned) for `chunk` pairs per chunk and `waves` cover waves.
struct ScratchPlan {
  uint64_t chunk = 0, words = 0, pitch = 0, ntiles = 0;
  uint32_t waves = 0, n_cstate = 0;
  uint64_t V = 0, VB = 0, M = 0, NZ = 0, ACC = 0, list = 0, cstate = 0, act = 0, cs = 0, hs = 0, bytes = 0;
};

ScratchPlan scratch_layout(uint64_t n, bool seq, uint64_t chunk, uint32_t waves) {
  ScratchPlan p;
  p.chunk = chunk;
  p.waves = waves;
  p.words = (chunk + 63) / 64;
  p.pitch = (p.words + 1 + 15) / 16 * 16;  // as pdf_grader_set_chunk lays out the planes
  p.ntiles = (p.words + 63) / 64;
  p.n_cstate = (uint32_t)(((3 * p.words + 2) & ~1ULL) + 2);
  uint64_t at = 0;
  auto take = [&](uint64_t bytes) {
    const uint64_t o = at;
    at += (bytes + 255) / 256 * 256;
    return o;
  };
  p.V = take(n * p.pitch * 8);
  p.VB = seq ? 0 : take(n * p.pitch * 8);
  p.M = take(n * p.pitch * 8);
  p.NZ = take(n * p.ntiles * 8);
  p.ACC = take(n * p.pitch * 8);
  p.list = take(n * p.words * 4);
  p.cstate = take((uint64_t)p.n_cstate * 4);
  p.act = take(chunk * 4);
  p.cs = take(n * waves * 8);
  p.hs = take(n * waves * 8);
  p.bytes = at;
  return p;
}

// The largest chunk (a multiple of 64 pairs, or all of them) whose buffers
// fit in the scratch array, with as many cover waves (64 down to 1) as fit.
bool plan_scratch(uint64_t n, bool seq, uint64_t pairs, ScratchPlan *out) {
  for (uint32_t waves : {64u, 16u, 4u, 1u}) {
    if (scratch_layout(n, seq, 64, waves).bytes > PDF_NOALLOC_SCRATCH_BYTES) continue;
    uint64_t lo = 1, hi = (pairs + 63) / 64;  // words per chunk
    while (lo < hi) {
      const uint64_t mid = (lo + hi + 1) / 2;
      if (scratch_layout(n, seq, mid * 64, waves).bytes <= PDF_NOALLOC_SCRATCH_BYTES) lo = mid;
      else hi = mid - 1;
    }
    *out = scratch_layout(n, seq, std::min(lo * 64, pairs), waves);
    return true;
  }
  return false;
}

}  // namespace

extern "C" {

int pdf_grade_noalloc(pdf_grader *g, int robust, void *stream, uint64_t *total_dev, uint32_t *err_dev) {
  if (!g || !stream || !total_dev) { set_error("null argument"); return PDF_ERR_ARG; }
  if (!g->num_pairs) { set_error("pdf_grade_noalloc: load vectors first"); return PDF_ERR_STATE; }
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev != g->device) HIPCHK(hipSetDevice(g->device));
  uint64_t *base = nullptr;
  int rc = scratch_address(g->device, &base);
  if (rc) return rc;
  ScratchPlan p;
  if (!plan_scratch(g->nl.n, g->seq, g->num_pairs, &p)) {
    set_error("pdf_grade_noalloc: " + std::to_string(g->nl.n) + " nodes need " +
              std::to_string(scratch_layout(g->nl.n, g->seq, 64, 1).bytes) +
              " bytes of chunk buffers even at 64 pairs per chunk; the static scratch holds " +
              std::to_string((unsigned long long)PDF_NOALLOC_SCRATCH_BYTES));
    return PDF_ERR_UNSUPPORTED;
  }
  // the grader's chunk fields point into the scratch array for this call only
  struct Saved {
    pdf_grader::Lane lane;
    uint64_t chunk_pairs, pitch, ntiles;
    uint32_t n_cstate, cover_waves, grade_segs;
    bool one_lane;
  } sv;
  lane_save(g, sv.lane);
  sv.chunk_pairs = g->chunk_pairs;
  sv.pitch = g->pitch;
  sv.ntiles = g->ntiles;
  sv.n_cstate = g->n_cstate;
  sv.cover_waves = g->cover_waves;
  sv.grade_segs = g->grade_segs;
  sv.one_lane = g->one_lane;
  char *b = reinterpret_cast<char *>(base);
  g->stream = static_cast<hipStream_t>(stream);
  g->d_V = reinterpret_cast<uint64_t *>(b + p.V);
  g->d_VB = g->seq ? nullptr : reinterpret_cast<uint64_t *>(b + p.VB);
  g->d_M = reinterpret_cast<uint64_t *>(b + p.M);
  g->d_NZ = reinterpret_cast<uint64_t *>(b + p.NZ);
  g->d_ACC = reinterpret_cast<uint64_t *>(b + p.ACC);
  g->d_list = reinterpret_cast<uint32_t *>(b + p.list);
  g->d_cstate = reinterpret_cast<uint32_t *>(b + p.cstate);
  g->d_actmap = reinterpret_cast<uint64_t *>(g->d_cstate);
  g->d_list_cnt = g->d_cstate + 2 * p.words;
  g->d_act_count = g->d_cstate + 3 * p.words;
  g->d_chunk_total = reinterpret_cast<uint64_t *>(g->d_cstate + (((3 * p.words + 2) & ~1ULL)));
  g->d_act_list = reinterpret_cast<uint32_t *>(b + p.act);
  g->d_cs = reinterpret_cast<uint64_t *>(b + p.cs);
  g->d_hs = reinterpret_cast<uint64_t *>(b + p.hs);
  g->d_NEW = nullptr;
  g->chunk_pairs = p.chunk;
  g->pitch = p.pitch;
  g->ntiles = p.ntiles;
  g->n_cstate = p.n_cstate;
  g->cover_waves = p.waves;
  g->d_VT = nullptr;  // no tile-major plane in the scratch: per-level launches
  g->d_rtl = nullptr;  // nor item lists: dense reverse launches
  g->d_rtl_cnt = nullptr;
  g->grade_segs = 1;
  g->one_lane = true;
  // the removal accumulator must start at zero and the array keeps whatever
  // the last call on this device left there; the mark-present words and the
  // per-chunk counters are rewritten by the passes themselves
  const uint64_t acc_words = (uint64_t)g->nl.n * p.pitch * 2;
  k_fill_u32<<<(uint32_t)std::min<uint64_t>(grid_for(acc_words, 256), 65536), 256, 0, g->stream>>>(
      reinterpret_cast<uint32_t *>(g->d_ACC), acc_words, 0u);
  if (!hip_ok(hipGetLastError(), "k_fill_u32")) rc = PDF_ERR_HIP;
  if (!rc) rc = queue_grade(g, robust);
  if (!rc) {
    k_copy_result<<<1, 1, 0, g->stream>>>(g->d_grand, g->d_err, total_dev, err_dev);
    if (!hip_ok(hipGetLastError(), "k_copy_result")) rc = PDF_ERR_HIP;
  }
  lane_load(g, sv.lane);
  g->chunk_pairs = sv.chunk_pairs;
  g->pitch = sv.pitch;
  g->ntiles = sv.ntiles;
  g->n_cstate = sv.n_cstate;
  g->cover_waves = sv.cover_waves;
  g->grade_segs = sv.grade_segs;
  g->one_lane = sv.one_lane;
  if (rc) return rc;
  // first / per-gate counts / total now belong to this grade; the grader's
  // own planes do not hold its chunks
  g->stage = 0;
  g->stage_chunk = -1;
  return PDF_OK;
}

int pdf_grade_segs(pdf_grader *g, int robust, int segs, uint64_t *total, pdf_kernel_times *times) {
  if (!g || !total) { set_error("null argument"); return PDF_ERR_ARG; }
  if (segs < 1 || segs > 8) { set_error("pdf_grade_segs: segment length must be 1..8"); return PDF_ERR_ARG; }
  if (!g->chunk_pairs) { set_error("pdf_grade_segs: load vectors first"); return PDF_ERR_STATE; }
  HIPCHK(hipSetDevice(g->device));
  int rc;
  if (segs >= 2 && (rc = ensure_segments(g, (uint32_t)segs))) return rc;
  g->grade_segs = (uint32_t)segs;
  rc = times ? grade_timed_impl(g, robust, total, times) : grade_impl(g, robust, total);
  g->grade_segs = 1;
  if (rc && rc != PDF_ERR_OVERFLOW && segs >= 2) drop_new_plane(g);
  if (!rc) g->merged_segs = (uint32_t)segs;
  return rc;
}

int pdf_get_sim_planes(const pdf_grader *gc, uint64_t *a_plane, uint64_t *b_plane) {
  pdf_grader *g = const_cast<pdf_grader *>(gc);
  if (!g) { set_error("null grader"); return PDF_ERR_ARG; }
  if (g->stage < 1) { set_error("nothing simulated yet"); return PDF_ERR_STATE; }
  HIPCHK(hipSetDevice(g->device));
  const uint64_t width = std::min(g->chunk_pairs, g->num_pairs - (uint64_t)g->stage_chunk * g->chunk_pairs);
  const uint64_t words = (width + 63) / 64, n = g->nl.n;
  const uint64_t last = width % 64 ? (1ULL << (width % 64)) - 1 : ~0ULL;
  if (g->walk_planes && !g->walk_full) {
    set_error("pdf_get_sim_planes: the level walk kept only the records it reads back; set a debug flag "
              "(pdf_grader_set_debug) before grading to keep every node's values");
    return PDF_ERR_STATE;
  }
  if (g->walk_planes) {
    // the level walk's tile-major records: [tiles][n][NV], TW + 1 frame
    // words (frame 1 of the tile's words and the halo) or 2 TW (both
    // frames), then the tile's mark-present bits, padded to an even count
    const uint64_t tw = g->walk_planes, nv = walk_record_words(g->seq, (uint32_t)tw), tiles = (words + tw - 1) / tw;
    std::vector<uint64_t> vt(tiles * n * nv);
    HIPCHK(hipMemcpy(vt.data(), g->d_VT, vt.size() * 8, hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < n; ++i)
      for (uint64_t k = 0; k < words; ++k) {
        const uint64_t *rec = vt.data() + ((k / tw) * n + i) * nv;
        const uint64_t u = k % tw, a = rec[u];
        uint64_t b = g->seq ? ((a >> 1) | (rec[u + 1] << 63)) : rec[tw + u];
        if (k + 1 == words) b = (b & last) | (a & ~last);  // lanes past the last pair read as stable
        if (a_plane) a_plane[i * words + k] = a;
        if (b_plane) b_plane[i * words + k] = b;
      }
    return PDF_OK;
  }
  std::vector<uint64_t> v(n * (words + 1)), vb;
  int rc = copy_rows(g, v.data(), g->d_V, words + 1);
  if (rc) return rc;
  if (!g->seq) {
    vb.resize(n * words);
    if ((rc = copy_rows(g, vb.data(), g->d_VB, words))) return rc;
  }
  for (uint64_t i = 0; i < n; ++i)
    for (uint64_t k = 0; k < words; ++k) {
      const uint64_t a = v[i * (words + 1) + k];
      uint64_t b = g->seq ? ((a >> 1) | (v[i * (words + 1) + k + 1] << 63)) : vb[i * words + k];
      if (k + 1 == words) b = (b & last) | (a & ~last);  // lanes past the last pair read as stable
      if (a_plane) a_plane[i * words + k] = a;
      if (b_plane) b_plane[i * words + k] = b;
    }
  return PDF_OK;
}

int pdf_get_mark_plane(const pdf_grader *gc, uint64_t *mark_plane) {
  pdf_grader *g = const_cast<pdf_grader *>(gc);
  if (!g || !mark_plane) { set_error("null argument"); return PDF_ERR_ARG; }
  if (g->stage < 2) { set_error("nothing marked yet"); return PDF_ERR_STATE; }
  HIPCHK(hipSetDevice(g->device));
  const uint64_t width = std::min(g->chunk_pairs, g->num_pairs - (uint64_t)g->stage_chunk * g->chunk_pairs);
  const uint64_t words = (width + 63) / 64, n = g->nl.n;
  int rc = copy_rows(g, mark_plane, g->d_M, words);
  if (rc) return rc;
  if (g->walk_planes) {  // the walk's mark-present bytes BT[tile][n], bit k % tw of tile k / tw
    const uint64_t tw = g->walk_planes, tiles = (words + tw - 1) / tw;
    std::vector<uint8_t> bt(n * tiles);
    HIPCHK(hipMemcpy(bt.data(), walk_bits(g, (uint32_t)tw), bt.size(), hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < n; ++i)
      for (uint64_t k = 0; k < words; ++k)
        if (!((bt[(k / tw) * n + i] >> (k % tw)) & 1)) mark_plane[i * words + k] = 0;
    return PDF_OK;
  }
  std::vector<uint64_t> nz(n * g->ntiles);
  HIPCHK(hipMemcpy(nz.data(), g->d_NZ, nz.size() * 8, hipMemcpyDeviceToHost));
  for (uint64_t i = 0; i < n; ++i)
    for (uint64_t k = 0; k < words; ++k)
      if (!((nz[i * g->ntiles + k / 64] >> (k % 64)) & 1)) mark_plane[i * words + k] = 0;
  return PDF_OK;
}

int pdf_grader_set_debug(pdf_grader *g, uint32_t flags) {
  if (!g) { set_error("null grader"); return PDF_ERR_ARG; }
  if (flags & ~(uint32_t)(PDF_DEBUG_FORWARD_MARKS | PDF_DEBUG_PI_COVER)) {
    set_error("pdf_grader_set_debug: unknown flag");
    return PDF_ERR_ARG;
  }
  if (flags != g->debug) {
    // captured grades hold the walk's keep-every-record flag as a launch
    // argument: capture again
    HIPCHK(hipSetDevice(g->device));
    if (g->stream) HIPCHK(hipStreamSynchronize(g->stream));
    for (auto &l : g->lanes)
      if (l.stream) HIPCHK(hipStreamSynchronize(l.stream));
    drop_graph(g);
  }
  g->debug = flags;
  return PDF_OK;
}

int pdf_get_forward_mark_plane(const pdf_grader *gc, uint64_t *mark_plane) {
  pdf_grader *g = const_cast<pdf_grader *>(gc);
  if (!g || !mark_plane) { set_error("null argument"); return PDF_ERR_ARG; }
  if (g->dbg_fwd_chunk < 0 || g->dbg_fwd_chunk != g->stage_chunk) {
    set_error("pdf_get_forward_mark_plane: mark this chunk with PDF_DEBUG_FORWARD_MARKS set first");
    return PDF_ERR_STATE;
  }
  HIPCHK(hipSetDevice(g->device));
  const uint64_t width = std::min(g->chunk_pairs, g->num_pairs - (uint64_t)g->stage_chunk * g->chunk_pairs);
  const uint64_t words = (width + 63) / 64, n = g->nl.n;
  std::vector<uint64_t> nz(n * g->ntiles);
  int rc = copy_rows(g, mark_plane, g->d_dbg_M, words);
  if (rc) return rc;
  HIPCHK(hipMemcpy(nz.data(), g->d_dbg_NZ, nz.size() * 8, hipMemcpyDeviceToHost));
  for (uint64_t i = 0; i < n; ++i)
    for (uint64_t k = 0; k < words; ++k)
      if (!((nz[i * g->ntiles + k / 64] >> (k % 64)) & 1)) mark_plane[i * words + k] = 0;
  return PDF_OK;
}

int pdf_get_pi_cover(const pdf_grader *g, uint64_t *pi_cover) {
  if (!g || !pi_cover) { set_error("null argument"); return PDF_ERR_ARG; }
  if (g->dbg_pic_chunk < 0 || g->dbg_pic_chunk != g->stage_chunk || g->stage != 4) {
    set_error("pdf_get_pi_cover: cover this chunk with PDF_DEBUG_PI_COVER set first");
    return PDF_ERR_STATE;
  }
  HIPCHK(hipSetDevice(g->device));
  const uint64_t width = std::min(g->chunk_pairs, g->num_pairs - (uint64_t)g->stage_chunk * g->chunk_pairs);
  HIPCHK(hipMemcpy2D(pi_cover, width * 8, g->d_dbg_pic, g->chunk_pairs * 8, width * 8, g->nl.num_pi,
                     hipMemcpyDeviceToHost));
  return PDF_OK;
}

int pdf_get_first(const pdf_grader *g, uint32_t *first_t0, uint32_t *first_t1) {
  if (!g) { set_error("null grader"); return PDF_ERR_ARG; }
  HIPCHK(hipSetDevice(g->device));
  if (first_t0) HIPCHK(hipMemcpy(first_t0, g->d_first0, sizeof(uint32_t) * g->nl.n, hipMemcpyDeviceToHost));
  if (first_t1) HIPCHK(hipMemcpy(first_t1, g->d_first1, sizeof(uint32_t) * g->nl.n, hipMemcpyDeviceToHost));
  return PDF_OK;
}

int pdf_get_segment_first(const pdf_grader *g, uint32_t *first_t0, uint32_t *first_t1, uint64_t capacity,
                          uint64_t *count) {
  if (!g || !count) { set_error("null argument"); return PDF_ERR_ARG; }
  HIPCHK(hipSetDevice(g->device));
  *count = g->seg_count;
  const uint64_t k = std::min(capacity, g->seg_count);
  if (!k) return PDF_OK;
  if (first_t0) HIPCHK(hipMemcpy(first_t0, g->d_sfirst0, sizeof(uint32_t) * k, hipMemcpyDeviceToHost));
  if (first_t1) HIPCHK(hipMemcpy(first_t1, g->d_sfirst1, sizeof(uint32_t) * k, hipMemcpyDeviceToHost));
  return PDF_OK;
}

int pdf_get_gate_counts(const pdf_grader *g, uint64_t *counts) {
  if (!g || !counts) { set_error("null argument"); return PDF_ERR_ARG; }
  HIPCHK(hipSetDevice(g->device));
  HIPCHK(hipMemcpy(counts, g->d_gate, sizeof(uint64_t) * g->nl.n, hipMemcpyDeviceToHost));
  return PDF_OK;
}

int pdf_get_total(const pdf_grader *g, uint64_t *total) {
  if (!g || !total) { set_error("null argument"); return PDF_ERR_ARG; }
  HIPCHK(hipSetDevice(g->device));
  HIPCHK(hipMemcpy(total, g->d_grand, sizeof(uint64_t), hipMemcpyDeviceToHost));
  return PDF_OK;
}

// ------------------------------------------------------- learned proposal

extern "C++" {  // helpers of the feature entry points (C++ linkage)

// The feature pass: every chunk's levels 0..last simulated and forward-marked,
// then the level and PI features, a group of `lanes` chunks at a time with
// their level launches interleaved across the lanes' streams.  After each
// group is queued, on_group(first_pair, end_pair, lanes, lanes_used) runs
// (the fused scoring pass chains the scorer there).  Queues only; the lanes'
// streams are left running.
template <typename OnGroup>
static int queue_features(pdf_grader *g, uint32_t last, int robust, uint16_t *feat_dev, uint64_t ld,
                          std::vector<pdf_grader::Lane> &ln, OnGroup &&on_group) {
  const uint64_t nchunks = (g->num_pairs + g->chunk_pairs - 1) / g->chunk_pairs;
  const uint32_t L = (uint32_t)ln.size();
  int rc = PDF_OK;
  if (last < 127) {  // level features past the simulated depth are zero
    for (uint64_t f0 : {(uint64_t)last + 1, (uint64_t)128 + last + 1})
      HIPCHK(hipMemset2DAsync(feat_dev + f0 * ld, ld * 2, 0, g->num_pairs * 2, 127 - last, g->stream));
  }
  if (L > 1) {
    HIPCHK(hipEventRecord(g->lane_start, ln[0].stream));
    for (uint32_t i = 1; i < L; ++i) HIPCHK(hipStreamWaitEvent(ln[i].stream, g->lane_start, 0));
  }
  for (uint64_t c0 = 0; c0 < nchunks && !rc; c0 += L) {
    const uint64_t c1 = std::min<uint64_t>(nchunks, c0 + L);
    uint64_t start = 0, width = 0;
    auto chunk = [&](uint64_t c) {
      lane_load(g, ln[c % L]);
      start = c * g->chunk_pairs;
      width = std::min(g->chunk_pairs, g->num_pairs - start);
    };
    for (uint32_t k = 0; k <= last && !rc; ++k)
      for (uint64_t c = c0; c < c1 && !rc; ++c) {
        chunk(c);
        rc = launch_fwd<true, true>(g, start, width, robust, false, last, k, k);
      }
    for (uint64_t c = c0; c < c1 && !rc; ++c) {
      chunk(c);
      const Planes pl = planes(g, width);
      const dim3 lgrid(grid_for(pl.words, kFeatWords), 2 * (last + 1)), pgrid(grid_for(pl.words, 256), 256);
      if (g->seq) {
        k_features<true><<<lgrid, 256, 0, g->stream>>>(pl, g->d_level_off, last + 1, start, g->num_pairs, feat_dev,
                                                       ld);
        k_features_pi<true><<<pgrid, 256, 0, g->stream>>>(g->d_rec, pl, g->nl.num_pi, start, g->num_pairs, feat_dev,
                                                         ld);
      } else {
        k_features<false><<<lgrid, 256, 0, g->stream>>>(pl, g->d_level_off, last + 1, start, g->num_pairs, feat_dev,
                                                        ld);
        k_features_pi<false><<<pgrid, 256, 0, g->stream>>>(g->d_rec, pl, g->nl.num_pi, start, g->num_pairs,
                                                          feat_dev, ld);
      }
      if (!hip_ok(hipGetLastError(), "k_features")) rc = PDF_ERR_HIP;
    }
    if (!rc)
      rc = on_group(c0 * g->chunk_pairs, std::min(c1 * g->chunk_pairs, g->num_pairs), (uint32_t)(c1 - c0));
  }
  lane_load(g, ln[0]);
  return rc;
}

static int check_features_args(pdf_grader *g, uint16_t *feat_dev, uint64_t ld, const char *who) {
  if (!g || !feat_dev) { set_error("null argument"); return PDF_ERR_ARG; }
  if (!g->chunk_pairs) { set_error(std::string(who) + ": load pairs first"); return PDF_ERR_STATE; }
  if (ld % 64 || ld < g->num_pairs || (uintptr_t)feat_dev % 16) {
    set_error(std::string(who) +
              ": ld must be a multiple of 64 and >= the pair count, feat_dev 16-byte aligned");
    return PDF_ERR_ARG;
  }
  return PDF_OK;
}

static std::vector<pdf_grader::Lane> all_lanes(pdf_grader *g) {
  std::vector<pdf_grader::Lane> ln(1 + g->lanes.size());
  lane_save(g, ln[0]);
  for (size_t i = 1; i < ln.size(); ++i) ln[i] = g->lanes[i - 1];
  return ln;
}

}  // extern "C++"

int pdf_pair_features(pdf_grader *g, uint32_t max_level, int robust, uint16_t *feat_dev, uint64_t ld) {
  int rc = check_features_args(g, feat_dev, ld, "pdf_pair_features");
  if (rc) return rc;
  HIPCHK(hipSetDevice(g->device));
  const uint32_t last = std::min(std::min(max_level, g->nl.L), 127u);
  std::vector<pdf_grader::Lane> ln = all_lanes(g);
  rc = queue_features(g, last, robust, feat_dev, ld, ln, [](uint64_t, uint64_t, uint32_t) { return PDF_OK; });
  for (size_t i = 1; i < ln.size(); ++i) HIPCHK(hipStreamSynchronize(ln[i].stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  if (rc) return rc;
  g->stage = 0;  // the planes hold a partial simulation
  g->stage_chunk = -1;
  return PDF_OK;
}

int pdf_score_pairs(pdf_grader *g, pdf_mlp *m, uint32_t max_level, int robust, uint16_t *feat_dev, uint64_t ld,
                    float *score_dev) {
  int rc = check_features_args(g, feat_dev, ld, "pdf_score_pairs");
  if (rc) return rc;
  if (!m || !score_dev) { set_error("null argument"); return PDF_ERR_ARG; }
  if (pdf::mlp_device(m) != g->device) { set_error("pdf_score_pairs: grader and scorer on different devices"); return PDF_ERR_ARG; }
  HIPCHK(hipSetDevice(g->device));
  const uint32_t last = std::min(std::min(max_level, g->nl.L), 127u);
  // 64-feature stages that hold only levels past `last`: stage 1 (transition
  // features of levels 64-127) and stage 3 (forward-mark features of levels
  // 64-127) when last < 64; zero in every row, so the scorer skips them
  const uint32_t zero_stages = last < 64 ? (1u << 1) | (1u << 3) : 0u;
  std::vector<pdf_grader::Lane> ln = all_lanes(g);
  if ((rc = pdf::mlp_prepare(m, feat_dev, ld, std::min<uint64_t>(g->num_pairs, g->chunk_pairs * ln.size())))) return rc;
  // one event per lane, re-recorded for every group (a wait takes the
  // record current when it is queued)
  std::vector<hipEvent_t> ev(ln.size(), nullptr);
  for (auto &e : ev)
    if (!hip_ok(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate")) rc = PDF_ERR_HIP;
  const hipStream_t ms = pdf::mlp_stream(m);
  if (!rc)
    rc = queue_features(g, last, robust, feat_dev, ld, ln, [&](uint64_t p0, uint64_t p1, uint32_t used) -> int {
      // the scorer takes the group once every lane that worked on it is done
      for (uint32_t i = 0; i < used; ++i) {
        if (!hip_ok(hipEventRecord(ev[i], ln[i].stream), "hipEventRecord") ||
            !hip_ok(hipStreamWaitEvent(ms, ev[i], 0), "hipStreamWaitEvent"))
          return PDF_ERR_HIP;
      }
      return pdf::mlp_queue_forward(m, feat_dev, ld, p0, p1, score_dev, zero_stages);
    });
  for (size_t i = 1; i < ln.size(); ++i) (void)hipStreamSynchronize(ln[i].stream);
  (void)hipStreamSynchronize(g->stream);
  const hipError_t se = hipStreamSynchronize(ms);
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  if (!rc && !hip_ok(se, "hipStreamSynchronize")) rc = PDF_ERR_HIP;
  if (rc) return rc;
  g->stage = 0;
  g->stage_chunk = -1;
  return PDF_OK;
}

int pdf_pair_gains(pdf_grader *g, int robust, uint32_t *gain_dev) {
  if (!g || !gain_dev) { set_error("null argument"); return PDF_ERR_ARG; }
  if (!g->chunk_pairs) { set_error("pdf_pair_gains: load pairs first"); return PDF_ERR_STATE; }
  HIPCHK(hipSetDevice(g->device));
  int rc = queue_reset(g);
  if (rc) return rc;
  const uint64_t nchunks = (g->num_pairs + g->chunk_pairs - 1) / g->chunk_pairs;
  for (uint64_t c = 0; c < nchunks; ++c) {
    const uint64_t start = c * g->chunk_pairs;
    const uint64_t width = std::min(g->chunk_pairs, g->num_pairs - start);
    if ((rc = launch_fwd<true, true>(g, start, width, robust, true))) return rc;
    if (!g->seq && (rc = reset_chunk(g))) return rc;
    if ((rc = launch_reverse<true, true>(g, start, width))) return rc;
    const Planes pl = planes(g, width);
    k_cover<true><<<g->cover_waves, 64, 0, g->stream>>>(g->net(), pl, (uint32_t)start, g->d_list, g->d_list_cnt,
                                                      g->nl.n, g->d_act_list, g->d_act_count, g->d_first0,
                                                      g->d_first1, g->d_cs, g->d_hs, g->d_gate, g->d_chunk_total,
                                                      g->d_grand, g->d_err, (uint32_t)width, gain_dev, nullptr,
                                                      nullptr, 0, g->cover_caps);
    HIPCHK(hipGetLastError());
  }
  uint32_t err = 0;
  HIPCHK(hipMemcpyAsync(&err, g->d_err, sizeof(err), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  g->stage = 0;
  g->stage_chunk = -1;
  if (err & kErrOverflow) { set_error("pdf_pair_gains: uint64 path count overflow"); return PDF_ERR_OVERFLOW; }
  return PDF_OK;
}

int pdf_grader_rank_pairs(pdf_grader *g, const float *score_dev, uint32_t *perm_dev) {
  if (!g || !score_dev) { set_error("null argument"); return PDF_ERR_ARG; }
  if (!g->chunk_pairs) { set_error("pdf_grader_rank_pairs: load pairs first"); return PDF_ERR_STATE; }
  if (g->seq) {
    set_error("pdf_grader_rank_pairs: needs explicit pairs (reordering consecutive vectors changes the pairs)");
    return PDF_ERR_STATE;
  }
  HIPCHK(hipSetDevice(g->device));
  const uint64_t num = g->num_pairs;
  const size_t vsize = (size_t)g->nl.num_pi * g->vstride;
  const uint32_t nb = (g->nl.num_pi + 63) / 64;
  const uint64_t pm_words = g->vstride * 64 * nb;  // pair-major copy of one vector set
  size_t temp_bytes = 0;
  HIPCHK(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, temp_bytes, score_dev, g->d_rk_keys, g->d_rk_idx,
                                                      g->d_rk_perm, (int)num, 0, 32, g->stream));
  if (g->rk_pairs < num || g->rk_vsize != vsize || g->rk_temp_bytes < temp_bytes) {
    HIPCHK(hipStreamSynchronize(g->stream));
    free_rank(g);
    int rc = dalloc(&g->d_rk_keys, num);
    if (!rc) rc = dalloc(&g->d_rk_idx, num);
    if (!rc) rc = dalloc(&g->d_rk_perm, num);
    if (!rc) rc = dalloc(&g->d_rk_va, vsize);
    if (!rc) rc = dalloc(&g->d_rk_vb, vsize);
    if (!rc) rc = dalloc(&g->d_rk_pm0, pm_words);
    if (!rc) rc = dalloc(&g->d_rk_pm1, pm_words);
    if (!rc && hipMalloc(&g->d_rk_temp, temp_bytes) != hipSuccess) {
      (void)hipGetLastError();
      g->d_rk_temp = nullptr;
      set_error("pdf_grader_rank_pairs: hipMalloc of sort scratch failed");
      rc = PDF_ERR_NOMEM;
    }
    if (rc) {
      free_rank(g);
      return rc;
    }
    g->rk_pairs = num;
    g->rk_vsize = vsize;
    g->rk_temp_bytes = temp_bytes;
  }
  k_iota<<<grid_for(num, 256), 256, 0, g->stream>>>(g->d_rk_idx, num);
  // stable descending radix sort: equal scores keep their pair order
  temp_bytes = g->rk_temp_bytes;
  HIPCHK(hipcub::DeviceRadixSort::SortPairsDescending(g->d_rk_temp, temp_bytes, score_dev, g->d_rk_keys,
                                                      g->d_rk_idx, g->d_rk_perm, (int)num, 0, 32, g->stream));
  // each vector set: to pair-major, gather, back to PI-major
  const uint64_t tiles = g->vstride * nb;  // one wave each
  for (int set = 0; set < 2; ++set) {
    const uint64_t *src = set ? g->d_vb : g->d_va;
    uint64_t *dst = set ? g->d_rk_vb : g->d_rk_va;
    k_pi_to_pair<<<grid_for(tiles * 64, 256), 256, 0, g->stream>>>(src, g->vstride, g->nl.num_pi, g->vstride, nb,
                                                                   g->d_rk_pm0);
    k_gather_pairs<<<grid_for(pm_words, 256), 256, 0, g->stream>>>(g->d_rk_pm0, g->d_rk_pm1, g->d_rk_perm, num, nb,
                                                                   pm_words);
    k_pair_to_pi<<<grid_for(tiles * 64, 256), 256, 0, g->stream>>>(g->d_rk_pm1, nb, g->nl.num_pi, g->vstride,
                                                                   dst, g->vstride);
  }
  HIPCHK(hipGetLastError());
  if (perm_dev) HIPCHK(hipMemcpyAsync(perm_dev, g->d_rk_perm, num * 4, hipMemcpyDeviceToDevice, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  // the permuted vectors become the loaded ones; the old buffers are the
  // next call's scratch
  std::swap(g->d_va, g->d_rk_va);
  std::swap(g->d_vb, g->d_rk_vb);
  drop_graph(g);
  g->stage = 0;
  g->stage_chunk = -1;
  return PDF_OK;
}

int pdf_grade_curve(pdf_grader *g, int robust, uint64_t *cum_total, float *cum_ms, uint32_t *pair_total_dev) {
  if (!g || !cum_total) { set_error("null argument"); return PDF_ERR_ARG; }
  if (!g->chunk_pairs) { set_error("pdf_grade_curve: load vectors first"); return PDF_ERR_STATE; }
  HIPCHK(hipSetDevice(g->device));
  if (pair_total_dev) HIPCHK(hipMemsetAsync(pair_total_dev, 0, g->num_pairs * 4, g->stream));
  const uint64_t nchunks = (g->num_pairs + g->chunk_pairs - 1) / g->chunk_pairs;
  std::vector<hipEvent_t> ev(nchunks + 1, nullptr);
  uint64_t *curve = nullptr;
  int rc = dalloc(&curve, nchunks);
  for (auto &e : ev)
    if (!rc && hipEventCreate(&e) != hipSuccess) { set_error("hipEventCreate failed"); rc = PDF_ERR_HIP; }
  if (!rc) {
    (void)hipEventRecord(ev[0], g->stream);
    rc = queue_reset(g);
  }
  for (uint64_t c = 0; c < nchunks && !rc; ++c) {
    const uint64_t start = c * g->chunk_pairs;
    const uint64_t width = std::min(g->chunk_pairs, g->num_pairs - start);
    if ((rc = launch_fwd<true, true>(g, start, width, robust, true))) break;
    if (!g->seq && (rc = reset_chunk(g))) break;
    if ((rc = launch_reverse<true, true>(g, start, width))) break;
    if ((rc = launch_cover(g, start, width, pair_total_dev))) break;
    k_copy_u64<<<1, 1, 0, g->stream>>>(g->d_grand, curve + c);
    (void)hipEventRecord(ev[c + 1], g->stream);
  }
  uint32_t err = 0;
  if (!rc && !hip_ok(hipStreamSynchronize(g->stream), "pdf_grade_curve")) rc = PDF_ERR_HIP;
  if (!rc && !hip_ok(hipMemcpy(cum_total, curve, nchunks * 8, hipMemcpyDeviceToHost), "hipMemcpy")) rc = PDF_ERR_HIP;
  if (!rc && !hip_ok(hipMemcpy(&err, g->d_err, 4, hipMemcpyDeviceToHost), "hipMemcpy")) rc = PDF_ERR_HIP;
  for (uint64_t c = 0; c < nchunks && !rc && cum_ms; ++c)
    if (!hip_ok(hipEventElapsedTime(&cum_ms[c], ev[0], ev[c + 1]), "hipEventElapsedTime")) rc = PDF_ERR_HIP;
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  dfree(curve);
  if (rc) return rc;
  g->stage = 4;
  g->stage_chunk = (int64_t)nchunks - 1;
  if (err & kErrOverflow) { set_error("pdf_grade_curve: uint64 path count overflow"); return PDF_ERR_OVERFLOW; }
  return PDF_OK;
}

}  // extern "C"

// ------------------------------------------------------------- autotuning

extern "C" int pdf_grader_autotune(pdf_grader *g, int robust, uint32_t reps, pdf_grader_tuning *best) {
  if (!g || reps == 0) { set_error("pdf_grader_autotune: bad argument"); return PDF_ERR_ARG; }
  if (!g->num_pairs) { set_error("pdf_grader_autotune: load vectors first"); return PDF_ERR_STATE; }
  HIPCHK(hipSetDevice(g->device));
  const uint64_t P = g->num_pairs;
  auto round64 = [](uint64_t x) { return (x + 63) / 64 * 64; };
  // the caller's configuration, restored if tuning fails
  const uint64_t chunk0 = g->chunk_pairs;
  const uint32_t lanes0 = g->tune_lanes, graph0 = g->tune_use_graph, wpl0 = g->tune_words_per_lane;
  const uint32_t walk0 = g->tune_walk;
  auto restore = [&](int rc) {
    g->tune_lanes = lanes0;
    g->tune_use_graph = graph0;
    g->tune_words_per_lane = wpl0;
    g->tune_walk = walk0;
    (void)pdf_grader_set_chunk(g, chunk0);
    return rc;
  };
  // Candidates: 1..4 lanes, each with the even split of the pairs and the
  // power-of-two chunks from 32,768 pairs below it (from 65,536 on, 2 words
  // per lane applies), kept when the lanes' chunk buffers fit in 85% of the
  // HBM that is free once this grader's own chunk buffers are released.
  HIPCHK(hipStreamSynchronize(g->stream));
  free_chunk(g);
  size_t free_b = 0, total_b = 0;
  if (!hip_ok(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo")) return restore(PDF_ERR_HIP);
  const double budget = 0.85 * (double)free_b;
  const double n = g->nl.n;
  auto lane_bytes = [&](uint64_t chunk, uint32_t walk) {  // what alloc_chunk_buffers takes for one lane
    const uint64_t words = (chunk + 63) / 64, pitch = (words + 1 + 15) / 16 * 16, ntiles = (words + 63) / 64;
    const double per_wave = 16.0 * n;  // the cover's c/h fallback scratch
    const double waves = std::max(1.0, std::min((double)g->max_cover_waves, std::floor((double)(1ull << 30) / per_wave)));
    double bytes = n * (8.0 * pitch * (g->seq ? 3 : 4) + 8.0 * ntiles + 4.0 * words) + 4.0 * chunk + 12.0 * words +
                   waves * per_wave;
    if (walk) bytes += 8.0 * (double)walk_plane_words(g->seq, walk, g->nl.n, words);
    if (g->tune_rev_compact)  // the compacted reverse pass's item lists
      bytes += (double)sizeof(RevTile) * (n - g->nl.level_off[1]) * rev_slots_per_gate(ntiles) +
               4.0 * kRevSubLists * kRevCntStride * (g->nl.L + 1);
    return bytes;
  };
  struct Cand { uint32_t lanes; uint64_t chunk; uint32_t walk; };
  std::vector<Cand> cands;
  auto add = [&](uint32_t lanes, uint64_t chunk) {
    if ((P + chunk - 1) / chunk < lanes) return;  // fewer chunks than lanes
    // per-level launches, and the level walk with tiles of 2 and 4 words
    // (1 word is never the faster walk, 8 leaves most CUs idle below 2^19 pairs)
    for (uint32_t walk : {0u, 2u, 4u})
      if (lanes * lane_bytes(chunk, walk) <= budget) cands.push_back({lanes, chunk, walk});
  };
  for (uint32_t lanes = 1; lanes <= 4; ++lanes) {
    const uint64_t even = lanes == 1 ? P : round64((P + lanes - 1) / lanes);
    if (lanes > 1 && even >= P) continue;
    add(lanes, even);
    for (uint64_t c = 32768; c < even; c *= 2) add(lanes, c);
  }
  if (cands.empty()) {
    set_error("pdf_grader_autotune: no chunk configuration fits in free HBM");
    return restore(PDF_ERR_STATE);
  }
  uint64_t ref = 0;
  bool have_ref = false;
  pdf_grader_tuning bt{};
  bt.ms = 1e30f;
  int rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!hip_ok(hipEventCreate(&e0), "hipEventCreate")) return restore(PDF_ERR_HIP);
  if (!hip_ok(hipEventCreate(&e1), "hipEventCreate")) {
    (void)hipEventDestroy(e0);
    return restore(PDF_ERR_HIP);
  }
  auto run = [&]() -> int {
    for (const Cand &c : cands)
      for (uint32_t graph = 0; graph < 2; ++graph)
        for (uint32_t wpl = 1; wpl <= 2; ++wpl) {
          if (wpl == 2 && words_per_lane(2, g->seq, true, true, c.chunk) != 2) continue;  // no effect
          if (wpl == 2 && c.walk) continue;  // the walk's tiles take the place of words per lane
          g->tune_lanes = c.lanes;
          g->tune_use_graph = graph;
          g->tune_words_per_lane = wpl;
          g->tune_walk = c.walk;
          rc = pdf_grader_set_chunk(g, c.chunk);
          if (rc == PDF_ERR_NOMEM) continue;  // the estimate was short: skip this candidate
          if (rc) return rc;
          std::vector<float> t(reps + 1);
          uint64_t tot = 0;
          for (uint32_t r = 0; r <= reps; ++r) {  // the first run warms up (and captures)
            HIPCHK(hipEventRecord(e0, g->stream));
            if ((rc = grade_impl(g, robust, &tot))) return rc;
            HIPCHK(hipEventRecord(e1, g->stream));
            HIPCHK(hipEventSynchronize(e1));
            HIPCHK(hipEventElapsedTime(&t[r], e0, e1));
          }
          if (!have_ref) {
            ref = tot;
            have_ref = true;
          } else if (tot != ref) {
            set_error("pdf_grader_autotune: totals differ between configurations");
            return PDF_ERR_STATE;
          }
          std::sort(t.begin() + 1, t.end());
          const float med = t[1 + reps / 2];
          if (med < bt.ms) bt = pdf_grader_tuning{c.lanes, c.chunk, graph, wpl, c.walk, med};
        }
    if (!have_ref) {
      set_error("pdf_grader_autotune: no candidate configuration could be allocated");
      return PDF_ERR_NOMEM;
    }
    return PDF_OK;
  };
  rc = run();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return restore(rc);
  g->tune_lanes = bt.lanes;
  g->tune_use_graph = bt.use_graph;
  g->tune_words_per_lane = bt.words_per_lane;
  g->tune_walk = bt.walk;
  if ((rc = pdf_grader_set_chunk(g, bt.chunk_pairs))) return restore(rc);
  if (best) *best = bt;
  return PDF_OK;
}

// ----------------------------------------------------------- batched grading
// A batch owns one grader per circuit (vectors, planes, first[] and lists),
// the per-circuit descriptors and the per-level segment tables in HBM, a
// shared cover scratch and one captured graph of the whole grade.

struct pdf_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<pdf_grader *> members;
  uint32_t C = 0, maxL = 0, total_nodes = 0, max_n = 0;
  bool wide = true;
  BatchCircuit *d_circ = nullptr;
  uint32_t *d_node_base = nullptr;
  BatchSeg *d_segs = nullptr;
  std::vector<uint32_t> fwd_off, fwd_cnt, fwd_grid, rev_off, rev_cnt, rev_grid;  // per level
  uint64_t *d_cs = nullptr, *d_hs = nullptr;
  uint32_t cover_waves = 0;
  uint64_t *d_totals = nullptr;
  uint32_t *d_errs = nullptr;
  hipGraphExec_t graph = nullptr;
  int graph_robust = -1;
  // compacted reverse pass: scan segments per level, the work list and its
  // per-level counters
  BatchScanSeg *d_scan = nullptr;
  std::vector<uint32_t> scan_off, scan_cnt, scan_grid, work_grid;  // per level
  RevItem *d_work = nullptr;
  uint32_t *d_work_cnt = nullptr;
  // the lists the forward launches build (rev_compact 2): union level k's
  // 64 sub-lists of blist_cap[k] items at d_blist + blist_off[k], their
  // counters at d_blist_cnt + k * kRevSubLists * kRevCntStride
  RevTile *d_blist = nullptr;
  uint32_t *d_blist_cnt = nullptr;
  std::vector<uint64_t> blist_off;
  std::vector<uint32_t> blist_cap;
  // tuning knobs (pdf_batch_set_tuning / pdf_batch_autotune)
  uint32_t tune_cover_waves = 512, tune_order = 1, tune_use_graph = 1, tune_rev_compact = 1, tune_wpl = 2;
  // (tune_rev_compact: 2 from pdf_batch_create where every chunk has whole-wave tiles)
  // words per lane of the forward launches: 2 only with whole-wave tiles
  uint32_t fwd_nw() const { return wide && tune_wpl == 2 ? 2 : 1; }
};

namespace {

void batch_drop_graph(pdf_batch *b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
  b->graph_robust = -1;
}

// Segment tables of every level.  order 0: circuits in index order; order 1:
// the segments with most blocks first (the longest-running blocks are
// dispatched first, so the launch's tail is short).
int batch_build_segments(pdf_batch *b) {
  std::vector<BatchSeg> all;
  b->fwd_off.assign(b->maxL + 1, 0);
  b->fwd_cnt.assign(b->maxL + 1, 0);
  b->fwd_grid.assign(b->maxL + 1, 0);
  b->rev_off = b->fwd_off;
  b->rev_cnt = b->fwd_cnt;
  b->rev_grid = b->fwd_grid;
  for (int pass = 0; pass < 2; ++pass) {
    for (uint32_t k = 0; k <= b->maxL; ++k) {
      std::vector<BatchSeg> lv;
      for (uint32_t c = 0; c < b->C; ++c) {
        const pdf_grader *g = b->members[c];
        if (k > g->nl.L) continue;
        const uint32_t lo = g->nl.level_off[k], count = g->nl.level_off[k + 1] - lo;
        const uint32_t words = (uint32_t)((g->num_pairs + 63) / 64);
        BatchSeg sg{};
        sg.mp = level_map(words, lo, count, pass == 0, &sg.nblocks);
        if (pass == 0 && b->fwd_nw() == 2) {  // 128-word tiles, one node per wave
          sg.mp.T = (words + 127) / 128;
          sg.mp.main_blocks = sg.mp.T * ((count + (kBlock >> 6) - 1) / (kBlock >> 6));
          sg.nblocks = sg.mp.main_blocks + grid_for(count, kBlock);
        }
        sg.circuit = c;
        sg.clear = pass == 0 && k == g->nl.L;
        lv.push_back(sg);
      }
      if (b->tune_order == 1)
        std::stable_sort(lv.begin(), lv.end(), [](const BatchSeg &x, const BatchSeg &y) { return x.nblocks > y.nblocks; });
      uint32_t blk = 0;
      for (auto &sg : lv) {
        sg.block0 = blk;
        blk += sg.nblocks;
      }
      (pass == 0 ? b->fwd_off : b->rev_off)[k] = (uint32_t)all.size();
      (pass == 0 ? b->fwd_cnt : b->rev_cnt)[k] = (uint32_t)lv.size();
      (pass == 0 ? b->fwd_grid : b->rev_grid)[k] = blk;
      all.insert(all.end(), lv.begin(), lv.end());
    }
  }
  // lists built by the forward launches: a block appends at most one item
  // per wave and word of a lane (one gate a wave), to sub-list block % 64
  dfree(b->d_blist);
  dfree(b->d_blist_cnt);
  b->blist_off.assign(b->maxL + 1, 0);
  b->blist_cap.assign(b->maxL + 1, 0);
  int rc = PDF_OK, list_rc = PDF_OK;
  if (b->tune_rev_compact == 2) {
    const uint64_t per_block = (kBlock / 64) * b->fwd_nw();
    uint64_t at = 0;
    for (uint32_t k = 0; k <= b->maxL && !list_rc; ++k) {
      const uint64_t cap = per_block * ((b->fwd_grid[k] + kRevSubLists - 1) / kRevSubLists);
      if (cap > 0xFFFFFFFFull) {
        set_error("pdf_batch: a union level has too many blocks for the forward-built reverse lists");
        list_rc = PDF_ERR_UNSUPPORTED;
      }
      b->blist_off[k] = at;
      b->blist_cap[k] = (uint32_t)std::min<uint64_t>(cap, 0xFFFFFFFFull);
      at += cap * kRevSubLists;
    }
    if (!list_rc) list_rc = dalloc(&b->d_blist, std::max<uint64_t>(1, at));
    if (!list_rc) list_rc = dalloc(&b->d_blist_cnt, ((uint64_t)b->maxL + 1) * kRevSubLists * kRevCntStride);
    if (list_rc) {
      // no room for the lists: the batch keeps a consistent state with the
      // scan-built lists (knob value 1), and the caller learns why
      dfree(b->d_blist);
      dfree(b->d_blist_cnt);
      b->tune_rev_compact = 1;
    }
  }
  if (b->tune_rev_compact == 2) {
    for (uint32_t k = 0; k <= b->maxL; ++k)
      for (uint32_t s = b->fwd_off[k]; s < b->fwd_off[k] + b->fwd_cnt[k]; ++s) {
        LevelMap &mp = all[s].mp;
        mp.items = b->d_blist + b->blist_off[k];
        mp.items_cnt = b->d_blist_cnt + (uint64_t)k * kRevSubLists * kRevCntStride;
        mp.items_cap = b->blist_cap[k];
        mp.items_nsub = kRevSubLists;
        mp.items_by_block = 1;
        mp.items_tag = all[s].circuit << kBatchTileBits;
      }
  }
  dfree(b->d_segs);
  rc = dalloc(&b->d_segs, all.size());
  if (!rc && !hip_ok(hipMemcpy(b->d_segs, all.data(), all.size() * sizeof(BatchSeg), hipMemcpyHostToDevice),
                     "hipMemcpy"))
    rc = PDF_ERR_HIP;
  if (rc) return rc;
  // the compacted reverse pass: per level, one scan segment per circuit with
  // that level (items = its gates x its 64-word tiles)
  std::vector<BatchScanSeg> scan;
  b->scan_off.assign(b->maxL + 1, 0);
  b->scan_cnt.assign(b->maxL + 1, 0);
  b->scan_grid.assign(b->maxL + 1, 0);
  b->work_grid.assign(b->maxL + 1, 0);
  uint64_t cap = 1;
  for (uint32_t k = 0; k <= b->maxL; ++k) {
    b->scan_off[k] = (uint32_t)scan.size();
    uint64_t blk = 0, items = 0;
    for (uint32_t c = 0; c < b->C; ++c) {
      const pdf_grader *g = b->members[c];
      if (k > g->nl.L) continue;
      const uint32_t lo = g->nl.level_off[k], count = g->nl.level_off[k + 1] - lo;
      const uint32_t tiles = (uint32_t)(((g->num_pairs + 63) / 64 + 63) / 64);
      BatchScanSeg sg{c, lo, count, tiles, (uint32_t)blk, 0};
      scan.push_back(sg);
      items += (uint64_t)count * tiles;
      blk += ((uint64_t)count * tiles + 255) / 256;
    }
    b->scan_cnt[k] = (uint32_t)scan.size() - b->scan_off[k];
    b->scan_grid[k] = (uint32_t)blk;
    // up to one item per wave, at most 2,048 blocks (8,192 waves: one round)
    b->work_grid[k] = (uint32_t)std::min<uint64_t>(2048, std::max<uint64_t>(1, (items + 3) / 4));
    cap = std::max(cap, items);
  }
  dfree(b->d_scan);
  dfree(b->d_work);
  dfree(b->d_work_cnt);
  rc = dalloc(&b->d_scan, std::max<size_t>(1, scan.size()));
  if (!rc) rc = dalloc(&b->d_work, cap);
  if (!rc) rc = dalloc(&b->d_work_cnt, b->maxL + 1);
  if (!rc && !scan.empty() &&
      !hip_ok(hipMemcpy(b->d_scan, scan.data(), scan.size() * sizeof(BatchScanSeg), hipMemcpyHostToDevice),
              "hipMemcpy"))
    rc = PDF_ERR_HIP;
  return rc ? rc : list_rc;
}

int batch_alloc_scratch(pdf_batch *b) {
  dfree(b->d_cs);
  dfree(b->d_hs);
  b->cover_waves = std::max(1u, b->tune_cover_waves);
  int rc = dalloc(&b->d_cs, (size_t)b->cover_waves * b->max_n);
  if (!rc) rc = dalloc(&b->d_hs, (size_t)b->cover_waves * b->max_n);
  return rc;
}

template <bool R>
void batch_queue_t(pdf_batch *b) {
  hipStream_t st = b->stream;
  k_reset_batch<<<grid_for(b->total_nodes, 256), 256, 0, st>>>(b->d_circ, b->d_node_base, b->C, b->total_nodes);
  const uint32_t nw = b->fwd_nw();
  if (b->tune_rev_compact == 2) {
    const uint64_t nc = ((uint64_t)b->maxL + 1) * kRevSubLists * kRevCntStride;
    k_fill_u32<<<(uint32_t)std::min<uint64_t>(grid_for(nc, 256), 1024), 256, 0, st>>>(b->d_blist_cnt, nc, 0u);
  }
  for (uint32_t k = 0; k <= b->maxL; ++k) {
    const BatchSeg *sg = b->d_segs + b->fwd_off[k];
    if (nw == 2)
      k_fwd_batch<R, true, 2><<<b->fwd_grid[k], kBlock, 0, st>>>(b->d_circ, sg, b->fwd_cnt[k]);
    else if (b->wide)
      k_fwd_batch<R, true, 1><<<b->fwd_grid[k], kBlock, 0, st>>>(b->d_circ, sg, b->fwd_cnt[k]);
    else
      k_fwd_batch<R, false, 1><<<b->fwd_grid[k], kBlock, 0, st>>>(b->d_circ, sg, b->fwd_cnt[k]);
  }
  if (b->tune_rev_compact == 2) {
    for (uint32_t k = b->maxL + 1; k-- > 0;) {
      if (!b->scan_grid[k]) continue;
      k_rev_list_batch<<<b->work_grid[k], 256, 0, st>>>(b->d_circ, b->d_blist + b->blist_off[k],
                                                        b->d_blist_cnt + (uint64_t)k * kRevSubLists * kRevCntStride,
                                                        b->blist_cap[k]);
    }
  } else if (b->tune_rev_compact) {
    k_fill_u32<<<1, 256, 0, st>>>(b->d_work_cnt, b->maxL + 1, 0u);
    for (uint32_t k = b->maxL + 1; k-- > 0;) {
      if (!b->scan_grid[k]) continue;
      k_rev_scan_batch<<<b->scan_grid[k], 256, 0, st>>>(b->d_circ, b->d_scan + b->scan_off[k], b->scan_cnt[k],
                                                       b->d_work, b->d_work_cnt + k);
      k_rev_work_batch<<<b->work_grid[k], 256, 0, st>>>(b->d_circ, b->d_work, b->d_work_cnt + k);
    }
  } else {
    for (uint32_t k = b->maxL + 1; k-- > 0;) {
      if (b->wide)
        k_rev_batch<true><<<b->rev_grid[k], kBlock, 0, st>>>(b->d_circ, b->d_segs + b->rev_off[k], b->rev_cnt[k]);
      else
        k_rev_batch<false><<<b->rev_grid[k], kBlock, 0, st>>>(b->d_circ, b->d_segs + b->rev_off[k], b->rev_cnt[k]);
    }
  }
  k_events_batch<<<grid_for(b->total_nodes, 256), 256, 0, st>>>(b->d_circ, b->d_node_base, b->C, b->total_nodes);
  k_cover_batch<<<b->cover_waves, 64, 0, st>>>(b->d_circ, b->C, b->d_cs, b->d_hs, b->max_n);
  k_gather_totals<<<grid_for(b->C, 256), 256, 0, st>>>(b->d_circ, b->C, b->d_totals, b->d_errs);
}

int batch_queue(pdf_batch *b, int robust) {
  if (robust) batch_queue_t<true>(b);
  else batch_queue_t<false>(b);
  HIPCHK(hipGetLastError());
  return PDF_OK;
}

// One grade of the whole batch on b->stream (graph replay or direct launches).
int batch_run(pdf_batch *b, int robust) {
  if (b->tune_use_graph) {
    if (!b->graph || b->graph_robust != (robust != 0)) {
      batch_drop_graph(b);
      hipGraph_t graph = nullptr;
      HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      const int rc = batch_queue(b, robust);
      const hipError_t e = hipStreamEndCapture(b->stream, &graph);
      if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      HIPCHK(e);
      const hipError_t ie = hipGraphInstantiate(&b->graph, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ie != hipSuccess) {
        b->graph = nullptr;
        HIPCHK(ie);
      }
      b->graph_robust = robust != 0;
    }
    HIPCHK(hipGraphLaunch(b->graph, b->stream));
    return PDF_OK;
  }
  return batch_queue(b, robust);
}

int batch_results(pdf_batch *b, uint64_t *totals) {
  std::vector<uint32_t> errs(b->C);
  HIPCHK(hipMemcpyAsync(errs.data(), b->d_errs, b->C * 4, hipMemcpyDeviceToHost, b->stream));
  if (totals) HIPCHK(hipMemcpyAsync(totals, b->d_totals, b->C * 8, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  for (uint32_t c = 0; c < b->C; ++c)
    if (errs[c] & kErrOverflow) {
      set_error("pdf_batch_grade: uint64 path count overflow in circuit " + std::to_string(c));
      return PDF_ERR_OVERFLOW;
    }
  return PDF_OK;
}

}  // namespace

extern "C" {

int pdf_batch_create(int device, uint32_t count, const pdf_netlist *const *nets, const uint64_t *const *vec_bits,
                     const uint64_t *num_vectors, pdf_batch **out) {
  if (!out || !nets || !vec_bits || !num_vectors || count == 0) { set_error("pdf_batch_create: bad argument"); return PDF_ERR_ARG; }
  *out = nullptr;
  auto *b = new (std::nothrow) pdf_batch;
  if (!b) { set_error("out of memory"); return PDF_ERR_NOMEM; }
  b->device = device;
  int rc = PDF_OK;
  for (uint32_t c = 0; c < count && !rc; ++c) {
    pdf_grader *g = nullptr;
    if ((rc = pdf_grader_create(device, nets[c], &g))) break;
    b->members.push_back(g);
    g->max_cover_waves = 1;  // the batch's own scratch serves the cover
    if ((rc = pdf_grader_set_vectors(g, vec_bits[c], num_vectors[c]))) break;
    if (g->chunk_pairs != g->num_pairs) {
      set_error("pdf_batch_create: circuit " + std::to_string(c) + " does not fit one chunk in HBM");
      rc = PDF_ERR_NOMEM;
    }
  }
  if (!rc && (!hip_ok(hipSetDevice(device), "hipSetDevice") ||
              !hip_ok(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate") ||
              !hip_ok(hipEventCreate(&b->ev0), "hipEventCreate") || !hip_ok(hipEventCreate(&b->ev1), "hipEventCreate")))
    rc = PDF_ERR_HIP;
  if (rc) {
    pdf_batch_destroy(b);
    return rc;
  }
  b->C = count;
  std::vector<BatchCircuit> circ(count);
  std::vector<uint32_t> base(count);
  uint64_t nodes = 0;
  for (uint32_t c = 0; c < count; ++c) {
    pdf_grader *g = b->members[c];
    base[c] = (uint32_t)nodes;
    nodes += g->nl.n;
    b->maxL = std::max(b->maxL, g->nl.L);
    b->max_n = std::max(b->max_n, g->nl.n);
    if (g->num_pairs < 4096) b->wide = false;  // fewer than 64 words: narrow tiles
    BatchCircuit &d = circ[c];
    d.net = g->net();
    d.recs = g->d_rec;
    d.pl = planes(g, g->num_pairs);
    d.va = g->d_va;
    d.vstride = g->vstride;
    d.first0 = g->d_first0;
    d.first1 = g->d_first1;
    d.list = g->d_list;
    d.list_cnt = g->d_list_cnt;
    d.list_stride = g->nl.n;
    d.cstate = g->d_cstate;
    d.n_cstate = g->n_cstate;
    d.actmap = g->d_actmap;
    d.act_list = g->d_act_list;
    d.act_count = g->d_act_count;
    d.gate = g->d_gate;
    d.chunk_total = g->d_chunk_total;
    d.grand = g->d_grand;
    d.err = g->d_err;
    d.width = (uint32_t)g->num_pairs;
  }
  if (nodes >= 0xFFFFFFFFull) {
    set_error("pdf_batch_create: more than 2^32 - 1 nodes in the batch");
    pdf_batch_destroy(b);
    return PDF_ERR_ARG;
  }
  b->total_nodes = (uint32_t)nodes;
  // the reverse pass over lists the forward launches build where it applies
  // (256-circuit batch: 35.7 ms a grade against 37.9 with a scan launch)
  b->tune_rev_compact = b->wide && count <= (1u << (32 - kBatchTileBits)) ? 2 : 1;
  rc = dalloc(&b->d_circ, count);
  if (!rc) rc = dalloc(&b->d_node_base, count);
  if (!rc) rc = dalloc(&b->d_totals, count);
  if (!rc) rc = dalloc(&b->d_errs, count);
  if (!rc && (!hip_ok(hipMemcpy(b->d_circ, circ.data(), count * sizeof(BatchCircuit), hipMemcpyHostToDevice), "hipMemcpy") ||
              !hip_ok(hipMemcpy(b->d_node_base, base.data(), count * 4, hipMemcpyHostToDevice), "hipMemcpy")))
    rc = PDF_ERR_HIP;
  if (!rc) rc = batch_build_segments(b);
  if (!rc) rc = batch_alloc_scratch(b);
  if (rc) {
    pdf_batch_destroy(b);
    return rc;
  }
  *out = b;
  return PDF_OK;
}

void pdf_batch_destroy(pdf_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_drop_graph(b);
  dfree(b->d_circ);
  dfree(b->d_node_base);
  dfree(b->d_segs);
  dfree(b->d_scan);
  dfree(b->d_work);
  dfree(b->d_work_cnt);
  dfree(b->d_blist);
  dfree(b->d_blist_cnt);
  dfree(b->d_cs);
  dfree(b->d_hs);
  dfree(b->d_totals);
  dfree(b->d_errs);
  for (pdf_grader *g : b->members) pdf_grader_destroy(g);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int pdf_batch_grade(pdf_batch *b, int robust, uint64_t *totals, float *ms) {
  if (!b) { set_error("null batch"); return PDF_ERR_ARG; }
  HIPCHK(hipSetDevice(b->device));
  if (ms) HIPCHK(hipEventRecord(b->ev0, b->stream));
  int rc = batch_run(b, robust);
  if (rc) return rc;
  if (ms) HIPCHK(hipEventRecord(b->ev1, b->stream));
  rc = batch_results(b, totals);
  if (ms) HIPCHK(hipEventElapsedTime(ms, b->ev0, b->ev1));
  for (pdf_grader *g : b->members) {  // the members' getters read this grade
    g->stage = 4;
    g->stage_chunk = 0;
  }
  return rc;
}

int pdf_batch_member(pdf_batch *b, uint32_t circuit, pdf_grader **out) {
  if (!b || !out || circuit >= b->C) { set_error("pdf_batch_member: bad argument"); return PDF_ERR_ARG; }
  *out = b->members[circuit];
  return PDF_OK;
}

int pdf_batch_set_tuning(pdf_batch *b, int key, int64_t value) {
  if (!b) { set_error("null batch"); return PDF_ERR_ARG; }
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  switch (key) {
    case PDF_TUNE_USE_GRAPH:
      if (value != 0 && value != 1) break;
      b->tune_use_graph = (uint32_t)value;
      batch_drop_graph(b);
      return PDF_OK;
    case PDF_TUNE_COVER_WAVES:
      if (value < 1 || value > 65535) break;
      b->tune_cover_waves = (uint32_t)value;
      batch_drop_graph(b);
      return batch_alloc_scratch(b);
    case PDF_TUNE_SEGMENT_ORDER:
      if (value != 0 && value != 1) break;
      b->tune_order = (uint32_t)value;
      batch_drop_graph(b);
      return batch_build_segments(b);
    case PDF_TUNE_REV_COMPACT: {
      if (value < 0 || value > 2) break;
      if (value == 2 && !(b->wide && b->C <= (1u << (32 - kBatchTileBits)))) {
        set_error("pdf_batch_set_tuning: forward-built reverse lists need chunks of >= 4,096 pairs and at most " +
                  std::to_string(1u << (32 - kBatchTileBits)) + " circuits");
        return PDF_ERR_UNSUPPORTED;
      }
      const uint32_t was = b->tune_rev_compact;
      b->tune_rev_compact = (uint32_t)value;
      batch_drop_graph(b);
      return (was == 2) != (value == 2) ? batch_build_segments(b) : PDF_OK;
    }
    case PDF_TUNE_WORDS_PER_LANE:
      if (value != 1 && value != 2) break;
      b->tune_wpl = (uint32_t)value;
      batch_drop_graph(b);
      return batch_build_segments(b);
    default:
      set_error("pdf_batch_set_tuning: unknown key");
      return PDF_ERR_ARG;
  }
  set_error("pdf_batch_set_tuning: value out of range");
  return PDF_ERR_ARG;
}

int pdf_batch_autotune(pdf_batch *b, int robust, uint32_t reps, pdf_batch_tuning *best) {
  if (!b || reps == 0) { set_error("pdf_batch_autotune: bad argument"); return PDF_ERR_ARG; }
  // every combination of the knobs' values (96 at most: segment order x
  // cover waves x graph replay x reverse pass x words per lane),
  // each configuration timed once, so interactions between the knobs (the
  // forward tile width against the reverse pass's waves, say) are measured
  // rather than assumed away
  struct Knob {
    int key;
    std::vector<uint32_t> values;
  };
  const Knob knobs[] = {{PDF_TUNE_SEGMENT_ORDER, {0, 1}},
                        {PDF_TUNE_COVER_WAVES, {256, 512, 1024, 2048}},
                        {PDF_TUNE_USE_GRAPH, {0, 1}},
                        {PDF_TUNE_REV_COMPACT, b->wide && b->C <= (1u << (32 - kBatchTileBits))
                                                   ? std::vector<uint32_t>{0, 1, 2}
                                                   : std::vector<uint32_t>{0, 1}},
                        {PDF_TUNE_WORDS_PER_LANE, b->wide ? std::vector<uint32_t>{1, 2} : std::vector<uint32_t>{1}}};
  constexpr int kKnobs = 5;
  uint32_t cur[kKnobs] = {b->tune_order, b->tune_cover_waves, b->tune_use_graph, b->tune_rev_compact,
                          b->wide ? b->tune_wpl : 1u};
  std::vector<std::pair<std::vector<uint32_t>, float>> seen;
  std::vector<uint64_t> ref(b->C), tot(b->C);
  bool have_ref = false;
  int rc;
  auto apply = [&](const uint32_t *v) {
    int r = PDF_OK;
    for (int i = 0; i < kKnobs && !r; ++i) r = pdf_batch_set_tuning(b, knobs[i].key, v[i]);
    return r;
  };
  auto timed = [&](const uint32_t *v, float *med) {
    const std::vector<uint32_t> key(v, v + kKnobs);
    for (const auto &e : seen)
      if (e.first == key) { *med = e.second; return PDF_OK; }
    int r = apply(v);
    if (r == PDF_ERR_NOMEM) {  // its buffers do not fit: never the choice
      *med = std::numeric_limits<float>::infinity();
      seen.emplace_back(key, *med);
      return PDF_OK;
    }
    if (r) return r;
    std::vector<float> t(reps + 1);
    for (uint32_t i = 0; i <= reps; ++i)  // the first run warms up (and captures)
      if ((r = pdf_batch_grade(b, robust, tot.data(), &t[i]))) return r;
    if (!have_ref) {
      ref = tot;
      have_ref = true;
    } else if (tot != ref) {
      set_error("pdf_batch_autotune: totals differ between configurations");
      return (int)PDF_ERR_STATE;
    }
    std::sort(t.begin() + 1, t.end());
    *med = t[1 + reps / 2];
    seen.emplace_back(key, *med);
    return PDF_OK;
  };
  float best_ms;
  if ((rc = timed(cur, &best_ms))) return rc;
  uint32_t idx[kKnobs] = {};  // odometer over the knobs' value lists
  for (;;) {
    uint32_t trial[kKnobs];
    for (int i = 0; i < kKnobs; ++i) trial[i] = knobs[i].values[idx[i]];
    float ms;
    if ((rc = timed(trial, &ms))) return rc;
    if (ms < best_ms) {
      best_ms = ms;
      std::copy(trial, trial + kKnobs, cur);
    }
    int i = 0;
    while (i < kKnobs && ++idx[i] == knobs[i].values.size()) idx[i++] = 0;
    if (i == kKnobs) break;
  }
  if ((rc = apply(cur))) return rc;
  if (best) *best = pdf_batch_tuning{cur[0], cur[1], cur[2], cur[3], cur[4], best_ms};
  return PDF_OK;
}

}  // extern "C"
