/*
 * oracle.c -- CPU restatement of the reference's grading path, used ONLY as
 * test infrastructure: the checker in tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  It is never linked into the product library.
 *
 * It follows the reference one pattern (one byte) at a time, in the same
 * layout (node rows x pattern columns, a chunk of W patterns at a time):
 *   sim    simkernel.cuh:96-156 (kernSimulateP1) with the tables of
 *          simkernel.cuh:157-165 (loadSimLUTs) and simLUT at 62-95
 *   mark   markkernel.cuh:36-72 (LUT_R/LUT_NR, LUT_16, robust/nonrobust),
 *          215-270 (kernMarkPathSegments), 271-293 (kernRemoveInvalidMarks),
 *          295-342 (level loops)
 *   merge  mergekernel.cuh:116-198 (kernSegmentReduce2, K=1 segments)
 *   cover  coverkernel.cuh:171-230 (kernCover), 77-105 (kernSumSingle),
 *          233-305 (gpuCountPaths level loop)
 *   K >= 2 mergekernel.cuh:91-110 (has_marked_segment), coverkernel.cuh:106-154
 *          (findActiveThreads<K>); the segment list itself is this project's
 *          enumeration (orc_segments), the reference's lives in util/
 *   loop   main.cu:171-231 (chunk loop, totals += coverage)
 *
 * Deliberate differences (resource-limit artefacts of the reference):
 *   - merge examines every pattern of the chunk; `merge_window` > 0 restores
 *     the reference's 4,224-pattern window (mergekernel.cuh:125-130);
 *   - path counts are exact uint64 instead of 16-bit fields
 *     (coverkernel.cuh:41-43) and the chunk total is exact instead of a
 *     uint32 partial (coverkernel.cuh:79-101); flags report when the
 *     reference would have left its range.
 *
 * Pinning (tests/test_oracle.py): the tables are checked against two-frame
 * Boolean evaluation; sim/mark against an independent Boolean restatement;
 * cover against brute-force enumeration of marked PI->PO paths on small
 * circuits and a hand-derived c17 case; chunk-size invariance and the
 * explicit-pair mode against sequence mode.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* codes (S0,S1,T0,T1) = (0,1,2,3): bit0 = final value, bit1 = transition */
enum { S0 = 0, S1 = 1, T0 = 2, T1 = 3 };
/* node types, numbered as in include/pdfgrade.h */
enum { UNKN = 0, INPT = 1, AND = 2, NAND = 3, OR = 4, NOR = 5, XOR = 6, XNOR = 7,
       BUFF = 8, NOT = 9, FROM = 10, DFF = 11 };

/* simkernel.cuh:158-165.  A 4x4 table is read as tex2D(lut, val, r), i.e.
 * element [r*4 + val]; stable is read as tex2D(stable, v[p], v[p+1]), i.e.
 * element [v[p+1]*2 + v[p]]. */
static const uint8_t nand2[16] = {S1, S1, S1, S1, S1, S0, T1, T0, S1, T1, T1, S1, S1, T0, S1, T0};
static const uint8_t and2[16] = {S0, S0, S0, S0, S0, S1, T0, T1, S0, T0, T0, S0, S0, T1, S0, T1};
static const uint8_t nor2[16] = {S1, S0, T1, T0, S0, S0, S0, S0, T1, S0, T1, S0, T0, S0, S0, T0};
static const uint8_t or2[16] = {S0, S1, T0, T1, S1, S1, S1, S1, T0, S1, T0, S1, T1, S1, S1, T1};
static const uint8_t xnor2[16] = {S1, S0, T1, T0, S0, S1, T0, T1, T1, T0, S1, S0, T0, T1, S0, S1};
static const uint8_t xor2[16] = {S0, S1, T0, T1, S1, S0, T1, T0, T0, T1, S0, S1, T1, T0, S1, S0};
static const uint8_t stable_lut[4] = {S0, T0, T1, S1};
static const uint8_t not_lut[4] = {S1, S0, T1, T0};

/* markkernel.cuh:36-41 */
static const uint16_t LUT_NR[6] = {0x2A00, 0x2A00, 0x5100, 0x5100, 0x0000, 0x0000};
static const uint16_t LUT_R[6] = {0x2200, 0x2200, 0x1100, 0x1100, 0x0000, 0x0000};
static uint8_t lut16(uint16_t table, uint16_t on, uint16_t off) {
  return (uint8_t)((table >> ((on << 2) + off)) & 1u);
}

const uint8_t *orc_table(int which) {
  switch (which) {
    case 0: return and2;
    case 1: return nand2;
    case 2: return or2;
    case 3: return nor2;
    case 4: return xor2;
    case 5: return xnor2;
    case 6: return stable_lut;
    case 7: return not_lut;
    default: return NULL;
  }
}
uint16_t orc_mark_lut(int robust, int idx) { return robust ? LUT_R[idx] : LUT_NR[idx]; }

/* simLUT(coalesce_t) of simkernel.cuh:62-95: XOR/XNOR (and anything else)
 * give S0 -- the table-driven scalar overload at 49-60 is never called. */
static uint8_t sim_lut(uint8_t type, uint8_t val, uint8_t r) {
  switch (type) {
    case OR: return or2[r * 4 + val];
    case NOR: return nor2[r * 4 + val];
    case AND: return and2[r * 4 + val];
    case NAND: return nand2[r * 4 + val];
    default: return S0;
  }
}

static inline uint8_t getbit(const uint64_t *rows, uint64_t words, uint32_t row, uint64_t q) {
  return (uint8_t)((rows[(uint64_t)row * words + (q >> 6)] >> (q & 63)) & 1u);
}

/* kernSimulateP1 over one chunk: codes[g*W + p] for pattern start+p. */
int orc_sim_chunk(const orc_netlist *nl, const uint64_t *a_bits, const uint64_t *b_bits,
                  uint64_t num_pairs, uint64_t start, uint64_t W, uint8_t *codes) {
  const uint64_t words = (num_pairs + 63) / 64;
  for (uint32_t g = 0; g < nl->n; ++g) {
    uint8_t *row = codes + (uint64_t)g * W;
    const uint8_t ty = nl->type[g];
    const uint32_t f0 = nl->fin_off[g];
    const uint32_t nfi = nl->fin_off[g + 1] - f0;
    for (uint64_t p = 0; p < W; ++p) {
      const uint64_t q = start + p;
      uint8_t val;
      if (ty == INPT) {
        /* simkernel.cuh:115-124: pair (V[q], V[q+1]), wrapping to V[0] */
        const uint8_t v1 = getbit(a_bits, words, g, q);
        const uint8_t v2 = b_bits ? getbit(b_bits, words, g, q)
                                  : getbit(a_bits, words, g, (q + 1) % num_pairs);
        val = stable_lut[v2 * 2 + v1];
      } else if (ty == DFF || ty == BUFF || ty == FROM) {
        val = codes[(uint64_t)nl->fin[f0] * W + p]; /* simkernel.cuh:126-130 */
      } else {
        val = codes[(uint64_t)nl->fin[f0] * W + p]; /* simkernel.cuh:135 */
        if (ty == NOT) val = not_lut[val];
        for (uint32_t j = 1; j < nfi; ++j) /* simkernel.cuh:142-148 */
          val = sim_lut(ty, val, codes[(uint64_t)nl->fin[f0 + j] * W + p]);
      }
      row[p] = val;
    }
  }
  return 0;
}

/* robust()/nonrobust() of markkernel.cuh:43-72, one pattern. */
static uint8_t sens(const uint16_t *lut, uint32_t f1, const uint8_t *codes, uint64_t W, uint64_t p,
                    const uint32_t *fins, uint32_t nfi, uint8_t type, uint8_t result, uint8_t oldmark) {
  uint8_t tmp = 0xFF;
  const uint8_t sim_fin1 = codes[(uint64_t)fins[f1] * W + p];
  for (uint32_t fin2 = 0; fin2 < nfi; ++fin2) {
    if (f1 == fin2) continue;
    const uint8_t sim_fin2 = codes[(uint64_t)fins[fin2] * W + p];
    tmp = (uint8_t)(tmp && lut16(lut[type - 2], sim_fin1, sim_fin2));
  }
  tmp = (uint8_t)(tmp & oldmark);
  return (uint8_t)(tmp | result);
}

int orc_mark_chunk(const orc_netlist *nl, const uint8_t *codes, uint64_t W, int robust,
                   uint8_t *marks, uint8_t *marks_fwd) {
  const uint16_t *lut = robust ? LUT_R : LUT_NR;
  /* forward, level order (markkernel.cuh:308-322) */
  for (uint32_t g = 0; g < nl->n; ++g) {
    const uint8_t ty = nl->type[g];
    const uint32_t *fins = nl->fin + nl->fin_off[g];
    const uint32_t nfi = nl->fin_off[g + 1] - nl->fin_off[g];
    for (uint64_t p = 0; p < W; ++p) {
      const uint8_t row_cache = codes[(uint64_t)g * W + p] > 1;
      uint8_t result = 0;
      switch (ty) {
        case INPT: result = row_cache; break;
        case FROM: case BUFF: case DFF: case NOT:
          result = marks[(uint64_t)fins[0] * W + p];
          break;
        case OR: case NOR: case XOR: case XNOR: case NAND: case AND:
          for (uint32_t fin1 = 0; fin1 < nfi; ++fin1) {
            const uint8_t oldmark = marks[(uint64_t)fins[0] * W + p]; /* markkernel.cuh:254 */
            result = sens(lut, fin1, codes, W, p, fins, nfi, ty, result, oldmark);
          }
          break;
        default: break;
      }
      marks[(uint64_t)g * W + p] = (uint8_t)(result & row_cache);
    }
  }
  if (marks_fwd) memcpy(marks_fwd, marks, (size_t)nl->n * W);
  /* reverse removal (markkernel.cuh:326-342, 279-291) */
  for (uint32_t gi = nl->n; gi-- > 0;) {
    const uint32_t *fots = nl->fout + nl->fout_off[gi];
    const uint32_t nfo = nl->fout_off[gi + 1] - nl->fout_off[gi];
    for (uint64_t p = 0; p < W; ++p) {
      uint8_t r = nl->po[gi] > 0;
      for (uint32_t j = 0; j < nfo; ++j) r |= marks[(uint64_t)fots[j] * W + p];
      marks[(uint64_t)gi * W + p] &= r;
    }
  }
  return 0;
}

/* kernSegmentReduce2<1> (mergekernel.cuh:125-197): earliest marked T0 and
 * T1 pattern per node, min-combined with the previous chunks' values. */
int orc_merge_chunk(const orc_netlist *nl, const uint8_t *codes, const uint8_t *marks, uint64_t W,
                    uint64_t start, uint32_t merge_window, uint32_t *first_t0, uint32_t *first_t1) {
  const uint64_t lim = merge_window ? (W < merge_window ? W : merge_window) : W;
  for (uint32_t g = 0; g < nl->n; ++g) {
    uint32_t bx = 0xFFFFFFFFu, by = 0xFFFFFFFFu;
    for (uint64_t p = 0; p < lim; ++p) {
      if (!marks[(uint64_t)g * W + p]) continue;
      const uint32_t pid = (uint32_t)(start + p);
      if (codes[(uint64_t)g * W + p] & 1u) { if (pid < by) by = pid; }
      else if (pid < bx) bx = pid;
    }
    if (bx < first_t0[g]) first_t0[g] = bx;
    if (by < first_t1[g]) first_t1[g] = by;
  }
  return 0;
}

/* kernCover<K> + kernSumSingle over one chunk, exact uint64; newm[g*W + p]
 * says node g counts as new at pattern start+p (coverkernel.cuh:214-220:
 * some segment holding g has start+p as its first T0 or T1 pattern).
 * pi_c (may be NULL) receives c(g, p) of every primary input g < num_pi,
 * [num_pi][W], the values debugCover prints (coverkernel.cuh:44-66). */
int orc_cover_new_chunk(const orc_netlist *nl, const uint8_t *marks, const uint8_t *newm, uint64_t W,
                        uint64_t *gate_counts, uint64_t *chunk_total, uint32_t *flags, uint64_t *pi_c) {
  uint64_t *c = (uint64_t *)malloc(sizeof(uint64_t) * nl->n);
  uint64_t *h = (uint64_t *)malloc(sizeof(uint64_t) * nl->n);
  if (!c || !h) { free(c); free(h); return -5; }
  uint64_t total = 0;
  for (uint64_t p = 0; p < W; ++p) {
    int any_new = 0; /* a pair with no new node contributes nothing: a sparse grader skips it */
    for (uint32_t gi = 0; gi < nl->n && !any_new; ++gi) any_new = newm[(uint64_t)gi * W + p] != 0;
    for (uint32_t gi = nl->n; gi-- > 0;) {
      const uint8_t cache = marks[(uint64_t)gi * W + p];
      uint64_t cc = 0, hh = 0;
      if (nl->po[gi] == 1) hh = cache > 0; /* coverkernel.cuh:198-201 */
      uint64_t rc = 0, hc = 0;
      for (uint32_t k = nl->fout_off[gi]; k < nl->fout_off[gi + 1]; ++k) {
        const uint32_t fo = nl->fout[k];
        if (__builtin_add_overflow(rc, c[fo], &rc)) *flags |= ORC_FLAG_U64_OVERFLOW;
        if (__builtin_add_overflow(hc, h[fo], &hc)) *flags |= ORC_FLAG_U64_OVERFLOW;
      }
      if (rc > 0xFFFF || hc > 0xFFFF) *flags |= ORC_FLAG_EXCEEDS_U16;
      cc += rc;
      if (__builtin_add_overflow(hh, hc, &hh)) *flags |= ORC_FLAG_U64_OVERFLOW;
      const int is_new = newm[(uint64_t)gi * W + p] != 0;
      /* coverkernel.cuh:214-220 */
      const uint64_t tmp = (cache > 0 && is_new) ? hh : 0;
      if (__builtin_add_overflow((cache > 0) ? cc : 0, tmp, &cc)) *flags |= ORC_FLAG_U64_OVERFLOW;
      hh = (cache > 0 && !is_new) ? hh : 0;
      if (nl->type[gi] == INPT) hh = 0; /* coverkernel.cuh:221-223 */
      if (cc > 0xFFFF || hh > 0xFFFF) *flags |= ORC_FLAG_EXCEEDS_U16;
      if (any_new && cache > 0 && (rc > 0xFFFF || hc > 0xFFFF || cc > 0xFFFF || hh > 0xFFFF))
        *flags |= ORC_FLAG_EXCEEDS_U16_DP;
      c[gi] = cc;
      h[gi] = hh;
      if (gate_counts && __builtin_add_overflow(gate_counts[gi], cc, &gate_counts[gi]))
        *flags |= ORC_FLAG_U64_OVERFLOW;
    }
    /* kernSumSingle: INPT rows among the first levelsize(0) ids */
    for (uint32_t gi = 0; gi < nl->level_off[1]; ++gi)
      if (nl->type[gi] == INPT && __builtin_add_overflow(total, c[gi], &total))
        *flags |= ORC_FLAG_U64_OVERFLOW;
    if (pi_c)
      for (uint32_t gi = 0; gi < nl->num_pi; ++gi) pi_c[(uint64_t)gi * W + p] = c[gi];
  }
  if (total > 0xFFFFFFFFull) *flags |= ORC_FLAG_EXCEEDS_U32;
  *chunk_total = total;
  free(c);
  free(h);
  return 0;
}

/* K = 1 (findActiveThreads<1>, coverkernel.cuh:156-169): node g is new at p
 * when first_T0[g] or first_T1[g] is p. */
int orc_cover_chunk(const orc_netlist *nl, const uint8_t *marks, uint64_t W, uint64_t start,
                    const uint32_t *first_t0, const uint32_t *first_t1, uint64_t *gate_counts,
                    uint64_t *chunk_total, uint32_t *flags) {
  uint8_t *newm = (uint8_t *)calloc((size_t)nl->n * W, 1);
  if (!newm) return -5;
  for (uint32_t g = 0; g < nl->n; ++g) {
    const uint32_t x[2] = {first_t0[g], first_t1[g]};
    for (int k = 0; k < 2; ++k)
      if (x[k] != 0xFFFFFFFFu && x[k] >= start && x[k] - start < W) newm[(uint64_t)g * W + (x[k] - start)] = 1;
  }
  const int rc = orc_cover_new_chunk(nl, marks, newm, W, gate_counts, chunk_total, flags, NULL);
  free(newm);
  return rc;
}

/* Segments of length K, 1 <= K <= 8: the reference's segment<K> list
 * (mergekernel.cuh:205-221) comes from generateSegmentList in its absent
 * util/ library, so this enumeration is the project's definition: for every
 * node g0 in id order, every fanout walk g0 -> g1 -> ... of K nodes, fanouts
 * taken in CSR order, depth first; a walk that reaches a node without
 * fanouts before K nodes ends there and the rest of its entries are
 * 0xFFFFFFFF (the reference skips negative ids, mergekernel.cuh:100-109).
 * K = 1 gives one segment per node, in id order.  Writes min(count, cap)
 * segments of K entries to out (may be NULL) and the total to *count. */
static void seg_walk(const orc_netlist *nl, uint32_t K, uint32_t *path, uint32_t d, uint32_t *out, uint64_t cap,
                     uint64_t *cnt) {
  const uint32_t g = path[d - 1];
  const uint32_t lo = nl->fout_off[g], hi = nl->fout_off[g + 1];
  if (d == K || lo == hi) {
    if (out && *cnt < cap)
      for (uint32_t i = 0; i < K; ++i) out[*cnt * K + i] = i < d ? path[i] : 0xFFFFFFFFu;
    ++*cnt;
    return;
  }
  for (uint32_t k = lo; k < hi; ++k) {
    path[d] = nl->fout[k];
    seg_walk(nl, K, path, d + 1, out, cap, cnt);
  }
}

int orc_segments(const orc_netlist *nl, uint32_t K, uint32_t *out, uint64_t cap, uint64_t *count) {
  if (!nl || !count || K < 1 || K > 8) return -1;
  uint32_t path[8];
  uint64_t cnt = 0;
  for (uint32_t g = 0; g < nl->n; ++g) {
    path[0] = g;
    seg_walk(nl, K, path, 1, out, cap, &cnt);
  }
  *count = cnt;
  return 0;
}

/* kernSegmentReduce2<K> (mergekernel.cuh:116-198): a segment is marked at a
 * pattern when every node of it is (has_marked_segment, 91-110); its earliest
 * marked pattern with node 0 falling (T0) and rising (T1), min-combined with
 * the previous chunks'. */
int orc_merge_segs_chunk(const orc_netlist *nl, const uint8_t *codes, const uint8_t *marks, uint64_t W,
                         uint64_t start, uint32_t merge_window, uint32_t K, const uint32_t *segs, uint64_t S,
                         uint32_t *seg_t0, uint32_t *seg_t1) {
  (void)nl;
  const uint64_t lim = merge_window ? (W < merge_window ? W : merge_window) : W;
  for (uint64_t s = 0; s < S; ++s) {
    const uint32_t *sg = segs + s * K;
    uint32_t bx = 0xFFFFFFFFu, by = 0xFFFFFFFFu;
    for (uint64_t p = 0; p < lim && (bx == 0xFFFFFFFFu || by == 0xFFFFFFFFu); ++p) {
      int all = 1;
      for (uint32_t i = 0; i < K && all; ++i)
        if (sg[i] != 0xFFFFFFFFu) all = marks[(uint64_t)sg[i] * W + p] != 0;
      if (!all) continue;
      const uint32_t pid = (uint32_t)(start + p);
      if (codes[(uint64_t)sg[0] * W + p] & 1u) { if (by == 0xFFFFFFFFu) by = pid; }
      else if (bx == 0xFFFFFFFFu) bx = pid;
    }
    if (bx < seg_t0[s]) seg_t0[s] = bx;
    if (by < seg_t1[s]) seg_t1[s] = by;
  }
  return 0;
}

/* findActiveThreads<K> (coverkernel.cuh:127-154): node g is new at pattern p
 * of the chunk when a segment holding g has p as its first T0 or T1. */
int orc_new_from_segments(const orc_netlist *nl, uint64_t W, uint64_t start, uint32_t K, const uint32_t *segs,
                          uint64_t S, const uint32_t *seg_t0, const uint32_t *seg_t1, uint8_t *newm) {
  memset(newm, 0, (size_t)nl->n * W);
  for (uint64_t s = 0; s < S; ++s) {
    const uint32_t x[2] = {seg_t0[s], seg_t1[s]};
    for (int k = 0; k < 2; ++k) {
      if (x[k] == 0xFFFFFFFFu || x[k] < start || x[k] - start >= W) continue;
      for (uint32_t i = 0; i < K; ++i)
        if (segs[s * K + i] != 0xFFFFFFFFu) newm[(uint64_t)segs[s * K + i] * W + (x[k] - start)] = 1;
    }
  }
  return 0;
}

/* The chunk loop with segments of length K (main.cu:185-223): sim, mark,
 * segment merge, cover.  seg_t0/seg_t1 [S] and gate_counts [n] are
 * accumulated (the caller initializes them). */
int orc_grade_segs(const orc_netlist *nl, const uint64_t *a_bits, const uint64_t *b_bits, uint64_t num_pairs,
                   uint64_t chunk, const orc_options *opt, uint32_t K, const uint32_t *segs, uint64_t S,
                   uint32_t *seg_t0, uint32_t *seg_t1, uint64_t *gate_counts, uint64_t *chunk_totals,
                   uint64_t *total, uint32_t *flags) {
  if (!nl || !a_bits || !segs || !seg_t0 || !seg_t1 || !total || !flags || num_pairs == 0 || K < 1 || K > 8)
    return -1;
  if (chunk == 0 || chunk > num_pairs) chunk = num_pairs;
  uint8_t *codes = (uint8_t *)malloc((size_t)nl->n * chunk);
  uint8_t *marks = (uint8_t *)malloc((size_t)nl->n * chunk);
  uint8_t *newm = (uint8_t *)malloc((size_t)nl->n * chunk);
  int rc = 0;
  if (!codes || !marks || !newm) rc = -5;
  *total = 0;
  *flags = 0;
  uint64_t idx = 0;
  for (uint64_t start = 0; !rc && start < num_pairs; start += chunk, ++idx) {
    const uint64_t W = (num_pairs - start) < chunk ? (num_pairs - start) : chunk;
    uint64_t ct = 0;
    orc_sim_chunk(nl, a_bits, b_bits, num_pairs, start, W, codes);
    orc_mark_chunk(nl, codes, W, opt ? opt->robust : 0, marks, NULL);
    orc_merge_segs_chunk(nl, codes, marks, W, start, opt ? opt->merge_window : 0, K, segs, S, seg_t0, seg_t1);
    orc_new_from_segments(nl, W, start, K, segs, S, seg_t0, seg_t1, newm);
    rc = orc_cover_new_chunk(nl, marks, newm, W, gate_counts, &ct, flags, NULL);
    if (rc) break;
    if (chunk_totals) chunk_totals[idx] = ct;
    if (__builtin_add_overflow(*total, ct, total)) *flags |= ORC_FLAG_U64_OVERFLOW;
  }
  free(codes);
  free(marks);
  free(newm);
  return rc;
}

int orc_grade(const orc_netlist *nl, const uint64_t *a_bits, const uint64_t *b_bits, uint64_t num_pairs,
              uint64_t chunk, const orc_options *opt, uint32_t *first_t0, uint32_t *first_t1,
              uint64_t *gate_counts, uint64_t *chunk_totals, uint64_t *total, uint32_t *flags) {
  if (!nl || !a_bits || !first_t0 || !first_t1 || !total || !flags || num_pairs == 0) return -1;
  if (chunk == 0 || chunk > num_pairs) chunk = num_pairs;
  uint8_t *codes = (uint8_t *)malloc((size_t)nl->n * chunk);
  uint8_t *marks = (uint8_t *)malloc((size_t)nl->n * chunk);
  if (!codes || !marks) { free(codes); free(marks); return -5; }
  *total = 0;
  *flags = 0;
  uint64_t idx = 0;
  for (uint64_t start = 0; start < num_pairs; start += chunk, ++idx) {
    const uint64_t W = (num_pairs - start) < chunk ? (num_pairs - start) : chunk;
    uint64_t ct = 0;
    orc_sim_chunk(nl, a_bits, b_bits, num_pairs, start, W, codes);
    orc_mark_chunk(nl, codes, W, opt ? opt->robust : 0, marks, NULL);
    orc_merge_chunk(nl, codes, marks, W, start, opt ? opt->merge_window : 0, first_t0, first_t1);
    int rc = orc_cover_chunk(nl, marks, W, start, first_t0, first_t1, gate_counts, &ct, flags);
    if (rc) { free(codes); free(marks); return rc; }
    if (chunk_totals) chunk_totals[idx] = ct;
    if (__builtin_add_overflow(*total, ct, total)) *flags |= ORC_FLAG_U64_OVERFLOW;
  }
  free(codes);
  free(marks);
  return 0;
}

/* gpuMergeSegments (mergekernel.cuh:229-250): remaining_blocks = ceil(S/20);
 * each pass launches block_y = min(BLOCK_STEP, remaining_blocks) blocks of
 * PARALLEL_SEGS = 20 segments at startSegment = count, then count +=
 * BLOCK_STEP (65535, not 65535 * 20) and remaining_blocks -= BLOCK_STEP (or
 * becomes 0).  Segment sid is examined iff some pass has count <= sid <
 * count + 20 * block_y.  Counted here segment by segment. */
uint64_t orc_ref_unmerged_segments(uint64_t S) {
  uint64_t remaining = S / 20 + (S % 20 > 0), count = 0, missed = 0;
  uint64_t passes = 0;
  uint64_t starts[64], ends[64];
  do {
    const uint64_t block_y = remaining > 65535 ? 65535 : remaining;
    if (passes < 64) {
      starts[passes] = count;
      ends[passes] = count + 20 * block_y;
    }
    ++passes;
    count += 65535;
    if (remaining < 65535) remaining = 0;
    if (remaining >= 65535) remaining -= 65535;
  } while (remaining > 0);
  if (passes > 64) return S; /* beyond any netlist this oracle is run on */
  for (uint64_t sid = 0; sid < S; ++sid) {
    int hit = 0;
    for (uint64_t k = 0; k < passes && !hit; ++k) hit = sid >= starts[k] && sid < ends[k];
    missed += !hit;
  }
  return missed;
}
