/* bitsliced.c -- a second CPU grader: 64 pairs per machine word, OpenMP over
 * pair words (sim, mark, removal) and over first-detection pairs (cover).
 *
 * Test infrastructure only: a cross-check of the literal restatement in
 * oracle.c (the two share no code beyond the netlist struct) and the
 * multi-core CPU baseline of the benchmark.  Semantics are the same as
 * oracle.c's, the reference's:
 *   frames     simkernel.cuh:115-148 (pair (V[p], V[p+1 mod P]) or explicit
 *              (A_p, B_p); NOT inverts; AND/NAND/OR/NOR fold left to right;
 *              BUFF/DFF/FROM copy fanin 0; other types with >= 2 fanins S0)
 *   marks      markkernel.cuh:228-267 (t and the fanin-0 mark and OR over on-path
 *              positions of AND over off-path positions of S), removal
 *              markkernel.cuh:279-291
 *   first      mergekernel.cuh:141-197 (earliest T0 / T1 pair per node)
 *   coverage   coverkernel.cuh:194-228 (c/h DP per pair; only pairs that are
 *              some node's first detection can count anything)
 * Word layout: word-major planes, plane[w * n + g] = 64 pairs of node g. */
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

enum { UNKN, INPT, AND, NAND, OR, NOR, XOR, XNOR, BUFF, NOT, FROM, DFF };

/* Bits q0 .. q0 + 63 of one PI row (words (P + 63) / 64), positions taken mod P. */
static uint64_t row_word(const uint64_t *row, uint64_t P, uint64_t q0) {
  q0 %= P;
  const unsigned s = (unsigned)(q0 % 64);
  if (q0 + 64 <= P) return s ? (row[q0 / 64] >> s) | (row[q0 / 64 + 1] << (64 - s)) : row[q0 / 64];
  uint64_t v = 0;
  for (int j = 0; j < 64; ++j) {
    const uint64_t q = (q0 + (uint64_t)j) % P;
    v |= ((row[q / 64] >> (q % 64)) & 1ULL) << j;
  }
  return v;
}

/* Off-path predicate over 64 pairs: on-path frames (ao, bo), off-path (af, bf). */
static uint64_t sens(int robust, int andlike, uint64_t ao, uint64_t bo, uint64_t af, uint64_t bf) {
  const uint64_t rise = ~ao & bo, fall = ao & ~bo;
  if (andlike) {
    if (robust) return (rise | fall) & af & bf;
    return (fall & bf) | (rise & af & bf);
  }
  if (robust) return (rise | fall) & ~af & ~bf;
  return (rise & ~bf) | (fall & ~af & ~bf);
}

static int add_ovf(uint64_t *acc, uint64_t v) { return __builtin_add_overflow(*acc, v, acc); }

int bsl_grade(const orc_netlist *nl, const uint64_t *a_bits, const uint64_t *b_bits, uint64_t num_pairs,
              uint64_t chunk, int robust, int threads, uint32_t *first_t0, uint32_t *first_t1,
              uint64_t *gate_counts, uint64_t *total, uint32_t *flags) {
  if (!nl || !a_bits || !first_t0 || !first_t1 || !total || !flags || num_pairs == 0) return -1;
  if (chunk == 0 || chunk > num_pairs) chunk = num_pairs;
  if (chunk % 64 && chunk != num_pairs) return -1;
  if (threads < 1) threads = 1;
  const uint32_t n = nl->n;
  const uint64_t vwords = (num_pairs + 63) / 64, cwords = (chunk + 63) / 64;
  uint64_t *A = malloc(sizeof(uint64_t) * cwords * n), *B = malloc(sizeof(uint64_t) * cwords * n);
  uint64_t *M = malloc(sizeof(uint64_t) * cwords * n);
  uint64_t *gc = calloc((size_t)threads * n, sizeof(uint64_t));
  uint64_t *cb = calloc((size_t)threads * n, sizeof(uint64_t));
  uint64_t *hb = calloc((size_t)threads * n, sizeof(uint64_t));
  uint8_t *act = malloc(chunk);
  uint32_t *plist = malloc(sizeof(uint32_t) * chunk);
  int rc = 0;
  if (!A || !B || !M || !gc || !cb || !hb || !act || !plist) {
    rc = -5;
    goto out;
  }
  for (uint32_t g = 0; g < n; ++g) first_t0[g] = first_t1[g] = 0xFFFFFFFFu;
  if (gate_counts) memset(gate_counts, 0, sizeof(uint64_t) * n);
  *total = 0;
  *flags = 0;
  uint32_t ovf = 0;
  for (uint64_t start = 0; start < num_pairs; start += chunk) {
    const uint64_t W = num_pairs - start < chunk ? num_pairs - start : chunk;
    const uint64_t words = (W + 63) / 64;
    /* sim + forward mark, then removal: every word independent */
#pragma omp parallel for num_threads(threads) schedule(static)
    for (uint64_t w = 0; w < words; ++w) {
      uint64_t *a = A + w * n, *b = B + w * n, *m = M + w * n;
      const uint64_t p0 = start + 64 * w;
      const uint64_t live = (W - 64 * w) >= 64 ? ~0ULL : ((1ULL << (W - 64 * w)) - 1);
      for (uint32_t g = 0; g < n; ++g) {
        const uint8_t ty = nl->type[g];
        const uint32_t f0 = nl->fin_off[g], nfi = nl->fin_off[g + 1] - f0;
        if (ty == INPT) {
          a[g] = row_word(a_bits + (uint64_t)g * vwords, num_pairs, p0);
          b[g] = b_bits ? row_word(b_bits + (uint64_t)g * vwords, num_pairs, p0)
                        : row_word(a_bits + (uint64_t)g * vwords, num_pairs, p0 + 1);
          m[g] = (a[g] ^ b[g]) & live;
          continue;
        }
        const uint32_t x0 = nl->fin[f0];
        uint64_t va = a[x0], vb = b[x0];
        if (ty == NOT) { va = ~va; vb = ~vb; }
        if (nfi >= 2) {
          if (ty == AND || ty == NAND || ty == OR || ty == NOR) {
            for (uint32_t j = 1; j < nfi; ++j) {
              const uint32_t x = nl->fin[f0 + j];
              if (ty == AND) { va &= a[x]; vb &= b[x]; }
              else if (ty == OR) { va |= a[x]; vb |= b[x]; }
              else if (ty == NAND) { va = ~(va & a[x]); vb = ~(vb & b[x]); }
              else { va = ~(va | a[x]); vb = ~(vb | b[x]); }
            }
          } else if (ty != BUFF && ty != DFF && ty != FROM) {
            va = vb = 0;
          }
        }
        a[g] = va;
        b[g] = vb;
        const uint64_t t = va ^ vb;
        uint64_t mk = 0;
        if (ty == BUFF || ty == DFF || ty == FROM || ty == NOT) {
          mk = t & m[x0];
        } else if (ty >= AND && ty <= XNOR) {
          mk = t & m[x0];
          if (nfi >= 2 && mk) {
            if (ty == XOR || ty == XNOR) {
              mk = 0;
            } else {
              const int andlike = ty == AND || ty == NAND;
              uint64_t any = 0;
              for (uint32_t i = 0; i < nfi; ++i) {
                const uint32_t xi = nl->fin[f0 + i];
                uint64_t all = ~0ULL;
                for (uint32_t j = 0; j < nfi && all; ++j) {
                  if (j == i) continue;
                  const uint32_t xj = nl->fin[f0 + j];
                  all &= sens(robust, andlike, a[xi], b[xi], a[xj], b[xj]);
                }
                any |= all;
              }
              mk &= any;
            }
          }
        }
        m[g] = mk;
      }
      for (uint32_t g = n; g-- > 0;) {
        if (!m[g] || nl->po[g]) continue;
        uint64_t acc = 0;
        for (uint32_t k = nl->fout_off[g]; k < nl->fout_off[g + 1]; ++k) acc |= m[nl->fout[k]];
        m[g] &= acc;
      }
    }
    /* merge: earliest T0 / T1 pair of each node (chunks run in order) */
#pragma omp parallel for num_threads(threads) schedule(static)
    for (uint32_t g = 0; g < n; ++g) {
      int have0 = 0, have1 = 0;
      for (uint64_t w = 0; w < words && !(have0 && have1); ++w) {
        const uint64_t m = M[w * n + g], b = B[w * n + g];
        const uint64_t t0 = m & ~b, t1 = m & b;
        if (t0 && !have0) {
          const uint32_t p = (uint32_t)(start + 64 * w + (uint64_t)__builtin_ctzll(t0));
          if (p < first_t0[g]) first_t0[g] = p;
          have0 = 1;
        }
        if (t1 && !have1) {
          const uint32_t p = (uint32_t)(start + 64 * w + (uint64_t)__builtin_ctzll(t1));
          if (p < first_t1[g]) first_t1[g] = p;
          have1 = 1;
        }
      }
    }
    /* the chunk's distinct first-detection pairs */
    memset(act, 0, W);
    uint64_t nact = 0;
    for (uint32_t g = 0; g < n; ++g) {
      const uint32_t x[2] = {first_t0[g], first_t1[g]};
      for (int k = 0; k < 2; ++k)
        if (x[k] != 0xFFFFFFFFu && x[k] >= start && x[k] - start < W && !act[x[k] - start]) {
          act[x[k] - start] = 1;
          plist[nact++] = (uint32_t)(x[k] - start);
        }
    }
    /* cover: the DP of each first-detection pair over its marked nodes */
    uint64_t chunk_total = 0;
#pragma omp parallel num_threads(threads) reduction(| : ovf)
    {
      const int tid = omp_get_thread_num();
      uint64_t *c = cb + (size_t)tid * n, *h = hb + (size_t)tid * n;
      uint64_t *tgc = gc + (size_t)tid * n;
      uint64_t local = 0;
#pragma omp for schedule(dynamic, 1)
      for (uint64_t k = 0; k < nact; ++k) {
        const uint64_t rel = plist[k], w = rel / 64;
        const int j = (int)(rel % 64);
        const uint32_t pid = (uint32_t)(start + rel);
        const uint64_t *m = M + w * n;
        for (uint32_t g = n; g-- > 0;) {
          if (!((m[g] >> j) & 1)) continue;
          uint64_t C = 0, H = nl->po[g] == 1 ? 1 : 0;
          for (uint32_t q = nl->fout_off[g]; q < nl->fout_off[g + 1]; ++q) {
            const uint32_t fo = nl->fout[q];
            if (!((m[fo] >> j) & 1)) continue;
            ovf |= (uint32_t)add_ovf(&C, c[fo]) | (uint32_t)add_ovf(&H, h[fo]);
          }
          if (first_t0[g] == pid || first_t1[g] == pid) {
            ovf |= (uint32_t)add_ovf(&C, H);
            H = 0;
          }
          if (nl->type[g] == INPT) H = 0;
          c[g] = C;
          h[g] = H;
          if (C) {
            ovf |= (uint32_t)add_ovf(&tgc[g], C);
            if (g < nl->level_off[1] && nl->type[g] == INPT) ovf |= (uint32_t)add_ovf(&local, C);
          }
        }
        for (uint32_t g = 0; g < n; ++g)
          if ((m[g] >> j) & 1) c[g] = h[g] = 0;
      }
#pragma omp critical
      ovf |= (uint32_t)add_ovf(&chunk_total, local);
    }
    if (add_ovf(total, chunk_total)) ovf |= 1u;
  }
  if (gate_counts)
    for (int t = 0; t < threads; ++t)
      for (uint32_t g = 0; g < n; ++g)
        if (add_ovf(&gate_counts[g], gc[(size_t)t * n + g])) ovf |= 1u;
  if (ovf) *flags |= ORC_FLAG_U64_OVERFLOW;
out:
  free(A); free(B); free(M); free(gc); free(cb); free(hb); free(act); free(plist);
  return rc;
}
