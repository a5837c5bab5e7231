/* oracle.h -- CPU restatement of the reference grading path (test
 * infrastructure only: the checker for tests, smoke() and the CPU baseline). */
#ifndef PDF_ORACLE_H
#define PDF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_netlist {
  uint32_t n, num_pi, L;
  const uint8_t *type, *po;
  const uint32_t *fin_off, *fin, *fout_off, *fout, *level_off;
} orc_netlist;

typedef struct orc_options {
  int robust;
  uint32_t merge_window; /* 0 = whole chunk; 4224 = the reference's window */
} orc_options;

#define ORC_FLAG_EXCEEDS_U16 1u   /* a c/h value left the reference's 16 bits */
#define ORC_FLAG_EXCEEDS_U32 2u   /* a chunk total left the reference's uint32 */
#define ORC_FLAG_U64_OVERFLOW 4u  /* exact uint64 accumulation overflowed */
/* the part of ORC_FLAG_EXCEEDS_U16 a sparse grader can see: a fanout sum or
 * final c/h past 16 bits at a marked node of a pair where some node is new */
#define ORC_FLAG_EXCEEDS_U16_DP 8u

/* Segments of S that the reference's merge launch loop never reaches
 * (gpuMergeSegments, mergekernel.cuh:229-250, count += BLOCK_STEP). */
uint64_t orc_ref_unmerged_segments(uint64_t S);

const uint8_t *orc_table(int which);
uint16_t orc_mark_lut(int robust, int idx);

int orc_sim_chunk(const orc_netlist *nl, const uint64_t *a_bits, const uint64_t *b_bits,
                  uint64_t num_pairs, uint64_t start, uint64_t W, uint8_t *codes);
int orc_mark_chunk(const orc_netlist *nl, const uint8_t *codes, uint64_t W, int robust,
                   uint8_t *marks, uint8_t *marks_fwd);
int orc_merge_chunk(const orc_netlist *nl, const uint8_t *codes, const uint8_t *marks, uint64_t W,
                    uint64_t start, uint32_t merge_window, uint32_t *first_t0, uint32_t *first_t1);
int orc_cover_chunk(const orc_netlist *nl, const uint8_t *marks, uint64_t W, uint64_t start,
                    const uint32_t *first_t0, const uint32_t *first_t1, uint64_t *gate_counts,
                    uint64_t *chunk_total, uint32_t *flags);
int orc_cover_new_chunk(const orc_netlist *nl, const uint8_t *marks, const uint8_t *newm, uint64_t W,
                        uint64_t *gate_counts, uint64_t *chunk_total, uint32_t *flags, uint64_t *pi_c);
int orc_segments(const orc_netlist *nl, uint32_t K, uint32_t *out, uint64_t cap, uint64_t *count);
int orc_merge_segs_chunk(const orc_netlist *nl, const uint8_t *codes, const uint8_t *marks, uint64_t W,
                         uint64_t start, uint32_t merge_window, uint32_t K, const uint32_t *segs, uint64_t S,
                         uint32_t *seg_t0, uint32_t *seg_t1);
int orc_new_from_segments(const orc_netlist *nl, uint64_t W, uint64_t start, uint32_t K, const uint32_t *segs,
                          uint64_t S, const uint32_t *seg_t0, const uint32_t *seg_t1, uint8_t *newm);
int orc_grade_segs(const orc_netlist *nl, const uint64_t *a_bits, const uint64_t *b_bits, uint64_t num_pairs,
                   uint64_t chunk, const orc_options *opt, uint32_t K, const uint32_t *segs, uint64_t S,
                   uint32_t *seg_t0, uint32_t *seg_t1, uint64_t *gate_counts, uint64_t *chunk_totals,
                   uint64_t *total, uint32_t *flags);
int orc_grade(const orc_netlist *nl, const uint64_t *a_bits, const uint64_t *b_bits, uint64_t num_pairs,
              uint64_t chunk, const orc_options *opt, uint32_t *first_t0, uint32_t *first_t1,
              uint64_t *gate_counts, uint64_t *chunk_totals, uint64_t *total, uint32_t *flags);

#ifdef __cplusplus
}
#endif
#endif
