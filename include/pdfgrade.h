/*
 * pdfgrade.h -- C ABI of the MI355X path-delay fault grader.
 *
 * The grading path is sim -> mark -> merge -> cover, called once per chunk of
 * vector pairs, exactly as the reference's chunk loop does (main.cu:171-231).
 * Every entry point returns PDF_OK (0) or a negative status; the text of the
 * last failure on the calling thread is available from pdf_last_error().
 * No entry point calls exit() or abort() (the reference does, through its
 * HANDLE_ERROR macros, e.g. simkernel.cuh:32-39).
 *
 * Data model
 *   - Netlist: levelized, node ids sorted by level; INPT nodes are exactly ids
 *     0..num_pi-1 and form level 0.  Fanins are ordered (fanin 0 is the on-path
 *     input for marking).  This replaces the util/ GPU_Circuit + GPUNODE
 *     contract used at simkernel.cuh:108,129; markkernel.cuh:220-223,277;
 *     coverkernel.cuh:184,204-205.
 *   - Vectors: bit-packed, PI-major.  Row i holds ceil(V/64) uint64 words; bit
 *     j of word k is the value of PI i in vector 64k+j.  Sequence mode grades
 *     pair p = (V[p], V[(p+1) mod V]) as simkernel.cuh:115-124 does; explicit
 *     mode grades caller-given (A_p, B_p).
 *   - Node codes: S0=0, S1=1, T0=2 (1->0), T1=3 (0->1).
 *
 * Threading: a grader handle belongs to one host thread at a time.  All work
 * is issued on the handle's own HIP stream.
 */
#ifndef PDFGRADE_H
#define PDFGRADE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PDF_ABI_VERSION 1

/* status codes */
#define PDF_OK 0
#define PDF_ERR_ARG (-1)         /* invalid argument / shape */
#define PDF_ERR_NETLIST (-2)     /* malformed netlist */
#define PDF_ERR_IO (-3)          /* file could not be read / parsed */
#define PDF_ERR_HIP (-4)         /* HIP runtime error */
#define PDF_ERR_NOMEM (-5)       /* allocation failed */
#define PDF_ERR_UNSUPPORTED (-6) /* feature not available (e.g. segs > 1) */
#define PDF_ERR_STATE (-7)       /* call out of order */
#define PDF_ERR_OVERFLOW (-8)    /* exact uint64 accumulation overflowed */

/* node types (numbering is this library's; the reference's util/defines.h
 * values are not available -- only AND..XNOR = 2..7 is implied by
 * markkernel.cuh:49 indexing LUT_R[type-2]) */
#define PDF_UNKN 0
#define PDF_INPT 1
#define PDF_AND 2
#define PDF_NAND 3
#define PDF_OR 4
#define PDF_NOR 5
#define PDF_XOR 6
#define PDF_XNOR 7
#define PDF_BUFF 8
#define PDF_NOT 9
#define PDF_FROM 10
#define PDF_DFF 11

/* node codes */
#define PDF_S0 0
#define PDF_S1 1
#define PDF_T0 2
#define PDF_T1 3

/* no first-detection pattern yet (the reference's int2 (-1,-1),
 * mergekernel.cuh:215) */
#define PDF_NONE 0xFFFFFFFFu

const char *pdf_last_error(void);
int pdf_abi_version(void);

/* ------------------------------------------------------------------------ */
/* Netlists (host objects; no GPU needed)                                    */
/* ------------------------------------------------------------------------ */
typedef struct pdf_netlist pdf_netlist;

/* Caller-supplied CSR view of a levelized netlist.  Validation rejects: ids
 * not sorted by level, level(g) != 1 + max level(fanin) for gates, INPT not
 * exactly 0..num_pi-1, gates without fanins, fanins that are not earlier
 * nodes, empty levels, more than 65535 fanins or fanouts on one node.
 * Duplicate fanins are allowed (each counts as its own edge). */
typedef struct pdf_netlist_desc {
  uint32_t n;               /* nodes */
  uint32_t num_pi;          /* INPT nodes = ids 0..num_pi-1 */
  const uint8_t *type;      /* [n] PDF_* type */
  const uint8_t *po;        /* [n] primary-output flag (0/1) */
  const uint32_t *fin_off;  /* [n+1] */
  const uint32_t *fin;      /* [fin_off[n]] ordered fanins */
} pdf_netlist_desc;

typedef struct pdf_netlist_info {
  uint32_t n, num_pi, num_po, levels; /* levels = max level index L */
  uint64_t edges;
  uint64_t structural_paths;          /* PI->PO paths, saturated at 2^64-1 */
} pdf_netlist_info;

/* Build from arrays: fanouts and levels are derived (ordered by fanout id). */
int pdf_netlist_create(const pdf_netlist_desc *desc, pdf_netlist **out);
/* ISCAS .bench text (INPUT/OUTPUT/assignments), levelized and renumbered so
 * inputs keep declaration order and gates are sorted by (level, file order).
 * Replaces GPU_Circuit::read_bench (main.cu:129-130; util, absent). */
int pdf_netlist_read_bench(const char *path, pdf_netlist **out);
/* Presorted ".level" text written by pdf_netlist_write_level (format defined
 * by this library; the reference's GPU_Circuit::load format is not
 * available, main.cu:131-133). */
int pdf_netlist_read_level(const char *path, pdf_netlist **out);
int pdf_netlist_write_level(const pdf_netlist *nl, const char *path);
void pdf_netlist_free(pdf_netlist *nl);
int pdf_netlist_get_info(const pdf_netlist *nl, pdf_netlist_info *info);
/* Copy out CSR arrays; any pointer may be NULL.  Sizes: type/po/level [n],
 * fin_off/fout_off [n+1], fin/fout [edges], level_off [levels+2]. */
int pdf_netlist_export(const pdf_netlist *nl, uint8_t *type, uint8_t *po,
                       uint32_t *level, uint32_t *fin_off, uint32_t *fin,
                       uint32_t *fout_off, uint32_t *fout,
                       uint32_t *level_off);

/* Segments of length segs (1..8), the list gpuMergeSegments<K> builds on its
 * first call (mergekernel.cuh:205-221, generateSegmentList in the absent
 * util/ library, so the enumeration is this library's): for every node g0 in
 * id order, every fanout walk g0 -> g1 -> ... of segs nodes, fanouts in id
 * order, depth first; a walk that reaches a node without fanouts first ends
 * there and its remaining entries are PDF_NONE.  segs = 1 is one segment per
 * node.  *count receives the number of segments (saturated at 2^64-1);
 * min(count, capacity) segments of segs entries each are written to out
 * (may be NULL). */
int pdf_netlist_segments(const pdf_netlist *nl, uint32_t segs, uint32_t *out, uint64_t capacity,
                         uint64_t *count);

/* Seeded synthetic netlists (SplitMix64-seeded xoshiro256**).  The
 * reference ships no generator; this one is defined by this library. */
typedef struct pdf_gen_params {
  uint64_t seed;
  uint32_t num_pi, num_po, num_gates, levels;
  double fanin_p[4];        /* P(fan-in = 1..4); fan-in 1 -> NOT/BUFF */
  double prev_level_frac;   /* share of fanins drawn from level-1 */
  double fanout_alpha;      /* power-law exponent of fanout weights; 0 = uniform */
  uint32_t fanout_cap;      /* max fanout of any node (0 = none) */
  uint32_t max_attempts;    /* regenerations allowed when paths >= 2^limit */
  uint32_t path_limit_log2; /* reject when PI->PO paths >= 2^this; 0 = keep any */
} pdf_gen_params;

int pdf_netlist_generate(const pdf_gen_params *params, pdf_netlist **out);

/* iid Bernoulli(1/2) vectors, PI-major packed as described above: writes
 * num_pi * ceil(num_vectors/64) words; bits past num_vectors are zero. */
int pdf_gen_vectors(uint64_t seed, uint32_t num_pi, uint64_t num_vectors,
                    uint64_t *out_bits);
/* Explicit pairs of varied toggle density: A_p uniform (as pdf_gen_vectors
 * with `seed`), B_p = A_p with each PI flipped independently with
 * probability q_p, q_p log-uniform in [q_lo, q_hi] per pair
 * (0 < q_lo <= q_hi <= 1).  Both outputs PI-major, num_pi x ceil(count/64)
 * words, bits past `count` zero. */
int pdf_gen_toggle_pairs(uint64_t seed, uint32_t num_pi, uint64_t count, double q_lo, double q_hi,
                         uint64_t *a_bits, uint64_t *b_bits);

/* Vector text files (the reference's reader, main.cu:146-156, lives in its
 * absent util/ submodule, so the format is this library's): one vector per
 * line, one '0'/'1' per PI in INPUT order; whitespace is ignored and '#'
 * starts a comment.  pdf_vectors_count gives the number of vectors;
 * pdf_vectors_read fills PI-major packed bits (num_pi x ceil(capacity/64)
 * words) with at most `capacity` vectors. */
int pdf_vectors_count(const char *path, uint32_t num_pi, uint64_t *count);
int pdf_vectors_read(const char *path, uint32_t num_pi, uint64_t *bits, uint64_t capacity, uint64_t *count);
int pdf_vectors_write(const char *path, uint32_t num_pi, const uint64_t *bits, uint64_t count);

/* ------------------------------------------------------------------------ */
/* Grader (GPU)                                                              */
/* ------------------------------------------------------------------------ */
typedef struct pdf_grader pdf_grader;

/* Create a grader on HIP device `device` and copy the netlist to HBM
 * (GPU_Circuit::copy, main.cu:137).  The grader keeps its own copy. */
int pdf_grader_create(int device, const pdf_netlist *nl, pdf_grader **out);
void pdf_grader_destroy(pdf_grader *g);

/* Sequence-mode vectors from host memory (read_vectors, main.cu:146-156).
 * Any count >= 1; when it is not a multiple of 64 the bits past the last
 * vector are ignored (the last pair wraps to vector 0, simkernel.cuh:116-120). */
int pdf_grader_set_vectors(pdf_grader *g, const uint64_t *bits,
                           uint64_t num_vectors);
/* Explicit pairs (A_p, B_p), both PI-major packed, num_pairs >= 1. */
int pdf_grader_set_pairs(pdf_grader *g, const uint64_t *a_bits,
                         const uint64_t *b_bits, uint64_t num_pairs);
/* Pairs per chunk (the reference's simul_patterns, main.cu:150-153); a
 * multiple of 64, or 0 (or >= the number of pairs) for one chunk.  Loading
 * vectors picks a default: one chunk if its buffers fit in half of the free
 * HBM, else the largest multiple of 64 that fits.  Allocates the chunk's
 * bit-planes.  Also resets the cross-chunk state. */
int pdf_grader_set_chunk(pdf_grader *g, uint64_t chunk_pairs);
/* Number of chunks for the loaded pairs and the width (pairs) of one. */
int pdf_grader_num_chunks(const pdf_grader *g, uint64_t *num_chunks);
int pdf_grader_chunk_width(const pdf_grader *g, uint64_t chunk,
                           uint64_t *pairs);
/* Forget first-detection state and per-gate counts (new vector set). */
int pdf_grader_reset(pdf_grader *g);

/* The four stages.  `chunk` selects the pairs [start, start+width);
 * start_pattern must equal chunk * chunk_pairs (the reference passes both,
 * main.cu:175,182,187,208).  `ms` (may be NULL) receives wall milliseconds
 * measured with HIP events on the grader's stream.  Each call returns after
 * its work has completed. */
/* gpuRunSimulation (simkernel.cuh:16,200-249) */
int pdf_sim(pdf_grader *g, uint64_t chunk, uint64_t start_pattern, float *ms);
/* gpuMarkPaths (markkernel.cuh:31,295-355); robust replaces the global
 * robust_flag (main.cu:7; markkernel.cuh:313) */
int pdf_mark(pdf_grader *g, uint64_t chunk, uint64_t start_pattern,
             int robust, float *ms);
/* gpuMergeSegments<K> (mergekernel.cuh:200-267), segs = K in 1..8.  For
 * segs >= 2 the first call builds the segment list (pdf_netlist_segments)
 * and keeps it, with each segment's first T0 / T1 pair, until the segment
 * length changes; a segment is detected at a pair when all its nodes are
 * marked, its polarity is node 0's.  Use one segment length for all chunks
 * of a vector set. */
int pdf_merge(pdf_grader *g, uint64_t chunk, uint64_t start_pattern, int segs,
              float *ms);
/* gpuCountPaths<K> (coverkernel.cuh:233-333): *coverage receives (overwrites)
 * this chunk's exact total; per-gate counts accumulate inside the grader.
 * segs must equal the merge's; a node counts as new at a pair when a
 * segment holding it was first detected there (coverkernel.cuh:127-154). */
int pdf_cover(pdf_grader *g, uint64_t chunk, uint64_t start_pattern, int segs,
              uint64_t *coverage, float *ms);

/* The four stages and the reset with the stream made explicit: each queues
 * exactly the launches of its synchronous counterpart above on the caller's
 * HIP stream (a hipStream_t passed as void *) and returns without waiting,
 * allocating, copying synchronously or recording a memset, so a chunk's
 * stages can be captured into the caller's HIP graph.  The same order rules
 * hold (sim, mark, merge, cover per chunk; one segment length per vector
 * set, its list set up once by a synchronous pdf_merge or pdf_grade_segs);
 * debug dumps (pdf_grader_set_debug) need the synchronous calls.
 * pdf_cover_async writes this chunk's exact total to *chunk_total_dev, the
 * per-gate counts accumulated since the last reset to per_gate_accum_dev[n]
 * (if not NULL) and, to *err_dev (if not NULL), non-zero when a uint64 count
 * overflowed, all device memory, once the stream has run it.  The caller
 * lets the stream finish before it reads the getters, changes the chunk or
 * destroys the grader.  (simkernel.cuh:200, markkernel.cuh:295,
 * mergekernel.cuh:200, coverkernel.cuh:233 with the legacy default stream
 * replaced by the caller's.) */
int pdf_grader_reset_async(pdf_grader *g, void *stream);
int pdf_sim_async(pdf_grader *g, uint64_t chunk, uint64_t start_pattern, void *stream);
int pdf_mark_async(pdf_grader *g, uint64_t chunk, uint64_t start_pattern, int robust, void *stream);
int pdf_merge_async(pdf_grader *g, uint64_t chunk, uint64_t start_pattern, int segs, void *stream);
int pdf_cover_async(pdf_grader *g, uint64_t chunk, uint64_t start_pattern, int segs, void *stream,
                    uint64_t *chunk_total_dev, uint64_t *per_gate_accum_dev, uint32_t *err_dev);

/* Whole chunk loop of main.cu:171-231 (sim, mark, merge, cover per chunk,
 * totals accumulated) using the fused kernels; *total receives the exact
 * total.  Equivalent to calling the four stages per chunk. */
int pdf_grade(pdf_grader *g, int robust, uint64_t *total);

/* The fused grade (pdf_grade_segs with segs = 1..8) queued on the caller's
 * HIP stream (a hipStream_t passed as void *) without waiting for it, so
 * that it can be captured into the caller's HIP graph: it allocates
 * nothing, copies nothing synchronously and records no memset.  On
 * completion *total_dev (device memory) holds the exact total and, if
 * err_dev is not NULL, *err_dev is non-zero when a uint64 count overflowed.
 * One lane.  The buffers it uses must exist: pdf_grader_set_chunk makes
 * those of segs = 1, one pdf_grade_segs call those of a segs >= 2.  Calls
 * that overlap on one grader race; the getters see the results once the
 * caller's stream has finished, and the caller lets it finish before it
 * changes the chunk or destroys the grader (the grader does not know the
 * stream) (main.cu:171-231 with the stream made explicit). */
int pdf_grade_async(pdf_grader *g, int robust, int segs, void *stream, uint64_t *total_dev,
                    uint32_t *err_dev);

/* pdf_grade_async (segs = 1) with no chunk buffers of the grader's own.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Every chunk buffer (planes, node lists,
 * counters, cover scratch) is carved out of one static device array of
 * PDF_NOALLOC_SCRATCH_BYTES per GPU, held by the library until the process
 * ends; the call allocates nothing, copies nothing synchronously, records
 * no memset and only queues kernels on `stream`.  The chunk width is the
 * largest multiple of 64 pairs whose buffers fit (results do not depend on
 * it).  A netlist whose buffers do not fit even at 64 pairs per chunk (about
 * 650,000 nodes) is refused with PDF_ERR_UNSUPPORTED.  Results land where
 * pdf_grade_async puts them (*total_dev, *err_dev, and the grader's first /
 * per-gate / total getters); the plane getters do not see this grade.  All
 * calls on one GPU share its array, which keeps the last call's contents:
 * calls that overlap on one GPU race. */
#define PDF_NOALLOC_SCRATCH_BYTES (256ull << 20)
int pdf_grade_noalloc(pdf_grader *g, int robust, void *stream, uint64_t *total_dev, uint32_t *err_dev);

/* The reference's resource limits (results of this library are exact; these
 * say where the reference's own run would have printed something else).
 * pdf_grader_set_merge_window: with pairs > 0 the merge takes first
 * detections only from the first `pairs` pairs of each chunk, as the
 * reference's kernSegmentReduce2 examines only its first 4,224 patterns
 * (mergekernel.cuh:125-130; pass 4224 to reproduce it, 0 = whole chunk, the
 * default).  Applies to the stage calls and to every fused grade.
 * pdf_grader_ref_flags, after a grade (or a chunk's cover): PDF_REF_EXCEEDS_U16
 * when a path count the cover DP computed -- a marked node's fanout sums or
 * its final c / h -- left the reference's 16-bit fields, where its asserts
 * fire (coverkernel.cuh:41-43, 210-211, 224-225; the reference also asserts
 * on unmarked nodes' fanout sums, which this library does not compute);
 * PDF_REF_EXCEEDS_U32 when a chunk's total passed the reference's uint32
 * partial sums (coverkernel.cuh:79-101); PDF_REF_SEGMENT_STRIDE when the
 * reference's merge launches would leave segments unmerged (the stride bug
 * of mergekernel.cuh:243, *unmerged_segments of them, for the segment length
 * of the last merge); PDF_REF_MERGE_WINDOW when chunks are longer than 4,224
 * pairs and no merge window is set.  pdf_ref_unmerged_segments gives that
 * count for S segments. */
#define PDF_REF_EXCEEDS_U16 1u
#define PDF_REF_EXCEEDS_U32 2u
#define PDF_REF_SEGMENT_STRIDE 4u
#define PDF_REF_MERGE_WINDOW 8u
int pdf_grader_set_merge_window(pdf_grader *g, uint32_t pairs);
int pdf_grader_ref_flags(const pdf_grader *g, uint32_t *flags, uint64_t *unmerged_segments);
uint64_t pdf_ref_unmerged_segments(uint64_t segments);

/* Tuning knobs (no reference counterpart; the reference fixes its shapes
 * with compile-time constants such as SIM_BLOCK, MARK_BLOCK, PARALLEL_SEGS,
 * simkernel.cuh:29-30, markkernel.cuh:7-8, mergekernel.cuh:49-52).
 * Results never depend on them. */
#define PDF_TUNE_WORDS_PER_LANE 1 /* 1 or 2 words per lane in the fused forward pass */
#define PDF_TUNE_USE_GRAPH 2      /* 1: pdf_grade / pdf_batch_grade replay a captured HIP graph */
#define PDF_TUNE_COVER_WAVES 3    /* batch: waves of the cover launch */
#define PDF_TUNE_SEGMENT_ORDER 4  /* batch: 0 circuit order, 1 largest segments first */
#define PDF_TUNE_LANES 5          /* grader: chunks graded concurrently on 1..4 streams */
#define PDF_TUNE_REV_COMPACT 6    /* batch: 1 reverse pass over a list of the marked (gate, tile) items
                                   * that a scan launch builds, 2 over lists the forward launches build
                                   * (every circuit >= 4,096 pairs, at most 4,096 circuits; the default
                                   * where it applies, else 1);
                                   * grader (per-level launches): 1 the forward pass appends each level's
                                   * marked (gate, 64-word tile) items to a list and the reverse launches of
                                   * levels >= 1 run over those lists only */
#define PDF_TUNE_REV_TILES 7      /* grader: 64-word tiles per reverse-pass wave, 1/2/4/8 (0: by waves per launch) */
#define PDF_TUNE_WALK 8           /* grader: 0 per-level launches (the reference's host level loops); 1/2/4/8 fused
                                   * grades run each chunk's forward and reverse passes as one launch each, a
                                   * workgroup per tile of that many 64-pair words walking every level */
#define PDF_TUNE_REV_WALK 11      /* grader: 0 per-level reverse launches; 1/2/4/8 after per-level forward launches
                                   * (PDF_TUNE_WALK 0) the reverse pass of a chunk is one level walk with tiles
                                   * of that many words */
#define PDF_TUNE_COVER_NODES 9    /* grader: marked nodes of a pair whose cover DP stays in LDS, 1..1280 (1280);
                                   * pairs with more keep c/h in global scratch */
#define PDF_TUNE_COVER_POOL 10    /* grader: fanout ids of those nodes kept in LDS, 0..4096 (4096); the rest are
                                   * read from the netlist in the DP (both knobs exist to exercise those paths) */
#define PDF_TUNE_FWD_ORDER 12     /* grader: the per-level forward launches take each level's gates 0 (default)
                                   * in id order, 1 grouped by shared fanins (a fanin row is read from HBM once
                                   * where its readers run together: 17% fewer bytes read at 2M gates, but not
                                   * faster there) */
int pdf_grader_set_tuning(pdf_grader *g, int key, int64_t value);

/* Diagnostic variant of pdf_grade (no reference counterpart): the same
 * launches, each bracketed by HIP events on the grader's stream, so the
 * summed device time of every kernel family is reported.  Slower than
 * pdf_grade because of the event records; meant for roofline reporting. */
typedef struct pdf_kernel_times {
  double fwd_ms, reverse_ms, events_ms, cover_ms, fill_ms; /* summed durations */
  uint32_t fwd_launches, reverse_launches, events_launches, cover_launches, fill_launches;
  double wall_ms; /* first to last event */
  double merge_ms;         /* segment merge (segs >= 2) */
  uint32_t merge_launches;
} pdf_kernel_times;
int pdf_grade_timed(pdf_grader *g, int robust, uint64_t *total, pdf_kernel_times *times);

/* pdf_grade with segments of length segs (1..8): per chunk the fused forward
 * and reverse passes, then for segs >= 2 the segment merge and the cover
 * over the nodes of newly detected segments (main.cu:185-223 with
 * gpuMergeSegments<K> / gpuCountPaths<K>); lanes and graph replay as
 * PDF_TUNE_LANES / PDF_TUNE_USE_GRAPH say.  Results equal the four stage calls per chunk
 * with the same segs.  times (may be NULL) receives per-family kernel times
 * as pdf_grade_timed does (slower, serialized launches). */
int pdf_grade_segs(pdf_grader *g, int robust, int segs, uint64_t *total, pdf_kernel_times *times);

/* Parity getters (the reference's --verbose dumps, simkernel.cuh:250-295;
 * markkernel.cuh:357-395; mergekernel.cuh:25-47; coverkernel.cuh:44-66).
 * Planes are node-major: [n][words] uint64, words = chunk width / 64; bit j
 * of word k is pair start+64k+j of the chunk last simulated / marked.  After
 * a fused grade in the level walk (PDF_TUNE_WALK) the sim planes are kept
 * only when a debug flag was set before it (pdf_grader_set_debug); otherwise
 * pdf_get_sim_planes returns PDF_ERR_STATE. */
int pdf_get_sim_planes(const pdf_grader *g, uint64_t *a_plane,
                       uint64_t *b_plane);
int pdf_get_mark_plane(const pdf_grader *g, uint64_t *mark_plane);
/* first_T0 / first_T1: [n], PDF_NONE when never detected. */
int pdf_get_first(const pdf_grader *g, uint32_t *first_t0, uint32_t *first_t1);
/* Stage-API debug copies for the reference's --verbose logs (no effect on
 * results or on pdf_grade): PDF_DEBUG_FORWARD_MARKS makes pdf_mark keep the
 * marks before the reverse removal (gpumark-full.log, markkernel.cuh:323-324);
 * PDF_DEBUG_PI_COVER makes pdf_cover keep c(g, p) of every primary input g
 * for every pair p of the chunk, zero where the DP did not reach
 * (gpucover.log, coverkernel.cuh:44-66, 309-312). */
#define PDF_DEBUG_FORWARD_MARKS 1u
#define PDF_DEBUG_PI_COVER 2u
int pdf_grader_set_debug(pdf_grader *g, uint32_t flags);
/* Forward marks of the chunk last marked, [n][words] as pdf_get_mark_plane. */
int pdf_get_forward_mark_plane(const pdf_grader *g, uint64_t *mark_plane);
/* c of every primary input for every pair of the chunk last covered:
 * [num_pi][width] (exact; the reference prints the low 16 bits). */
int pdf_get_pi_cover(const pdf_grader *g, uint64_t *pi_cover);
/* Each segment's first T0 / T1 pair (PDF_NONE when never detected), in the
 * order of pdf_netlist_segments, for the kept segment list (the length of
 * the last merge or grade with segs >= 2): min(count, capacity) entries
 * each; *count gets the number of segments (0 when no list is kept).
 * Grades with segs >= 2 and pdf_grader_reset start them over; grades with
 * segs = 1 leave them alone. */
int pdf_get_segment_first(const pdf_grader *g, uint32_t *first_t0, uint32_t *first_t1, uint64_t capacity,
                          uint64_t *count);
/* Per-gate exact counts sum_p c(g,p) over all chunks graded so far: [n]. */
int pdf_get_gate_counts(const pdf_grader *g, uint64_t *counts);
/* Exact total over all chunks graded so far. */
int pdf_get_total(const pdf_grader *g, uint64_t *total);

/* ------------------------------------------------------------------------ */
/* Learned pair ranking (no reference counterpart).  Device buffers are      */
/* caller-allocated device pointers on the grader's device.                  */
/* ------------------------------------------------------------------------ */
#define PDF_FEATURES 512

/* Per-pair features of the loaded pairs, BF16 feature-major:
 * feat_dev[f * ld + pair] (ld a multiple of 64, >= num_pairs):
 * [0,128) fraction of the nodes of each level with a transition; [128,256)
 * fraction of the nodes of each level marked by the forward pass (robust
 * or non-robust as `robust` says), both for levels 0..min(max_level, 127);
 * [256,384) fraction of each of 128 PI buckets that toggles
 * (bucket b = PIs [floor(b I / 128), floor((b + 1) I / 128)));
 * [384,512) fanout-weighted toggling share of each PI bucket.  Simulates and
 * marks only levels 0..max_level. */
int pdf_pair_features(pdf_grader *g, uint32_t max_level, int robust, uint16_t *feat_dev, uint64_t ld);
/* Exact per-pair new-path gain against an empty detection history (each
 * pair graded alone): its number of marked PI->PO paths, saturated to u32. */
int pdf_pair_gains(pdf_grader *g, int robust, uint32_t *gain_dev);
/* Reorder the loaded explicit pairs by descending score (stable: equal
 * scores keep their order); perm_dev (may be NULL) receives, for each new
 * position, the pair's old index. */
int pdf_grader_rank_pairs(pdf_grader *g, const float *score_dev, uint32_t *perm_dev);
/* Exact grading of the loaded pairs in order, chunk by chunk: cum_total[c]
 * and cum_ms[c] (may be NULL) hold the running total and the wall time since
 * the start after chunk c; pair_total_dev (device, may be NULL) receives each
 * pair's contribution to the total (saturated to uint32). */
int pdf_grade_curve(pdf_grader *g, int robust, uint64_t *cum_total, float *cum_ms, uint32_t *pair_total_dev);

/* BF16 MFMA scorer for the pair ranking: 512 -> 256 -> 256 -> 1 with ReLU,
 * FP32 accumulation, trained with Adam by the library's own forward and
 * backward kernels.  Parameters (FP32 master copy) are one flat array:
 * W1[256][512] b1[256] W2[256][256] b2[256] w3[256] b3[1].  Every call is
 * synchronous on the scorer's own stream except pdf_mlp_train_step with
 * loss == NULL (finish with pdf_mlp_sync). */
#define PDF_MLP_PARAMS 197377
typedef struct pdf_mlp pdf_mlp;
int pdf_mlp_create(int device, uint64_t seed, pdf_mlp **out);
void pdf_mlp_destroy(pdf_mlp *m);
int pdf_mlp_set_params(pdf_mlp *m, const float *host); /* also clears the Adam state */
int pdf_mlp_get_params(pdf_mlp *m, float *host);
/* Features are BF16 feature-major, as pdf_pair_features writes them:
 * feature f of row r at feat_dev[f * ld + r], ld a multiple of 64 with
 * ld >= rows; feat_dev 16-byte aligned (a column offset that is a multiple
 * of 8 keeps this).  Columns up to the next multiple of 8 past the last row
 * must be readable and finite.
 * score_dev[r] = network output for row r. */
int pdf_mlp_forward(pdf_mlp *m, const uint16_t *feat_dev, uint64_t ld, uint64_t rows, float *score_dev);
/* Mean squared error against target_dev and its gradient (flat layout, to
 * grad_host when not NULL); no update. */
int pdf_mlp_gradients(pdf_mlp *m, const uint16_t *feat_dev, uint64_t ld, const float *target_dev, uint64_t rows,
                      float *loss, float *grad_host);
/* One Adam step with the last computed gradient. */
int pdf_mlp_adam(pdf_mlp *m, float lr, float beta1, float beta2, float eps);
/* Gradient + Adam (betas 0.9 / 0.999, eps 1e-8) on one batch. */
int pdf_mlp_train_step(pdf_mlp *m, const uint16_t *feat_dev, uint64_t ld, const float *target_dev, uint64_t rows,
                       float lr, float *loss);
int pdf_mlp_sync(pdf_mlp *m);

/* pdf_pair_features followed by pdf_mlp_forward on every pair, pipelined (no
 * reference counterpart): each group of `lanes` chunks is scored on the
 * scorer's stream as soon as its features are written, while the next
 * group's features are computed.  feat_dev receives the features as
 * pdf_pair_features writes them and score_dev [num_pairs] the scores; both
 * equal the two separate calls.  Grader and scorer on the same device. */
int pdf_score_pairs(pdf_grader *g, pdf_mlp *m, uint32_t max_level, int robust, uint16_t *feat_dev, uint64_t ld,
                    float *score_dev);

/* Autotuner (no reference counterpart): times the grade (reps runs each,
 * median) over chunk lanes 1..4, each with chunk = pairs / lanes and the
 * power-of-two chunks from 32,768 pairs below it (only configurations whose
 * chunk buffers fit in 85% of the free HBM), graph replay on and off, words
 * per lane, and per-level launches against the level walk with 2- and 4-word
 * tiles (PDF_TUNE_WALK); checks that every configuration gives the same
 * total and leaves the grader in the fastest one. */
typedef struct {
  uint32_t lanes;
  uint64_t chunk_pairs;
  uint32_t use_graph, words_per_lane;
  uint32_t walk; /* PDF_TUNE_WALK: 0 = per-level launches */
  float ms;      /* median grade time of this configuration */
} pdf_grader_tuning;
int pdf_grader_autotune(pdf_grader *g, int robust, uint32_t reps, pdf_grader_tuning *best);

/* ------------------------------------------------------------------------ */
/* Batched grading (no reference counterpart): many circuits, each with its  */
/* own sequence-mode vectors graded as one chunk, in one captured graph of   */
/* level launches over the union of the circuits' levels.  Results equal     */
/* grading each circuit alone with pdf_grade.                                */
/* ------------------------------------------------------------------------ */
typedef struct pdf_batch pdf_batch;
typedef struct {
  uint32_t segment_order, cover_waves, use_graph, rev_compact, words_per_lane;
  float ms; /* median grade time of this configuration */
} pdf_batch_tuning;

/* vec_bits[c]: PI-major packed vectors of circuit c (as pdf_grader_set_vectors),
 * num_vectors[c] of them. */
int pdf_batch_create(int device, uint32_t count, const pdf_netlist *const *nets, const uint64_t *const *vec_bits,
                     const uint64_t *num_vectors, pdf_batch **out);
void pdf_batch_destroy(pdf_batch *b);
/* Grade every circuit; totals[c] (may be NULL) = circuit c's path total;
 * ms (may be NULL) = GPU time of the grade. */
int pdf_batch_grade(pdf_batch *b, int robust, uint64_t *totals, float *ms);
/* Circuit c's grader, for the getters (pdf_get_first, pdf_get_gate_counts,
 * pdf_get_total) after pdf_batch_grade; owned by the batch. */
int pdf_batch_member(pdf_batch *b, uint32_t circuit, pdf_grader **out);
/* PDF_TUNE_USE_GRAPH, PDF_TUNE_COVER_WAVES, PDF_TUNE_SEGMENT_ORDER, PDF_TUNE_REV_COMPACT and
 * PDF_TUNE_WORDS_PER_LANE (2 takes effect only when every circuit has at least 4,096 pairs).
 * When the lists of PDF_TUNE_REV_COMPACT = 2 do not fit in device memory (after this call or
 * a call that changes the segments), the batch keeps running with value 1 and the call
 * returns PDF_ERR_NOMEM. */
int pdf_batch_set_tuning(pdf_batch *b, int key, int64_t value);
/* Time every combination of the knobs' values (segment order, cover waves,
 * graph replay, reverse pass, words per lane: 96 at most; each
 * timed over reps runs, median), check that every configuration gives the
 * same totals, keep the fastest and report it in *best. */
int pdf_batch_autotune(pdf_batch *b, int robust, uint32_t reps, pdf_batch_tuning *best);

#ifdef __cplusplus
}
#endif

#endif /* PDFGRADE_H */
