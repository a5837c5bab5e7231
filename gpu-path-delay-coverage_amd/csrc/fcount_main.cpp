// fcount: command-line grader with the reference's flags and result line
// (main.cu:32-237), driving the C ABI of libpdfgrade.
//
//   fcount [--robust] [--segs N] [--num_patterns N] [--verbose [--sim]
//          [--mark] [--merge] [--cover]] [--fused] --bench CIRCUIT VECTORS...
//
// CIRCUIT: a .bench file when the name contains "bench", otherwise a
// presorted .level file (main.cu:129-134).  VECTORS: text files with one
// vector per line as '0'/'1' characters (one per PI, in INPUT order);
// whitespace is ignored and '#' starts a comment (pdf_vectors_read).
//
// For every vector file one line goes to stdout, as main.cu:233-234 prints it:
//   file:V,n,K;gpu;sim;mark;merge;cover;total
// with gpu = summed stage milliseconds and sim..cover = the last chunk's
// (stage times are HIP-event wall time here, not process CPU time).

#include <getopt.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "pdfgrade.h"

namespace {

int verbose_flag = 0, sim_flag = 0, mark_flag = 0, merge_flag = 0, cover_flag = 0, robust_flag = 0,
    fused_flag = 0;

[[noreturn]] void die(const std::string &what) {
  std::cerr << "fcount: " << what << "\n";
  std::exit(EXIT_FAILURE);
}

void check(int rc, const char *what) {
  if (rc != PDF_OK) die(std::string(what) + ": " + pdf_last_error());
}

// Reads a vector file (pdf_vectors_read's format); returns PI-major packed bits.
std::vector<uint64_t> read_vectors(const std::string &path, uint32_t num_pi, uint64_t *count) {
  uint64_t V = 0;
  check(pdf_vectors_count(path.c_str(), num_pi, &V), "pdf_vectors_count");
  if (V == 0) die("no vectors in " + path);
  std::vector<uint64_t> packed((size_t)num_pi * ((V + 63) / 64), 0);
  check(pdf_vectors_read(path.c_str(), num_pi, packed.data(), V, count), "pdf_vectors_read");
  return packed;
}

const char *code_name(int c) {
  static const char *n[] = {"S0", "S1", "T0", "T1"};
  return n[c & 3];
}

// --verbose logs in the reference's formats (column width OUTJUST = 4,
// main.cu:26): a "Gate:" header row on the first chunk, then one row per
// pair, appended chunk by chunk.
//   gpusim-p2.log    (--sim)  node codes S0/S1/T0/T1, simkernel.cuh:250-295
//   gpumark-full.log (--mark) marks before the reverse removal, and
//   gpumark.log      (--mark) after it, Y/N, markkernel.cuh:323-349, 357-395
//   gpucover.log              c of every PI, coverkernel.cuh:44-66, 309-312
//                             (exact; the reference prints the low 16 bits)
// gpumerge.log (--merge) is written after the chunk loop (the reference
// rewrites it at every merge, so its final content is the same).
constexpr int kOutJust = 4;

std::ofstream open_log(const char *name, bool first_chunk) {
  return std::ofstream(name, first_chunk ? std::ios::out : std::ios::app);
}

void gate_header(std::ofstream &f, const char *label, uint32_t count) {
  f << label << "\t";
  for (uint32_t i = 0; i < count; ++i) f << std::setw(kOutJust) << i << " ";
  f << "\n";
}

void dump_marks(const char *name, const std::vector<uint64_t> &m, uint32_t n, uint64_t width, uint64_t start,
                bool first_chunk) {
  const uint64_t words = (width + 63) / 64;
  std::ofstream f = open_log(name, first_chunk);
  if (first_chunk) gate_header(f, "Gate:     ", n);
  for (uint64_t p = 0; p < width; ++p) {
    f << "Vector " << start + p << ":\t";
    for (uint32_t i = 0; i < n; ++i)
      f << std::setw(kOutJust) << (((m[i * words + p / 64] >> (p % 64)) & 1) ? "Y" : "N") << " ";
    f << "\n";
  }
}

void dump_planes(pdf_grader *g, uint32_t n, uint64_t width, uint64_t start, bool first_chunk) {
  const uint64_t words = (width + 63) / 64;
  if (sim_flag) {
    std::vector<uint64_t> a(n * words), b(n * words);
    check(pdf_get_sim_planes(g, a.data(), b.data()), "pdf_get_sim_planes");
    std::ofstream f = open_log("gpusim-p2.log", first_chunk);
    if (first_chunk) gate_header(f, "Gate:     ", n);
    for (uint64_t p = 0; p < width; ++p) {
      f << "Vector " << start + p << ":\t";
      for (uint32_t i = 0; i < n; ++i) {
        const int av = (a[i * words + p / 64] >> (p % 64)) & 1, bv = (b[i * words + p / 64] >> (p % 64)) & 1;
        f << std::setw(kOutJust + 1) << (std::string(code_name(bv | ((av ^ bv) << 1))) + " ");
      }
      f << "\n";
    }
  }
  if (mark_flag) {
    std::vector<uint64_t> m(n * words);
    check(pdf_get_forward_mark_plane(g, m.data()), "pdf_get_forward_mark_plane");
    dump_marks("gpumark-full.log", m, n, width, start, first_chunk);
    check(pdf_get_mark_plane(g, m.data()), "pdf_get_mark_plane");
    dump_marks("gpumark.log", m, n, width, start, first_chunk);
  }
}

void dump_cover(pdf_grader *g, uint32_t num_pi, uint64_t width, uint64_t start, bool first_chunk) {
  std::vector<uint64_t> c((uint64_t)num_pi * width);
  check(pdf_get_pi_cover(g, c.data()), "pdf_get_pi_cover");
  std::ofstream f = open_log("gpucover.log", first_chunk);
  if (first_chunk) gate_header(f, "Gate:   ", num_pi);
  for (uint64_t p = 0; p < width; ++p) {
    f << "Vector " << start + p << ":\t";
    for (uint32_t i = 0; i < num_pi; ++i) f << std::setw(kOutJust) << c[(uint64_t)i * width + p] << " ";
    f << "\n";
  }
}

}  // namespace

int main(int argc, char **argv) {
  std::string infile;
  std::vector<std::string> benchmarks;
  long override_patterns = 0;
  long merge_window = 0;
  int segs = 1;
  static struct option long_options[] = {{"verbose", no_argument, &verbose_flag, 1},
                                         {"cover", no_argument, &cover_flag, 1},
                                         {"mark", no_argument, &mark_flag, 1},
                                         {"merge", no_argument, &merge_flag, 1},
                                         {"sim", no_argument, &sim_flag, 1},
                                         {"robust", no_argument, &robust_flag, 1},
                                         {"brief", no_argument, &verbose_flag, 0},
                                         {"fused", no_argument, &fused_flag, 1},
                                         {"help", no_argument, 0, 'h'},
                                         {"bench", required_argument, 0, 'b'},
                                         {"num_patterns", required_argument, 0, 'o'},
                                         {"segs", required_argument, 0, 's'},
                                         {"merge_window", required_argument, 0, 'w'},
                                         {"ref_window", no_argument, 0, 'r'},
                                         {0, 0, 0, 0}};
  while (true) {
    int idx = 0;
    const int c = getopt_long(argc, argv, "b:s:", long_options, &idx);
    if (c == -1) break;
    switch (c) {
      case 0: break;
      case 'b': infile = optarg; break;
      case 'o': override_patterns = std::atol(optarg); break;
      case 'w': merge_window = std::atol(optarg); break;
      case 'r': merge_window = 4224; break;
      case 's':
        segs = std::atoi(optarg);
        if (segs > 8) die("Segment lengths over 8 are not supported!");
        if (segs < 1) die("Segment length must be at least 1");
        break;
      case 'h':
        std::printf("Usage: %s (options) /path/to/vectors1 /path/to/vectors2\n", argv[0]);
        std::printf("\t--segs N : Run fault grading over segments of length N, where 1 <= N <= 8\n");
        std::printf("\t--bench /path/to/ckt : the circuit to grade (.bench or .level)\n");
        std::printf("\t--robust : robust sensitization (default non-robust)\n");
        std::printf("\t--num_patterns N : pairs per chunk = N-1, rounded down to a multiple of 64\n");
        std::printf("\t--fused : whole chunk loop in one call (pdf_grade)\n");
        std::printf("\t--ref_window : merge only the first 4224 pairs of each chunk, as the reference does\n");
        std::printf("\t--merge_window N : merge only the first N pairs of each chunk (0 = all)\n");
        return 0;
      default: return EXIT_FAILURE;
    }
  }
  for (int i = optind; i < argc; ++i) benchmarks.push_back(argv[i]);
  if (infile.empty()) die("--bench argument is required.");
  if (fused_flag && segs != 1) die("--fused grades segments of length 1 only");

  pdf_netlist *nl = nullptr;
  if (infile.find("bench") != std::string::npos) check(pdf_netlist_read_bench(infile.c_str(), &nl), "read_bench");
  else check(pdf_netlist_read_level(infile.c_str(), &nl), "read_level");
  pdf_netlist_info info;
  check(pdf_netlist_get_info(nl, &info), "get_info");
  std::clog << "Circuit size is: " << info.n << " Levels: " << info.levels << std::endl;
  pdf_grader *g = nullptr;
  check(pdf_grader_create(0, nl, &g), "pdf_grader_create");
  if (merge_window < 0 || merge_window > 0xFFFFFFFFL) die("--merge_window must be 0..4294967295");
  check(pdf_grader_set_merge_window(g, (uint32_t)merge_window), "set_merge_window");
  if (verbose_flag && !fused_flag)
    check(pdf_grader_set_debug(g, (mark_flag ? PDF_DEBUG_FORWARD_MARKS : 0u) | PDF_DEBUG_PI_COVER), "set_debug");

  for (const std::string &vf : benchmarks) {
    uint64_t V = 0;
    std::vector<uint64_t> bits = read_vectors(vf, info.num_pi, &V);
    check(pdf_grader_set_vectors(g, bits.data(), V), "set_vectors");
    // pairs per chunk: --num_patterns N gives N-1 (main.cu:151-153), rounded
    // to whole 64-pair words; otherwise as many as fit in half of free HBM
    uint64_t chunk = 0;
    if (override_patterns > 0) {
      chunk = (uint64_t)(override_patterns - 1) / 64 * 64;
      if (chunk == 0) chunk = 64;
    } else {
      size_t free_b = 0, total_b = 0;
      if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
        const double per_pair = info.n * (3.0 / 8.0 + 4.0 / 64.0) + 16.0;
        chunk = (uint64_t)(free_b * 0.5 / per_pair) / 64 * 64;
      }
      if (chunk == 0 || chunk >= V) chunk = 0;
    }
    check(pdf_grader_set_chunk(g, chunk), "set_chunk");
    uint64_t nchunks = 0;
    check(pdf_grader_num_chunks(g, &nchunks), "num_chunks");
    std::clog << "Maximum patterns per pass: " << (chunk ? chunk : V) << " / " << V << std::endl;
    float sim = 0, mark = 0, merge = 0, cover = 0, gpu = 0;
    uint64_t totals = 0;
    if (fused_flag) {
      hipEvent_t e0, e1;
      (void)hipEventCreate(&e0);
      (void)hipEventCreate(&e1);
      (void)hipEventRecord(e0, 0);
      check(pdf_grade(g, robust_flag, &totals), "pdf_grade");
      (void)hipEventRecord(e1, 0);
      (void)hipEventSynchronize(e1);
      (void)hipEventElapsedTime(&gpu, e0, e1);
    } else {
      uint64_t start = 0;
      for (uint64_t c = 0; c < nchunks; ++c) {
        uint64_t width = 0, coverage = 0;
        check(pdf_grader_chunk_width(g, c, &width), "chunk_width");
        check(pdf_sim(g, c, start, &sim), "pdf_sim");
        check(pdf_mark(g, c, start, robust_flag, &mark), "pdf_mark");
        if (verbose_flag) dump_planes(g, info.n, width, start, c == 0);
        check(pdf_merge(g, c, start, segs, &merge), "pdf_merge");
        check(pdf_cover(g, c, start, segs, &coverage, &cover), "pdf_cover");
        if (verbose_flag) dump_cover(g, info.num_pi, width, start, c == 0);
        gpu += sim + mark + merge + cover;
        totals += coverage;
        std::cerr << "GPU Coverage: " << coverage << ", total: " << totals << std::endl;
        start += width;
      }
    }
    if (verbose_flag && (merge_flag || cover_flag)) {
      std::vector<uint32_t> f0(info.n), f1(info.n);
      std::vector<uint64_t> counts(info.n);
      check(pdf_get_first(g, f0.data(), f1.data()), "get_first");
      check(pdf_get_gate_counts(g, counts.data()), "get_gate_counts");
      if (merge_flag && segs == 1) {
        std::ofstream f("gpumerge.log");
        for (uint32_t i = 0; i < info.n; ++i)
          f << "Segment " << i << "(" << i << "): " << (int32_t)f0[i] << "," << (int32_t)f1[i] << " \n";
      } else if (merge_flag) {
        // segment r (its node ids, -1 padded): first T0, first T1
        uint64_t S = 0;
        check(pdf_netlist_segments(nl, (uint32_t)segs, nullptr, 0, &S), "segments");
        std::vector<uint32_t> ids(S * segs), s0(S), s1(S);
        check(pdf_netlist_segments(nl, (uint32_t)segs, ids.data(), S, &S), "segments");
        check(pdf_get_segment_first(g, s0.data(), s1.data(), S, &S), "get_segment_first");
        std::ofstream f("gpumerge.log");
        for (uint64_t r = 0; r < S; ++r) {
          f << "Segment " << r << "(";
          for (int j = 0; j < segs; ++j) f << (int32_t)ids[r * segs + j] << (j + 1 < segs ? "," : "");
          f << "): " << (int32_t)s0[r] << "," << (int32_t)s1[r] << " \n";
        }
      }
      if (cover_flag) {  // per-node totals over all pairs (no reference counterpart)
        std::ofstream f("gpucover-gates.log");
        for (uint32_t i = 0; i < info.n; ++i) f << "Gate " << i << ": " << counts[i] << "\n";
      }
    }
    // where the reference's own run would differ (its resource limits)
    uint32_t ref = 0;
    uint64_t unmerged = 0;
    check(pdf_grader_ref_flags(g, &ref, &unmerged), "ref_flags");
    if (ref & PDF_REF_EXCEEDS_U16) std::cerr << "note: a path count exceeds the reference's 16-bit fields\n";
    if (ref & PDF_REF_EXCEEDS_U32) std::cerr << "note: a chunk total exceeds the reference's uint32 sums\n";
    if (ref & PDF_REF_SEGMENT_STRIDE)
      std::cerr << "note: the reference's merge would skip " << unmerged << " segments\n";
    if (ref & PDF_REF_MERGE_WINDOW)
      std::cerr << "note: chunks exceed the reference's 4224-pattern merge window (see --ref_window)\n";
    std::cerr << "   GPU: " << gpu << " ms" << std::endl;
    std::cout << vf << ":" << V << "," << info.n << "," << segs << ";" << gpu << ";" << sim << ";" << mark << ";"
              << merge << ";" << cover << ";" << totals << std::endl;
  }
  pdf_grader_destroy(g);
  pdf_netlist_free(nl);
  return 0;
}
