// Host-side netlist handling: validation/levelization, the seeded synthetic
// generator, the .bench and .level readers, and packed vector generation.
//
// These replace the util/ contract the reference builds on (GPU_Circuit,
// read_bench, load, levelsize; main.cu:129-138, simkernel.cuh:213-214) whose
// source is not available.  Nodes are kept sorted by level with INPT nodes
// first, which the reference's kernels assume (simkernel.cuh:115 reads PI row
// = node id; coverkernel.cuh:299 sums the first levelsize(0) ids).

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "pdf_internal.hpp"

namespace pdf {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

static uint64_t sat_add(uint64_t a, uint64_t b) {
  uint64_t s = a + b;
  return s < a ? UINT64_MAX : s;
}

int finalize_netlist(Netlist &nl, std::string &err) {
  const uint32_t n = nl.n;
  if (n == 0) { err = "netlist has no nodes"; return PDF_ERR_NETLIST; }
  if (nl.num_pi == 0 || nl.num_pi > n) { err = "num_pi must be in [1, n]"; return PDF_ERR_NETLIST; }
  if (nl.type.size() != n || nl.po.size() != n || nl.fin_off.size() != (size_t)n + 1) {
    err = "array sizes disagree with n"; return PDF_ERR_NETLIST;
  }
  if (nl.fin_off[0] != 0 || nl.fin.size() != nl.fin_off[n]) {
    err = "fin_off does not describe fin"; return PDF_ERR_NETLIST;
  }
  nl.level.assign(n, 0);
  nl.num_po = 0;
  for (uint32_t g = 0; g < n; ++g) {
    if (nl.fin_off[g + 1] < nl.fin_off[g]) { err = "fin_off not monotone"; return PDF_ERR_NETLIST; }
    const bool is_pi = g < nl.num_pi;
    if (is_pi != (nl.type[g] == PDF_INPT)) {
      err = "node " + std::to_string(g) + ": INPT nodes must be exactly ids 0..num_pi-1";
      return PDF_ERR_NETLIST;
    }
    if (nl.type[g] > PDF_DFF) { err = "node " + std::to_string(g) + ": unknown type"; return PDF_ERR_NETLIST; }
    if (nl.po[g] > 1) { err = "po flags must be 0/1"; return PDF_ERR_NETLIST; }
    nl.num_po += nl.po[g];
    const uint32_t nfi = nl.fin_off[g + 1] - nl.fin_off[g];
    if (is_pi) {
      if (nfi != 0) { err = "INPT node " + std::to_string(g) + " has fanins"; return PDF_ERR_NETLIST; }
      continue;
    }
    if (nfi == 0) { err = "gate " + std::to_string(g) + " has no fanins"; return PDF_ERR_NETLIST; }
    if (nfi > 65535) { err = "gate " + std::to_string(g) + " has more than 65535 fanins"; return PDF_ERR_NETLIST; }
    uint32_t lv = 0;
    for (uint32_t k = nl.fin_off[g]; k < nl.fin_off[g + 1]; ++k) {
      const uint32_t f = nl.fin[k];
      if (f >= g) {
        err = "gate " + std::to_string(g) + ": fanin " + std::to_string(f) +
              " is not an earlier node (ids must be sorted by level)";
        return PDF_ERR_NETLIST;
      }
      lv = std::max(lv, nl.level[f] + 1);
    }
    nl.level[g] = lv;
    if (lv < nl.level[g - 1]) {
      err = "node " + std::to_string(g) + ": ids are not sorted by level";
      return PDF_ERR_NETLIST;
    }
  }
  nl.L = nl.level[n - 1];
  nl.level_off.assign(nl.L + 2, 0);
  for (uint32_t g = 0; g < n; ++g) nl.level_off[nl.level[g] + 1]++;
  for (uint32_t k = 0; k <= nl.L; ++k) {
    if (nl.level_off[k + 1] == 0) { err = "level " + std::to_string(k) + " is empty"; return PDF_ERR_NETLIST; }
    nl.level_off[k + 1] += nl.level_off[k];
  }
  // fanouts, sorted by fanout id (CSR transpose in id order)
  nl.fout_off.assign(n + 1, 0);
  for (uint32_t f : nl.fin) nl.fout_off[f + 1]++;
  for (uint32_t g = 0; g < n; ++g)
    if (nl.fout_off[g + 1] > 65535) {
      err = "node " + std::to_string(g) + " has more than 65535 fanouts";
      return PDF_ERR_NETLIST;
    }
  for (uint32_t g = 0; g < n; ++g) nl.fout_off[g + 1] += nl.fout_off[g];
  nl.fout.assign(nl.fin.size(), 0);
  std::vector<uint32_t> fill(nl.fout_off.begin(), nl.fout_off.end() - 1);
  for (uint32_t g = 0; g < n; ++g)
    for (uint32_t k = nl.fin_off[g]; k < nl.fin_off[g + 1]; ++k) nl.fout[fill[nl.fin[k]]++] = g;
  // structural PI->PO paths (saturating)
  std::vector<uint64_t> paths(n, 0);
  uint64_t total = 0;
  for (uint32_t g = 0; g < n; ++g) {
    uint64_t p = g < nl.num_pi ? 1 : 0;
    for (uint32_t k = nl.fin_off[g]; k < nl.fin_off[g + 1]; ++k) p = sat_add(p, paths[nl.fin[k]]);
    paths[g] = p;
    if (nl.po[g]) total = sat_add(total, p);
  }
  nl.structural_paths = total;
  return PDF_OK;
}

// ---------------------------------------------------------------- generator

static const uint8_t kMultiTypes[6] = {PDF_AND, PDF_NAND, PDF_OR, PDF_NOR, PDF_XOR, PDF_XNOR};

struct LevelSampler {
  std::vector<uint64_t> cum;  // prefix sums of weights over the level's ids
  uint32_t first = 0;
  uint32_t pick(Rng &rng) const {
    const uint64_t r = rng.below(cum.back());
    const size_t i = std::upper_bound(cum.begin(), cum.end(), r) - cum.begin();
    return first + (uint32_t)i;
  }
};

static int generate_once(const pdf_gen_params &p, Rng &rng, Netlist &nl, std::string &err) {
  const uint32_t I = p.num_pi, G = p.num_gates, L = p.levels;
  const uint32_t n = I + G;
  // level sizes: as even as possible, the remainder on the lowest levels
  std::vector<uint32_t> lvl_first(L + 2, 0);
  lvl_first[0] = 0;
  lvl_first[1] = I;
  for (uint32_t k = 1; k <= L; ++k) {
    const uint32_t sz = G / L + ((k - 1) < (G % L) ? 1u : 0u);
    lvl_first[k + 1] = lvl_first[k] + sz;
  }
  // fanout weights (power law on 1..cap) drawn per node
  std::vector<uint32_t> weight(n, 1);
  if (p.fanout_alpha > 0) {
    const uint32_t cap = p.fanout_cap ? p.fanout_cap : 64;
    std::vector<double> cdf(cap);
    double acc = 0;
    for (uint32_t k = 1; k <= cap; ++k) { acc += std::pow((double)k, -p.fanout_alpha); cdf[k - 1] = acc; }
    for (auto &c : cdf) c /= acc;
    for (uint32_t g = 0; g < n; ++g) {
      const double u = rng.uniform();
      uint32_t k = (uint32_t)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
      weight[g] = std::min(k, cap - 1) + 1;
    }
  }
  std::vector<LevelSampler> samp(L + 1);
  for (uint32_t k = 0; k <= L; ++k) {
    samp[k].first = lvl_first[k];
    uint64_t acc = 0;
    for (uint32_t g = lvl_first[k]; g < lvl_first[k + 1]; ++g) { acc += weight[g]; samp[k].cum.push_back(acc); }
  }
  double fcdf[4];
  {
    double acc = 0, tot = p.fanin_p[0] + p.fanin_p[1] + p.fanin_p[2] + p.fanin_p[3];
    for (int i = 0; i < 4; ++i) { acc += p.fanin_p[i] / tot; fcdf[i] = acc; }
  }
  std::vector<uint32_t> fo_count(n, 0);
  const uint32_t cap = p.fanout_cap;
  nl = Netlist();
  nl.n = n;
  nl.num_pi = I;
  nl.type.assign(n, PDF_INPT);
  nl.po.assign(n, 0);
  nl.fin_off.assign(n + 1, 0);
  nl.fin.reserve((size_t)G * 3);
  std::vector<uint32_t> chosen;
  for (uint32_t k = 1; k <= L; ++k) {
    for (uint32_t g = lvl_first[k]; g < lvl_first[k + 1]; ++g) {
      const double u = rng.uniform();
      uint32_t f = 1;
      while (f < 4 && u >= fcdf[f - 1]) ++f;
      const uint8_t ty = f == 1 ? (rng.below(2) ? PDF_NOT : PDF_BUFF) : kMultiTypes[rng.below(6)];
      chosen.clear();
      auto ok = [&](uint32_t c) {
        if (cap && fo_count[c] >= cap) return false;
        return std::find(chosen.begin(), chosen.end(), c) == chosen.end();
      };
      for (uint32_t pos = 0; pos < f; ++pos) {
        for (int tries = 0; tries < 64; ++tries) {
          const uint32_t src = (k == 1 || rng.uniform() < p.prev_level_frac) ? k - 1 : (uint32_t)rng.below(k - 1);
          const uint32_t c = samp[src].pick(rng);
          if (ok(c)) { chosen.push_back(c); break; }
        }
      }
      bool has_prev = false;
      for (uint32_t c : chosen) has_prev |= (c >= lvl_first[k - 1] && c < lvl_first[k]);
      if (!has_prev) {
        bool placed = false;
        for (int tries = 0; tries < 256 && !placed; ++tries) {
          const uint32_t c = samp[k - 1].pick(rng);
          if (!ok(c)) continue;
          if (chosen.empty()) chosen.push_back(c);
          else chosen[rng.below(chosen.size())] = c;
          placed = true;
        }
        if (!placed) { err = "generator could not connect a gate to the previous level"; return PDF_ERR_NETLIST; }
      }
      nl.type[g] = ty;
      for (uint32_t c : chosen) { nl.fin.push_back(c); fo_count[c]++; }
      nl.fin_off[g + 1] = (uint32_t)nl.fin.size();
    }
  }
  // primary outputs: drawn among the gates that drive nothing, then (if too
  // few) among the remaining gates
  std::vector<uint32_t> dangling, others;
  for (uint32_t g = I; g < n; ++g) (fo_count[g] == 0 ? dangling : others).push_back(g);
  uint32_t need = std::min<uint32_t>(p.num_po, G);
  auto take = [&](std::vector<uint32_t> &pool) {
    for (size_t i = 0; i < pool.size() && need > 0; ++i, --need) {
      const size_t j = i + rng.below(pool.size() - i);
      std::swap(pool[i], pool[j]);
      nl.po[pool[i]] = 1;
    }
  };
  take(dangling);
  take(others);
  return finalize_netlist(nl, err);
}

// Walks counted level by level: cnt_k(g) = 1 when k = 1 or g has no fanouts,
// else the sum over its fanouts of cnt_{k-1}.
uint64_t count_segments(const Netlist &nl, uint32_t K) {
  std::vector<uint64_t> cnt(nl.n, 1), nxt(nl.n);
  for (uint32_t k = 1; k < K; ++k) {
    for (uint32_t g = 0; g < nl.n; ++g) {
      const uint32_t lo = nl.fout_off[g], hi = nl.fout_off[g + 1];
      uint64_t s = lo == hi ? 1 : 0;
      for (uint32_t e = lo; e < hi; ++e)
        if (__builtin_add_overflow(s, cnt[nl.fout[e]], &s)) s = UINT64_MAX;
      nxt[g] = s;
    }
    cnt.swap(nxt);
  }
  uint64_t total = 0;
  for (uint64_t c : cnt)
    if (__builtin_add_overflow(total, c, &total)) return UINT64_MAX;
  return total;
}

void list_segments(const Netlist &nl, uint32_t K, uint32_t *out, uint64_t cap) {
  // explicit stack: path[d] and the next fanout position to try below it
  uint32_t path[8], next[8];
  uint64_t row = 0;
  for (uint32_t g0 = 0; g0 < nl.n && row < cap; ++g0) {
    uint32_t d = 1;
    path[0] = g0;
    next[0] = nl.fout_off[g0];
    bool emit = true;  // emit when a walk has just been extended
    while (d > 0 && row < cap) {
      const uint32_t g = path[d - 1];
      const uint32_t lo = nl.fout_off[g], hi = nl.fout_off[g + 1];
      if (emit && (d == K || lo == hi)) {
        for (uint32_t i = 0; i < K; ++i) out[row * K + i] = i < d ? path[i] : PDF_NONE;
        ++row;
        --d;  // back to the parent
        emit = false;
        continue;
      }
      if (emit) next[d - 1] = lo;
      if (next[d - 1] < hi) {
        path[d] = nl.fout[next[d - 1]++];
        ++d;
        emit = true;
      } else {
        --d;
        emit = false;
      }
    }
  }
}

}  // namespace pdf

using namespace pdf;

extern "C" {

const char *pdf_last_error(void) { return g_last_error.c_str(); }
int pdf_abi_version(void) { return PDF_ABI_VERSION; }

int pdf_netlist_create(const pdf_netlist_desc *d, pdf_netlist **out) {
  if (!d || !out || !d->type || !d->po || !d->fin_off || (!d->fin && d->n && d->fin_off[d->n])) {
    set_error("pdf_netlist_create: null argument");
    return PDF_ERR_ARG;
  }
  auto *h = new (std::nothrow) pdf_netlist;
  if (!h) { set_error("out of memory"); return PDF_ERR_NOMEM; }
  Netlist &nl = h->nl;
  nl.n = d->n;
  nl.num_pi = d->num_pi;
  nl.type.assign(d->type, d->type + d->n);
  nl.po.assign(d->po, d->po + d->n);
  nl.fin_off.assign(d->fin_off, d->fin_off + d->n + 1);
  nl.fin.assign(d->fin, d->fin + d->fin_off[d->n]);
  std::string err;
  int rc = finalize_netlist(nl, err);
  if (rc != PDF_OK) { delete h; set_error("pdf_netlist_create: " + err); return rc; }
  *out = h;
  return PDF_OK;
}

int pdf_netlist_generate(const pdf_gen_params *p, pdf_netlist **out) {
  if (!p || !out) { set_error("pdf_netlist_generate: null argument"); return PDF_ERR_ARG; }
  if (p->num_pi == 0 || p->levels == 0 || p->num_gates < p->levels) {
    set_error("pdf_netlist_generate: need num_pi >= 1, levels >= 1, num_gates >= levels");
    return PDF_ERR_ARG;
  }
  if (p->fanout_cap == 1) { set_error("pdf_netlist_generate: fanout_cap must be 0 or >= 2"); return PDF_ERR_ARG; }
  double tot = 0;
  for (double v : p->fanin_p) { if (v < 0) tot = -1; else tot += v; }
  if (!(tot > 0)) { set_error("pdf_netlist_generate: fanin_p must be non-negative with a positive sum"); return PDF_ERR_ARG; }
  auto *h = new (std::nothrow) pdf_netlist;
  if (!h) { set_error("out of memory"); return PDF_ERR_NOMEM; }
  const uint32_t attempts = p->max_attempts ? p->max_attempts : 1;
  std::string err;
  for (uint32_t a = 0; a < attempts; ++a) {
    Rng rng(p->seed + 0xD1B54A32D192ED03ULL * a);
    int rc = generate_once(*p, rng, h->nl, err);
    if (rc != PDF_OK) { delete h; set_error("pdf_netlist_generate: " + err); return rc; }
    if (p->path_limit_log2 == 0 || p->path_limit_log2 >= 64 ||
        h->nl.structural_paths < (1ULL << p->path_limit_log2)) {
      *out = h;
      return PDF_OK;
    }
  }
  delete h;
  set_error("pdf_netlist_generate: every attempt had >= 2^" + std::to_string(p->path_limit_log2) +
            " structural paths");
  return PDF_ERR_NETLIST;
}

void pdf_netlist_free(pdf_netlist *nl) { delete nl; }

int pdf_netlist_get_info(const pdf_netlist *h, pdf_netlist_info *info) {
  if (!h || !info) { set_error("pdf_netlist_get_info: null argument"); return PDF_ERR_ARG; }
  const Netlist &nl = h->nl;
  info->n = nl.n;
  info->num_pi = nl.num_pi;
  info->num_po = nl.num_po;
  info->levels = nl.L;
  info->edges = nl.fin.size();
  info->structural_paths = nl.structural_paths;
  return PDF_OK;
}

int pdf_netlist_segments(const pdf_netlist *h, uint32_t segs, uint32_t *out, uint64_t capacity, uint64_t *count) {
  if (!h || !count) { set_error("pdf_netlist_segments: null argument"); return PDF_ERR_ARG; }
  if (segs < 1 || segs > 8) { set_error("pdf_netlist_segments: segment length must be 1..8"); return PDF_ERR_ARG; }
  *count = count_segments(h->nl, segs);
  if (out && capacity) list_segments(h->nl, segs, out, capacity);
  return PDF_OK;
}

int pdf_netlist_export(const pdf_netlist *h, uint8_t *type, uint8_t *po, uint32_t *level,
                       uint32_t *fin_off, uint32_t *fin, uint32_t *fout_off, uint32_t *fout,
                       uint32_t *level_off) {
  if (!h) { set_error("pdf_netlist_export: null netlist"); return PDF_ERR_ARG; }
  const Netlist &nl = h->nl;
  auto cp = [](auto *dst, const auto &v) { if (dst) std::copy(v.begin(), v.end(), dst); };
  cp(type, nl.type);
  cp(po, nl.po);
  cp(level, nl.level);
  cp(fin_off, nl.fin_off);
  cp(fin, nl.fin);
  cp(fout_off, nl.fout_off);
  cp(fout, nl.fout);
  cp(level_off, nl.level_off);
  return PDF_OK;
}

int pdf_gen_vectors(uint64_t seed, uint32_t num_pi, uint64_t num_vectors, uint64_t *out) {
  if (!out && num_pi && num_vectors) { set_error("pdf_gen_vectors: null output"); return PDF_ERR_ARG; }
  const uint64_t words = (num_vectors + 63) / 64;
  Rng rng(seed);
  const uint64_t tail = num_vectors % 64 ? ((1ULL << (num_vectors % 64)) - 1) : ~0ULL;
  for (uint64_t i = 0; i < num_pi; ++i)
    for (uint64_t k = 0; k < words; ++k) {
      uint64_t w = rng.next();
      if (k + 1 == words) w &= tail;
      out[i * words + k] = w;
    }
  return PDF_OK;
}

int pdf_gen_toggle_pairs(uint64_t seed, uint32_t num_pi, uint64_t count, double q_lo, double q_hi,
                         uint64_t *a_bits, uint64_t *b_bits) {
  if ((!a_bits || !b_bits) && num_pi && count) { set_error("pdf_gen_toggle_pairs: null output"); return PDF_ERR_ARG; }
  if (!(q_lo > 0.0) || !(q_lo <= q_hi) || !(q_hi <= 1.0)) {
    set_error("pdf_gen_toggle_pairs: need 0 < q_lo <= q_hi <= 1");
    return PDF_ERR_ARG;
  }
  const uint64_t words = (count + 63) / 64;
  int rc = pdf_gen_vectors(seed, num_pi, count, a_bits);
  if (rc) return rc;
  std::copy(a_bits, a_bits + (size_t)num_pi * words, b_bits);
  // per pair: q log-uniform in [q_lo, q_hi]; the flipped PIs are found by
  // geometric skipping (gap ~ Geometric(q)), O(flips) per pair
  Rng rng(seed ^ 0x5851F42D4C957F2DULL);
  const double l0 = std::log(q_lo), l1 = std::log(q_hi);
  for (uint64_t p = 0; p < count; ++p) {
    const double q = std::exp(l0 + rng.uniform() * (l1 - l0));
    const double lq = std::log1p(-std::min(q, 1.0 - 1e-16));
    for (int64_t i = -1;;) {
      const double gap = std::floor(std::log1p(-rng.uniform()) / lq);
      if (!(gap < (double)num_pi)) break;
      i += 1 + (int64_t)gap;
      if (i >= (int64_t)num_pi) break;
      b_bits[(uint64_t)i * words + (p >> 6)] ^= 1ULL << (p & 63);
    }
  }
  return PDF_OK;
}

// ------------------------------------------------------------- vector files
// One vector per line, one '0'/'1' per PI in INPUT order; whitespace is
// ignored and '#' starts a comment (the reference's reader lives in its
// absent util/ submodule, main.cu:146-156, so this format is the library's).

static int scan_vectors(const char *path, uint32_t num_pi, uint64_t *bits, uint64_t capacity, uint64_t *count) {
  std::ifstream in(path);
  if (!in) { set_error(std::string("cannot open vector file ") + path); return PDF_ERR_IO; }
  const uint64_t words = (capacity + 63) / 64;
  std::string line;
  uint64_t v = 0, lineno = 0;
  while (std::getline(in, line)) {
    ++lineno;
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    uint32_t width = 0;
    for (char c : line) {
      if (c == '0' || c == '1') {
        if (bits && width < num_pi && v < capacity && c == '1')
          bits[(uint64_t)width * words + v / 64] |= 1ULL << (v % 64);
        ++width;
      } else if (!std::isspace((unsigned char)c)) {
        set_error(std::string(path) + ":" + std::to_string(lineno) + ": bad character");
        return PDF_ERR_IO;
      }
    }
    if (width == 0) continue;
    if (width != num_pi) {
      set_error(std::string(path) + ":" + std::to_string(lineno) + ": vector of width " + std::to_string(width) +
                ", circuit has " + std::to_string(num_pi) + " inputs");
      return PDF_ERR_IO;
    }
    if (bits && v >= capacity) { set_error("pdf_vectors_read: more vectors than capacity"); return PDF_ERR_ARG; }
    ++v;
  }
  *count = v;
  return PDF_OK;
}

int pdf_vectors_count(const char *path, uint32_t num_pi, uint64_t *count) {
  if (!path || !count) { set_error("null argument"); return PDF_ERR_ARG; }
  return scan_vectors(path, num_pi, nullptr, 0, count);
}

int pdf_vectors_read(const char *path, uint32_t num_pi, uint64_t *bits, uint64_t capacity, uint64_t *count) {
  if (!path || !bits || !count) { set_error("null argument"); return PDF_ERR_ARG; }
  std::fill(bits, bits + (size_t)num_pi * ((capacity + 63) / 64), 0ULL);
  return scan_vectors(path, num_pi, bits, capacity, count);
}

int pdf_vectors_write(const char *path, uint32_t num_pi, const uint64_t *bits, uint64_t count) {
  if (!path || (!bits && count)) { set_error("null argument"); return PDF_ERR_ARG; }
  std::ofstream out(path);
  if (!out) { set_error(std::string("cannot write ") + path); return PDF_ERR_IO; }
  const uint64_t words = (count + 63) / 64;
  std::string row(num_pi, '0');
  out << "# " << count << " vectors x " << num_pi << " inputs\n";
  for (uint64_t v = 0; v < count; ++v) {
    for (uint32_t i = 0; i < num_pi; ++i) row[i] = ((bits[(uint64_t)i * words + v / 64] >> (v % 64)) & 1) ? '1' : '0';
    out << row << '\n';
  }
  if (!out) { set_error(std::string("write failed: ") + path); return PDF_ERR_IO; }
  return PDF_OK;
}

// ------------------------------------------------------------- file formats

static std::string trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

static int type_from_name(std::string t) {
  for (auto &c : t) c = (char)std::toupper((unsigned char)c);
  static const std::map<std::string, int> m = {
      {"AND", PDF_AND}, {"NAND", PDF_NAND}, {"OR", PDF_OR},   {"NOR", PDF_NOR},
      {"XOR", PDF_XOR}, {"XNOR", PDF_XNOR}, {"BUFF", PDF_BUFF}, {"BUF", PDF_BUFF},
      {"NOT", PDF_NOT}, {"FROM", PDF_FROM}, {"DFF", PDF_DFF},  {"INPT", PDF_INPT}};
  auto it = m.find(t);
  return it == m.end() ? -1 : it->second;
}

static const char *type_name(int t) {
  static const char *names[] = {"UNKN", "INPT", "AND", "NAND", "OR", "NOR",
                                "XOR",  "XNOR", "BUFF", "NOT", "FROM", "DFF"};
  return (t >= 0 && t <= PDF_DFF) ? names[t] : "UNKN";
}

int pdf_netlist_read_bench(const char *path, pdf_netlist **out) {
  if (!path || !out) { set_error("pdf_netlist_read_bench: null argument"); return PDF_ERR_ARG; }
  std::ifstream in(path);
  if (!in) { set_error(std::string("pdf_netlist_read_bench: cannot open ") + path); return PDF_ERR_IO; }
  struct Def { int type; std::vector<std::string> ins; size_t order; };
  std::vector<std::string> inputs, outputs;
  std::unordered_map<std::string, Def> defs;
  std::vector<std::string> def_order;
  std::string line;
  size_t lineno = 0;
  while (std::getline(in, line)) {
    ++lineno;
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    line = trim(line);
    if (line.empty()) continue;
    auto bad = [&](const std::string &why) {
      set_error(std::string("pdf_netlist_read_bench: ") + path + ":" + std::to_string(lineno) + ": " + why);
      return PDF_ERR_IO;
    };
    const size_t eq = line.find('=');
    const size_t lp = line.find('('), rp = line.rfind(')');
    if (lp == std::string::npos || rp == std::string::npos || rp < lp) return bad("expected '(...)'");
    if (eq == std::string::npos) {
      std::string kw = trim(line.substr(0, lp));
      for (auto &c : kw) c = (char)std::toupper((unsigned char)c);
      const std::string name = trim(line.substr(lp + 1, rp - lp - 1));
      if (kw == "INPUT") inputs.push_back(name);
      else if (kw == "OUTPUT") outputs.push_back(name);
      else return bad("unknown declaration " + kw);
      continue;
    }
    const std::string lhs = trim(line.substr(0, eq));
    const std::string tname = trim(line.substr(eq + 1, lp - eq - 1));
    const int ty = type_from_name(tname);
    if (ty < 0 || ty == PDF_INPT) return bad("unknown gate type " + tname);
    Def d{ty, {}, def_order.size()};
    std::stringstream ss(line.substr(lp + 1, rp - lp - 1));
    std::string tok;
    while (std::getline(ss, tok, ',')) {
      tok = trim(tok);
      if (!tok.empty()) d.ins.push_back(tok);
    }
    if (d.ins.empty()) return bad("gate " + lhs + " has no inputs");
    if (defs.count(lhs)) return bad("signal " + lhs + " defined twice");
    defs.emplace(lhs, d);
    def_order.push_back(lhs);
  }
  std::unordered_map<std::string, uint32_t> level;
  for (auto &s : inputs) {
    if (defs.count(s)) { set_error("pdf_netlist_read_bench: input " + s + " is also driven by a gate"); return PDF_ERR_NETLIST; }
    if (!level.emplace(s, 0).second) { set_error("pdf_netlist_read_bench: input " + s + " declared twice"); return PDF_ERR_NETLIST; }
  }
  // levelize (iterative DFS, cycle detection)
  std::unordered_map<std::string, int> state;  // 1 = on stack, 2 = done
  for (auto &root : def_order) {
    if (level.count(root)) continue;
    std::vector<std::pair<std::string, size_t>> st{{root, 0}};
    state[root] = 1;
    while (!st.empty()) {
      auto &top = st.back();
      const Def &d = defs.at(top.first);
      if (top.second < d.ins.size()) {
        const std::string &s = d.ins[top.second++];
        if (level.count(s)) continue;
        if (!defs.count(s)) { set_error("pdf_netlist_read_bench: undriven signal " + s); return PDF_ERR_NETLIST; }
        if (state[s] == 1) { set_error("pdf_netlist_read_bench: combinational cycle through " + s); return PDF_ERR_NETLIST; }
        state[s] = 1;
        st.push_back({s, 0});
      } else {
        uint32_t lv = 0;
        for (auto &s : d.ins) lv = std::max(lv, level.at(s) + 1);
        level[top.first] = lv;
        state[top.first] = 2;
        st.pop_back();
      }
    }
  }
  std::vector<std::string> gates = def_order;
  std::stable_sort(gates.begin(), gates.end(),
                   [&](const std::string &a, const std::string &b) { return level.at(a) < level.at(b); });
  std::unordered_map<std::string, uint32_t> id;
  for (uint32_t i = 0; i < inputs.size(); ++i) id[inputs[i]] = i;
  for (uint32_t i = 0; i < gates.size(); ++i) id[gates[i]] = (uint32_t)inputs.size() + i;
  auto *h = new (std::nothrow) pdf_netlist;
  if (!h) { set_error("out of memory"); return PDF_ERR_NOMEM; }
  Netlist &nl = h->nl;
  nl.n = (uint32_t)(inputs.size() + gates.size());
  nl.num_pi = (uint32_t)inputs.size();
  nl.type.assign(nl.n, PDF_INPT);
  nl.po.assign(nl.n, 0);
  nl.fin_off.assign(nl.n + 1, 0);
  for (uint32_t i = 0; i < gates.size(); ++i) {
    const Def &d = defs.at(gates[i]);
    const uint32_t g = nl.num_pi + i;
    nl.type[g] = (uint8_t)d.type;
    for (auto &s : d.ins) nl.fin.push_back(id.at(s));
    nl.fin_off[g + 1] = (uint32_t)nl.fin.size();
  }
  for (uint32_t g = 0; g < nl.num_pi; ++g) nl.fin_off[g + 1] = 0;
  for (auto &s : outputs) {
    auto it = id.find(s);
    if (it == id.end()) { delete h; set_error("pdf_netlist_read_bench: output " + s + " is undefined"); return PDF_ERR_NETLIST; }
    nl.po[it->second] = 1;
  }
  std::string err;
  int rc = finalize_netlist(nl, err);
  if (rc != PDF_OK) { delete h; set_error("pdf_netlist_read_bench: " + err); return rc; }
  *out = h;
  return PDF_OK;
}

int pdf_netlist_write_level(const pdf_netlist *h, const char *path) {
  if (!h || !path) { set_error("pdf_netlist_write_level: null argument"); return PDF_ERR_ARG; }
  FILE *f = std::fopen(path, "w");
  if (!f) { set_error(std::string("pdf_netlist_write_level: cannot open ") + path); return PDF_ERR_IO; }
  const Netlist &nl = h->nl;
  std::fprintf(f, "pdflevel 1\n%u %u\n", nl.n, nl.num_pi);
  for (uint32_t g = 0; g < nl.n; ++g) {
    std::fprintf(f, "%u %s %u %u", g, type_name(nl.type[g]), (unsigned)nl.po[g], nl.fin_off[g + 1] - nl.fin_off[g]);
    for (uint32_t k = nl.fin_off[g]; k < nl.fin_off[g + 1]; ++k) std::fprintf(f, " %u", nl.fin[k]);
    std::fputc('\n', f);
  }
  const bool ok = std::fclose(f) == 0;
  if (!ok) { set_error("pdf_netlist_write_level: write failed"); return PDF_ERR_IO; }
  return PDF_OK;
}

int pdf_netlist_read_level(const char *path, pdf_netlist **out) {
  if (!path || !out) { set_error("pdf_netlist_read_level: null argument"); return PDF_ERR_ARG; }
  std::ifstream in(path);
  if (!in) { set_error(std::string("pdf_netlist_read_level: cannot open ") + path); return PDF_ERR_IO; }
  std::string magic;
  int ver = 0;
  uint32_t n = 0, npi = 0;
  if (!(in >> magic >> ver >> n >> npi) || magic != "pdflevel" || ver != 1) {
    set_error(std::string("pdf_netlist_read_level: bad header in ") + path);
    return PDF_ERR_IO;
  }
  auto *h = new (std::nothrow) pdf_netlist;
  if (!h) { set_error("out of memory"); return PDF_ERR_NOMEM; }
  Netlist &nl = h->nl;
  nl.n = n;
  nl.num_pi = npi;
  nl.type.assign(n, 0);
  nl.po.assign(n, 0);
  nl.fin_off.assign(n + 1, 0);
  for (uint32_t g = 0; g < n; ++g) {
    uint32_t id = 0, po = 0, nfi = 0;
    std::string tn;
    if (!(in >> id >> tn >> po >> nfi) || id != g) {
      delete h;
      set_error(std::string("pdf_netlist_read_level: bad record for node ") + std::to_string(g));
      return PDF_ERR_IO;
    }
    const int ty = type_from_name(tn);
    nl.type[g] = (uint8_t)(ty < 0 ? PDF_UNKN : ty);
    nl.po[g] = (uint8_t)po;
    for (uint32_t k = 0; k < nfi; ++k) {
      uint32_t f = 0;
      if (!(in >> f)) { delete h; set_error("pdf_netlist_read_level: truncated fanin list"); return PDF_ERR_IO; }
      nl.fin.push_back(f);
    }
    nl.fin_off[g + 1] = (uint32_t)nl.fin.size();
  }
  std::string err;
  int rc = finalize_netlist(nl, err);
  if (rc != PDF_OK) { delete h; set_error("pdf_netlist_read_level: " + err); return rc; }
  *out = h;
  return PDF_OK;
}

}  // extern "C"
