// Internal host-side types shared by the netlist code and the HIP grader.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "pdfgrade.h"

namespace pdf {

// Levelized netlist in CSR form.  Node ids are sorted by level; INPT nodes are
// ids 0..num_pi-1 (level 0).  fin lists keep the caller's order (fanin 0 is
// the on-path input for marking); fout lists are sorted by fanout id.
struct Netlist {
  uint32_t n = 0, num_pi = 0, num_po = 0, L = 0;  // L = max level index
  std::vector<uint8_t> type, po;
  std::vector<uint32_t> level;      // [n]
  std::vector<uint32_t> level_off;  // [L+2]
  std::vector<uint32_t> fin_off, fin, fout_off, fout;
  uint64_t structural_paths = 0;    // saturating PI->PO path count
};

// Build derived arrays (levels, fanouts, path count) and validate.  Returns
// PDF_OK or PDF_ERR_NETLIST with a message in `err`.
int finalize_netlist(Netlist &nl, std::string &err);

// Segments of length K (1..8): for every node g0 in id order, every fanout
// walk of K nodes, fanouts in CSR order, depth first; a walk reaching a node
// without fanouts first ends there, the rest of its K entries PDF_NONE.
// count_segments saturates at UINT64_MAX; list_segments writes at most `cap`.
uint64_t count_segments(const Netlist &nl, uint32_t K);
void list_segments(const Netlist &nl, uint32_t K, uint32_t *out, uint64_t cap);

// Thread-local last-error text.
void set_error(const std::string &msg);

// SplitMix64 (Steele, Lea & Flood 2014) -> xoshiro256** (Blackman & Vigna
// 2018).
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) s[i] = splitmix(x);
  }
  static uint64_t splitmix(uint64_t &x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  uint64_t next() {
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  // uniform in [0, m) by the multiply-shift method (Lemire 2019)
  uint64_t below(uint64_t m) {
    return (uint64_t)(((unsigned __int128)next() * m) >> 64);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace pdf

struct pdf_netlist {
  pdf::Netlist nl;
};
