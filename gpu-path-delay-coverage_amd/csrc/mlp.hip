// BF16 multilayer perceptron on MFMA for the learned pair ranking (no
// reference counterpart): 512 -> 256 -> 256 -> 1 with ReLU, FP32
// accumulation.  The forward pass, the backward pass and Adam are
// hand-written kernels; every matrix product goes through one LDS-tiled
// kernel built on v_mfma_f32_16x16x32_bf16 (gfx950).
//
// Precision contract (the CPU gradient oracle follows it step by step):
//   inputs X and the weight copies W1, W2 are BF16; products accumulate in
//   FP32; H1 = bf16(relu(X W1^T + b1)), H2 = bf16(relu(H1 W2^T + b2)),
//   y = H2 . w3 + b3 in FP32.  Loss = mean (y - t)^2.  Backward:
//   dy = 2 (y - t) / rows; dZ2 = bf16(dy w3 [H2 > 0]); dZ1 = bf16((dZ2 W2)
//   [H1 > 0]); weight and bias gradients accumulate in FP32.  The FP32 master
//   parameters, Adam moments and gradients share one flat layout:
//   W1[256][512] b1[256] W2[256][256] b2[256] w3[256] b3[1].

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mlp_internal.hpp"
#include "pdf_internal.hpp"

namespace {

using bf16x8 = __bf16 __attribute__((ext_vector_type(8)));
using floatx4 = float __attribute__((ext_vector_type(4)));

constexpr uint32_t kIn = PDF_FEATURES, kHid = 256;
constexpr size_t kW1 = 0, kB1 = kW1 + (size_t)kHid * kIn, kW2 = kB1 + kHid, kB2 = kW2 + (size_t)kHid * kHid,
                 kW3 = kB2 + kHid, kB3 = kW3 + kHid, kNumParams = kB3 + 1;
static_assert(kNumParams == PDF_MLP_PARAMS, "flat parameter layout");

constexpr uint64_t kSlab = 131072;   // rows per forward slab (bounds the H1/H2 workspace)
constexpr uint32_t kHeadRows = 64;   // rows per block of the column-reduction kernels
constexpr uint32_t kMaxSplits = 32;  // K splits of the weight-gradient products

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite values)
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ bool bf_pos(uint16_t h) { return !(h & 0x8000u) && h; }

// ------------------------------------------------------------------- GEMM
enum Epi { EPI_BIAS_RELU = 0, EPI_RELU_GRAD = 1, EPI_PARTIAL = 2 };

struct EpiArgs {
  const float *bias;     // EPI_BIAS_RELU: [N]
  const uint16_t *mask;  // EPI_RELU_GRAD: bf16 [M][ldc]; the product passes where mask > 0
  uint16_t *out16;       // EPI_BIAS_RELU / EPI_RELU_GRAD: bf16 [M][ldc]
  float *out32;          // EPI_PARTIAL: [splits][M][ldc]
  uint32_t ldc;
};

constexpr int kBK = 64;        // k per LDS stage (two MFMAs deep)
constexpr int kLd = kBK + 8;   // LDS row pitch in bf16 (144 bytes, keeps 16-byte alignment)

// LDS position of k-chunk `chunk` (8 bf16, 16 bytes) of tile row `row`: the
// chunk index is XORed with bits 3..5 of the row, so the transposed stores
// (eight rows 8 apart per lane group) and the MFMA operand reads (16 rows,
// one chunk) spread over the banks.
__device__ __forceinline__ uint32_t swz(uint32_t row, uint32_t chunk) {
  return row * kLd + ((chunk ^ ((row >> 3) & 7)) << 3);
}

// One operand tile of R rows x kBK k into registers.  Rows are the M (or N)
// index, k the reduction index.  !T: element (r, k) at P[r * ld + k] (k
// contiguous); T: at P[k * ld + r] (r contiguous).  Loads are 16 bytes;
// those outside [0, rmax) x [k0, kend) read as zero.
template <bool T, int R>
struct Tile {
  static constexpr int kPer = R * kBK / 8 / 256;  // uint4 per thread
  uint4 reg[kPer];

  __device__ __forceinline__ void load(const uint16_t *__restrict__ P, uint32_t ld, uint32_t r0, uint32_t rmax,
                                       uint32_t k0, uint32_t kend) {
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const uint32_t q = threadIdx.x + 256u * u;
      uint32_t r, k;
      if (!T) {
        r = q / (kBK / 8);
        k = (q % (kBK / 8)) * 8;
      } else {
        k = q / (R / 8);
        r = (q % (R / 8)) * 8;
      }
      const bool ok = r0 + r < rmax && k0 + k < kend;
      const uint16_t *src = !T ? P + (uint64_t)(r0 + r) * ld + k0 + k : P + (uint64_t)(k0 + k) * ld + r0 + r;
      reg[u] = ok ? *reinterpret_cast<const uint4 *>(src) : make_uint4(0, 0, 0, 0);
    }
  }

  // LDS layout: row r, k contiguous within 16-byte chunks placed by swz().
  __device__ __forceinline__ void store(uint16_t *S) const {
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const uint32_t q = threadIdx.x + 256u * u;
      if (!T) {
        *reinterpret_cast<uint4 *>(&S[swz(q / (kBK / 8), q % (kBK / 8))]) = reg[u];
      } else {
        const uint32_t k = q / (R / 8), r = (q % (R / 8)) * 8;
        const uint32_t w[4] = {reg[u].x, reg[u].y, reg[u].z, reg[u].w};
#pragma unroll
        for (int e = 0; e < 8; ++e) S[swz(r + e, k >> 3) + (k & 7)] = (uint16_t)(w[e >> 1] >> (16 * (e & 1)));
      }
    }
  }
};

// C[M][N] = sum_k A(i, k) B(k, j) over the block's k split, with
// A(i, k) = TA ? A[k * lda + i] : A[i * lda + k] and
// B(k, j) = TB ? B[j * ldb + k] : B[k * ldb + j].
// Block = BM x BN tile of C, 4 waves as 2 x 2, each wave (BM/2) x (BN/2) in
// 16 x 16 MFMA tiles.  Split z covers k in [z * kchunk, (z + 1) * kchunk).
// Bit s of `skip` drops the k stage [s kBK, (s + 1) kBK): the caller knows
// those rows of A are zero (their products would add nothing).
// Host-checked: leading dimensions are multiples of 8, kchunk a multiple of
// kBK, pointers 16-byte aligned.  Loads move 8 elements along an operand's
// contiguous dimension and are taken whole when their first element is in
// range, so each operand must be readable (and finite) up to the next
// multiple of 8 of that dimension; the extra products meet zeros of the
// other operand (k) or land in rows and columns the epilogue drops (i, j).
template <bool TA, bool TB, int BM, int BN, int EPI>
__global__ __launch_bounds__(256) void k_gemm(const uint16_t *__restrict__ A, uint32_t lda,
                                              const uint16_t *__restrict__ B, uint32_t ldb, uint32_t M, uint32_t N,
                                              uint32_t K, uint32_t kchunk, uint32_t skip, EpiArgs ep) {
  static_assert(BM == 128 && BN == 128, "the epilogue stages 64-row halves of a 128 x 128 tile");
  // operand tiles during the k loop; the epilogue reuses the space for one
  // half of the C tile in FP32 (64 x (BN + 4) floats)
  constexpr int kTileBytes = (BM + BN) * kLd * 2;
  constexpr int kStageLd = BN + 4;
  constexpr int kStageBytes = 64 * kStageLd * 4;
  __shared__ __attribute__((aligned(16))) unsigned char lds[kTileBytes > kStageBytes ? kTileBytes : kStageBytes];
  uint16_t *As = reinterpret_cast<uint16_t *>(lds);
  uint16_t *Bs = As + BM * kLd;
  float *stage = reinterpret_cast<float *>(lds);
  constexpr int MI = BM / 32, NI = BN / 32;
  const uint32_t n0 = blockIdx.x * BN, m0 = blockIdx.y * BM;
  const uint32_t kbeg = blockIdx.z * kchunk, kend = min(K, kbeg + kchunk);
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t wr = wave >> 1, wc = wave & 1;

  floatx4 acc[MI][NI];
#pragma unroll
  for (int mi = 0; mi < MI; ++mi)
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) acc[mi][ni] = floatx4{0.f, 0.f, 0.f, 0.f};

  Tile<TA, BM> ta;
  Tile<!TB, BN> tb;  // B's rows are the N index: k contiguous when TB
  auto next = [&](uint32_t k) {  // the first stage at or after k that is not skipped
    while (k < kend && ((skip >> (k / kBK)) & 1)) k += kBK;
    return k;
  };
  uint32_t k0 = next(kbeg);
  if (k0 < kend) {
    ta.load(A, lda, m0, M, k0, kend);
    tb.load(B, ldb, n0, N, k0, kend);
  }
  for (; k0 < kend;) {
    const uint32_t kn = next(k0 + kBK);
    __syncthreads();  // the previous stage's MFMA operands have been read
    ta.store(As);
    tb.store(Bs);
    __syncthreads();
    if (kn < kend) {  // prefetch the next stage while this one computes
      ta.load(A, lda, m0, M, kn, kend);
      tb.load(B, ldb, n0, N, kn, kend);
    }
#pragma unroll
    for (int ks = 0; ks < kBK / 32; ++ks) {
      bf16x8 a[MI], b[NI];
      // lane l supplies row (l & 15) and k 8 (l >> 4) .. 8 (l >> 4) + 7 of the substep
#pragma unroll
      for (int mi = 0; mi < MI; ++mi)
        a[mi] = *reinterpret_cast<const bf16x8 *>(
            &As[swz(wr * (BM / 2) + mi * 16 + (lane & 15), 4 * ks + (lane >> 4))]);
#pragma unroll
      for (int ni = 0; ni < NI; ++ni)
        b[ni] = *reinterpret_cast<const bf16x8 *>(
            &Bs[swz(wc * (BN / 2) + ni * 16 + (lane & 15), 4 * ks + (lane >> 4))]);
#pragma unroll
      for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni)
          acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mi], b[ni], acc[mi][ni], 0, 0, 0);
    }
    k0 = kn;
  }

  // Epilogue, one 64-row half at a time: the waves holding that half write
  // their accumulators (lane l: rows 4 (l >> 4) + r of column l & 15 of each
  // 16 x 16 tile) into the FP32 stage, then every thread stores 8 consecutive
  // columns of a row with 16-byte accesses.
#pragma unroll
  for (uint32_t half = 0; half < 2; ++half) {
    __syncthreads();  // operand tiles (or the previous half) are no longer read
    if (wr == half) {
#pragma unroll
      for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            stage[(mi * 16 + 4 * (lane >> 4) + r) * kStageLd + wc * (BN / 2) + ni * 16 + (lane & 15)] =
                acc[mi][ni][r];
    }
    __syncthreads();
    for (uint32_t item = threadIdx.x; item < 64 * (BN / 8); item += 256) {
      const uint32_t r = item / (BN / 8), c8 = (item % (BN / 8)) * 8;
      const uint32_t i = m0 + half * 64 + r, j = n0 + c8;
      if (i >= M || j >= N) continue;  // N is a multiple of 8: whole groups
      const float *v = &stage[r * kStageLd + c8];
      const uint64_t o = (uint64_t)i * ep.ldc + j;
      if (EPI == EPI_PARTIAL) {
        float *dst = ep.out32 + (uint64_t)blockIdx.z * M * ep.ldc + o;
        *reinterpret_cast<float4 *>(dst) = make_float4(v[0], v[1], v[2], v[3]);
        *reinterpret_cast<float4 *>(dst + 4) = make_float4(v[4], v[5], v[6], v[7]);
      } else {
        uint32_t w[4];
        if (EPI == EPI_BIAS_RELU) {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            w[e] = (uint32_t)f2bf(fmaxf(v[2 * e] + ep.bias[j + 2 * e], 0.f)) |
                   ((uint32_t)f2bf(fmaxf(v[2 * e + 1] + ep.bias[j + 2 * e + 1], 0.f)) << 16);
        } else {
          const uint4 mk = *reinterpret_cast<const uint4 *>(ep.mask + o);
          const uint32_t mw[4] = {mk.x, mk.y, mk.z, mk.w};
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const uint16_t lo = bf_pos((uint16_t)mw[e]) ? f2bf(v[2 * e]) : (uint16_t)0;
            const uint16_t hi = bf_pos((uint16_t)(mw[e] >> 16)) ? f2bf(v[2 * e + 1]) : (uint16_t)0;
            w[e] = (uint32_t)lo | ((uint32_t)hi << 16);
          }
        }
        *reinterpret_cast<uint4 *>(ep.out16 + o) = make_uint4(w[0], w[1], w[2], w[3]);
      }
    }
  }
}

// ------------------------------------------------------------ head + sums

// y = H2 . w3 + b3, one wave per row.  With a target: dy = 2 (y - t) / rows
// and sq = (y - t)^2.
__global__ __launch_bounds__(256) void k_head_fwd(const uint16_t *__restrict__ H2, const float *__restrict__ w,
                                                  uint32_t rows, float *__restrict__ y,
                                                  const float *__restrict__ target, float *__restrict__ dy,
                                                  float *__restrict__ sq) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= rows) return;
  const uint2 h = *reinterpret_cast<const uint2 *>(&H2[(uint64_t)row * kHid + 4 * lane]);
  const float *w3 = w + kW3 + 4 * lane;
  float s = bf2f((uint16_t)h.x) * w3[0] + bf2f((uint16_t)(h.x >> 16)) * w3[1] + bf2f((uint16_t)h.y) * w3[2] +
            bf2f((uint16_t)(h.y >> 16)) * w3[3];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) {
    const float v = s + w[kB3];
    y[row] = v;
    if (target) {
      const float d = v - target[row];
      dy[row] = 2.f * d / (float)rows;
      sq[row] = d * d;
    }
  }
}

// Per block of kHeadRows rows, thread j = column: dZ2 = bf16(dy w3_j
// [H2 > 0]) and the partial sums of dw3 = sum dy H2, db2 = sum dZ2 and
// db3 = sum dy.
__global__ __launch_bounds__(256) void k_head_bwd(const uint16_t *__restrict__ H2, const float *__restrict__ w,
                                                  const float *__restrict__ dy, uint32_t rows,
                                                  uint16_t *__restrict__ dZ2, float *__restrict__ part) {
  const uint32_t j = threadIdx.x, r0 = blockIdx.x * kHeadRows, r1 = min(rows, r0 + kHeadRows);
  const float w3 = w[kW3 + j];
  float dw3 = 0.f, db2 = 0.f, db3 = 0.f;
  for (uint32_t i = r0; i < r1; ++i) {
    const uint64_t o = (uint64_t)i * kHid + j;
    const uint16_t h = H2[o];
    const float d = dy[i];
    dw3 += d * bf2f(h);
    const uint16_t z = bf_pos(h) ? f2bf(d * w3) : (uint16_t)0;
    dZ2[o] = z;
    db2 += bf2f(z);
    db3 += d;
  }
  // part rows: [0, nblk) dw3, [nblk, 2 nblk) db2, [2 nblk, 3 nblk) db3 (col 0)
  const uint32_t nblk = gridDim.x;
  part[(uint64_t)blockIdx.x * kHid + j] = dw3;
  part[(uint64_t)(nblk + blockIdx.x) * kHid + j] = db2;
  if (j == 0) part[(uint64_t)(2 * nblk + blockIdx.x) * kHid] = db3;
}

// Partial column sums of a bf16 [rows][256] matrix.
__global__ __launch_bounds__(256) void k_colsum(const uint16_t *__restrict__ Z, uint32_t rows,
                                                float *__restrict__ part) {
  const uint32_t j = threadIdx.x, r0 = blockIdx.x * kHeadRows, r1 = min(rows, r0 + kHeadRows);
  float s = 0.f;
  for (uint32_t i = r0; i < r1; ++i) s += bf2f(Z[(uint64_t)i * kHid + j]);
  part[(uint64_t)blockIdx.x * kHid + j] = s;
}

// out[c] = sum over p of part[p * pstride + c], in p order (deterministic).
__global__ void k_reduce(const float *__restrict__ part, uint32_t nparts, uint64_t pstride, uint64_t ncol,
                         float *__restrict__ out) {
  const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  float s = 0.f;
  for (uint32_t p = 0; p < nparts; ++p) s += part[p * pstride + c];
  out[c] = s;
}

// Mean of sq[0, rows) by one block, fixed order.
__global__ __launch_bounds__(1024) void k_mean(const float *__restrict__ sq, uint32_t rows, float *__restrict__ out) {
  __shared__ float s[1024];
  float v = 0.f;
  for (uint32_t i = threadIdx.x; i < rows; i += 1024) v += sq[i];
  s[threadIdx.x] = v;
  __syncthreads();
  for (uint32_t h = 512; h > 0; h >>= 1) {
    if (threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0] / (float)rows;
}

// Adam (Kingma & Ba 2015) in the arrangement of torch.optim.Adam:
// p -= (lr / bc1) m / (sqrt(v) / sqrt(bc2) + eps); refreshes the BF16 weight
// copies.
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float beta1, float beta2, float step_size, float inv_sqrt_bc2,
                       float eps, uint16_t *__restrict__ w1b, uint16_t *__restrict__ w2b) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kNumParams) return;
  const float gi = g[i];
  const float mi = beta1 * m[i] + (1.f - beta1) * gi;
  const float vi = beta2 * v[i] + (1.f - beta2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float p = w[i] - step_size * mi / (sqrtf(vi) * inv_sqrt_bc2 + eps);
  w[i] = p;
  if (i < kB1) w1b[i - kW1] = f2bf(p);
  else if (i >= kW2 && i < kB2) w2b[i - kW2] = f2bf(p);
}

__global__ void k_to_bf16(const float *__restrict__ w, uint16_t *__restrict__ w1b, uint16_t *__restrict__ w2b) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kB1) w1b[i - kW1] = f2bf(w[i]);
  else if (i >= kW2 && i < kB2) w2b[i - kW2] = f2bf(w[i]);
}

// ------------------------------------------------------------- host side

bool ok(hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  pdf::set_error(std::string(what) + ": " + hipGetErrorString(e));
  return false;
}
#define MLPCHK(expr)                                \
  do {                                              \
    if (!ok((expr), #expr)) return PDF_ERR_HIP;     \
  } while (0)

template <class T>
int alloc(T **p, size_t count) {
  if (hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    pdf::set_error("hipMalloc failed (" + std::to_string(count * sizeof(T)) + " bytes)");
    return PDF_ERR_NOMEM;
  }
  return PDF_OK;
}
template <class T>
void release(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

inline uint32_t blocks(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

}  // namespace

struct pdf_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;  // kNumParams each
  uint16_t *w1b = nullptr, *w2b = nullptr;
  uint64_t cap = 0;  // rows of the activation workspace
  uint16_t *h1 = nullptr, *h2 = nullptr, *dz2 = nullptr, *dz1 = nullptr;
  float *y = nullptr, *dy = nullptr, *sq = nullptr, *loss = nullptr;
  float *part = nullptr;  // split-K and column-sum partials
  size_t part_cap = 0;
  uint64_t steps = 0;
};

namespace {

void free_work(pdf_mlp *m) {
  release(m->h1);
  release(m->h2);
  release(m->dz2);
  release(m->dz1);
  release(m->y);
  release(m->dy);
  release(m->sq);
  m->cap = 0;
}

int ensure_rows(pdf_mlp *m, uint64_t rows) {
  if (rows <= m->cap) return PDF_OK;
  MLPCHK(hipStreamSynchronize(m->stream));
  free_work(m);
  int rc = alloc(&m->h1, rows * kHid);
  if (!rc) rc = alloc(&m->h2, rows * kHid);
  if (!rc) rc = alloc(&m->dz2, rows * kHid);
  if (!rc) rc = alloc(&m->dz1, rows * kHid);
  if (!rc) rc = alloc(&m->y, rows);
  if (!rc) rc = alloc(&m->dy, rows);
  if (!rc) rc = alloc(&m->sq, rows);
  if (rc) {
    free_work(m);
    return rc;
  }
  m->cap = rows;
  return PDF_OK;
}

int ensure_part(pdf_mlp *m, size_t floats) {
  if (floats <= m->part_cap) return PDF_OK;
  MLPCHK(hipStreamSynchronize(m->stream));
  release(m->part);
  m->part_cap = 0;
  int rc = alloc(&m->part, floats);
  if (!rc) m->part_cap = floats;
  return rc;
}

template <bool TA, bool TB, int EPI>
int gemm(pdf_mlp *m, const uint16_t *A, uint32_t lda, const uint16_t *B, uint32_t ldb, uint32_t M, uint32_t N,
         uint32_t K, uint32_t splits, const EpiArgs &ep, uint32_t skip = 0) {
  constexpr int BM = 128, BN = 128;
  if (lda % 8 || ldb % 8 || ((uintptr_t)A | (uintptr_t)B) % 16 || N % 8 || ep.ldc % 8) {
    pdf::set_error("mlp gemm: operand not 16-byte granular");
    return PDF_ERR_ARG;
  }
  const uint32_t kchunk = (uint32_t)(((K + splits - 1) / splits + kBK - 1) / kBK * kBK);
  const dim3 grid(blocks(N, BN), blocks(M, BM), splits);
  if (grid.y > 65535) {
    pdf::set_error("mlp gemm: too many rows in one launch");
    return PDF_ERR_ARG;
  }
  if (K > 32 * kBK && skip) {
    pdf::set_error("mlp gemm: stage mask covers 32 stages");
    return PDF_ERR_ARG;
  }
  k_gemm<TA, TB, BM, BN, EPI><<<grid, 256, 0, m->stream>>>(A, lda, B, ldb, M, N, K, kchunk, skip, ep);
  MLPCHK(hipGetLastError());
  return PDF_OK;
}

// H1, H2 (and y) of `rows` rows of X (feature-major: X[f * ld + r]).
// zero_stages: bit s says features [64 s, 64 s + 64) are zero in every row.
int forward(pdf_mlp *m, const uint16_t *X, uint64_t ld, uint32_t rows, float *y, const float *target,
            uint32_t zero_stages = 0) {
  static_assert(kBK == 64, "zero_stages counts 64-feature stages");
  int rc;
  EpiArgs e1{m->w + kB1, nullptr, m->h1, nullptr, kHid};
  if ((rc = gemm<true, true, EPI_BIAS_RELU>(m, X, (uint32_t)ld, m->w1b, kIn, rows, kHid, kIn, 1, e1, zero_stages)))
    return rc;
  EpiArgs e2{m->w + kB2, nullptr, m->h2, nullptr, kHid};
  if ((rc = gemm<false, true, EPI_BIAS_RELU>(m, m->h1, kHid, m->w2b, kHid, rows, kHid, kHid, 1, e2))) return rc;
  k_head_fwd<<<blocks(rows, 4), 256, 0, m->stream>>>(m->h2, m->w, rows, y, target, m->dy, m->sq);
  MLPCHK(hipGetLastError());
  return PDF_OK;
}

// Forward + backward of one batch; the gradient lands in m->g.
int gradients(pdf_mlp *m, const uint16_t *X, uint64_t ld, const float *target, uint32_t rows) {
  int rc;
  if ((rc = ensure_rows(m, rows))) return rc;
  const uint32_t nblk = blocks(rows, kHeadRows);
  const uint32_t splits = std::max(1u, std::min(kMaxSplits, rows / 256));
  if ((rc = ensure_part(m, std::max<size_t>((size_t)splits * kHid * kIn, (size_t)3 * nblk * kHid)))) return rc;
  if ((rc = forward(m, X, ld, rows, m->y, target))) return rc;
  k_mean<<<1, 1024, 0, m->stream>>>(m->sq, rows, m->loss);
  // output layer: dZ2, dw3, db2, db3
  k_head_bwd<<<nblk, 256, 0, m->stream>>>(m->h2, m->w, m->dy, rows, m->dz2, m->part);
  k_reduce<<<1, 256, 0, m->stream>>>(m->part, nblk, kHid, kHid, m->g + kW3);
  k_reduce<<<1, 256, 0, m->stream>>>(m->part + (size_t)nblk * kHid, nblk, kHid, kHid, m->g + kB2);
  k_reduce<<<1, 1, 0, m->stream>>>(m->part + (size_t)2 * nblk * kHid, nblk, kHid, 1, m->g + kB3);
  MLPCHK(hipGetLastError());
  // dW2 = dZ2^T H1 (before the partial buffer is reused)
  EpiArgs ep{nullptr, nullptr, nullptr, m->part, kHid};
  if ((rc = gemm<true, false, EPI_PARTIAL>(m, m->dz2, kHid, m->h1, kHid, kHid, kHid, rows, splits, ep))) return rc;
  k_reduce<<<blocks((size_t)kHid * kHid, 256), 256, 0, m->stream>>>(m->part, splits, (uint64_t)kHid * kHid,
                                                                   (uint64_t)kHid * kHid, m->g + kW2);
  // dZ1 = bf16((dZ2 W2) [H1 > 0]), db1
  EpiArgs e1{nullptr, m->h1, m->dz1, nullptr, kHid};
  if ((rc = gemm<false, false, EPI_RELU_GRAD>(m, m->dz2, kHid, m->w2b, kHid, rows, kHid, kHid, 1, e1))) return rc;
  k_colsum<<<nblk, 256, 0, m->stream>>>(m->dz1, rows, m->part);
  k_reduce<<<1, 256, 0, m->stream>>>(m->part, nblk, kHid, kHid, m->g + kB1);
  // dW1 = dZ1^T X
  EpiArgs ew{nullptr, nullptr, nullptr, m->part, kIn};
  if ((rc = gemm<true, true, EPI_PARTIAL>(m, m->dz1, kHid, X, (uint32_t)ld, kHid, kIn, rows, splits, ew))) return rc;
  k_reduce<<<blocks((size_t)kHid * kIn, 256), 256, 0, m->stream>>>(m->part, splits, (uint64_t)kHid * kIn,
                                                                  (uint64_t)kHid * kIn, m->g + kW1);
  MLPCHK(hipGetLastError());
  return PDF_OK;
}

int adam(pdf_mlp *m, float lr, float beta1, float beta2, float eps) {
  m->steps++;
  const double bc1 = 1.0 - std::pow((double)beta1, (double)m->steps);
  const double bc2 = 1.0 - std::pow((double)beta2, (double)m->steps);
  k_adam<<<blocks(kNumParams, 256), 256, 0, m->stream>>>(m->w, m->g, m->m, m->v, beta1, beta2,
                                                         (float)(lr / bc1), (float)(1.0 / std::sqrt(bc2)), eps,
                                                         m->w1b, m->w2b);
  MLPCHK(hipGetLastError());
  return PDF_OK;
}

int check_rows(const pdf_mlp *m, const void *X, uint64_t ld, uint64_t rows) {
  if (!m || !X) {
    pdf::set_error("null argument");
    return PDF_ERR_ARG;
  }
  if (rows == 0 || rows > 0xFFFFFFFFull || ld % 64 || ld < rows || ld > 0xFFFFFFFFull || (uintptr_t)X % 16) {
    pdf::set_error("mlp: need 1 <= rows <= ld < 2^32, ld a multiple of 64, features 16-byte aligned");
    return PDF_ERR_ARG;
  }
  return PDF_OK;
}

}  // namespace

namespace pdf {

hipStream_t mlp_stream(const pdf_mlp *m) { return m->stream; }
int mlp_device(const pdf_mlp *m) { return m->device; }

int mlp_prepare(pdf_mlp *m, const uint16_t *feat_dev, uint64_t ld, uint64_t rows) {
  int rc = check_rows(m, feat_dev, ld, rows);
  if (rc) return rc;
  MLPCHK(hipSetDevice(m->device));
  return ensure_rows(m, std::min<uint64_t>(rows, kSlab));
}

int mlp_queue_forward(pdf_mlp *m, const uint16_t *feat_dev, uint64_t ld, uint64_t r0, uint64_t r1, float *score_dev,
                      uint32_t zero_stages) {
  for (uint64_t a = r0; a < r1; a += kSlab) {
    const uint32_t r = (uint32_t)std::min<uint64_t>(kSlab, r1 - a);
    const int rc = forward(m, feat_dev + a, ld, r, score_dev + a, nullptr, zero_stages);
    if (rc) return rc;
  }
  return PDF_OK;
}

}  // namespace pdf

extern "C" {

int pdf_mlp_create(int device, uint64_t seed, pdf_mlp **out) {
  if (!out) {
    pdf::set_error("null argument");
    return PDF_ERR_ARG;
  }
  *out = nullptr;
  pdf_mlp *m = new pdf_mlp();
  m->device = device;
  int rc = PDF_OK;
  if (!ok(hipSetDevice(device), "hipSetDevice") || !ok(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking),
                                                        "hipStreamCreate"))
    rc = PDF_ERR_HIP;
  if (!rc) rc = alloc(&m->w, kNumParams);
  if (!rc) rc = alloc(&m->m, kNumParams);
  if (!rc) rc = alloc(&m->v, kNumParams);
  if (!rc) rc = alloc(&m->g, kNumParams);
  if (!rc) rc = alloc(&m->w1b, (size_t)kHid * kIn);
  if (!rc) rc = alloc(&m->w2b, (size_t)kHid * kHid);
  if (!rc) rc = alloc(&m->loss, 1);
  if (!rc) {
    // He-uniform weights U(-sqrt(6 / fan_in), +sqrt(6 / fan_in)), zero biases
    std::vector<float> p(kNumParams, 0.f);
    pdf::Rng rng(seed);
    auto fill = [&](size_t off, size_t count, uint32_t fan_in) {
      const double a = std::sqrt(6.0 / fan_in);
      for (size_t i = 0; i < count; ++i) p[off + i] = (float)((2.0 * rng.uniform() - 1.0) * a);
    };
    fill(kW1, (size_t)kHid * kIn, kIn);
    fill(kW2, (size_t)kHid * kHid, kHid);
    fill(kW3, kHid, kHid);
    rc = pdf_mlp_set_params(m, p.data());
  }
  if (rc) {
    pdf_mlp_destroy(m);
    return rc;
  }
  *out = m;
  return PDF_OK;
}

void pdf_mlp_destroy(pdf_mlp *m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  free_work(m);
  release(m->part);
  release(m->w);
  release(m->m);
  release(m->v);
  release(m->g);
  release(m->w1b);
  release(m->w2b);
  release(m->loss);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

int pdf_mlp_set_params(pdf_mlp *m, const float *host) {
  if (!m || !host) {
    pdf::set_error("null argument");
    return PDF_ERR_ARG;
  }
  MLPCHK(hipSetDevice(m->device));
  MLPCHK(hipMemcpyAsync(m->w, host, kNumParams * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MLPCHK(hipMemsetAsync(m->m, 0, kNumParams * sizeof(float), m->stream));
  MLPCHK(hipMemsetAsync(m->v, 0, kNumParams * sizeof(float), m->stream));
  MLPCHK(hipMemsetAsync(m->g, 0, kNumParams * sizeof(float), m->stream));
  k_to_bf16<<<blocks(kB2, 256), 256, 0, m->stream>>>(m->w, m->w1b, m->w2b);
  MLPCHK(hipGetLastError());
  MLPCHK(hipStreamSynchronize(m->stream));
  m->steps = 0;
  return PDF_OK;
}

int pdf_mlp_get_params(pdf_mlp *m, float *host) {
  if (!m || !host) {
    pdf::set_error("null argument");
    return PDF_ERR_ARG;
  }
  MLPCHK(hipSetDevice(m->device));
  MLPCHK(hipMemcpyAsync(host, m->w, kNumParams * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLPCHK(hipStreamSynchronize(m->stream));
  return PDF_OK;
}

int pdf_mlp_forward(pdf_mlp *m, const uint16_t *feat_dev, uint64_t ld, uint64_t rows, float *score_dev) {
  int rc = check_rows(m, feat_dev, ld, rows);
  if (rc) return rc;
  if (!score_dev) {
    pdf::set_error("null argument");
    return PDF_ERR_ARG;
  }
  MLPCHK(hipSetDevice(m->device));
  if ((rc = ensure_rows(m, std::min<uint64_t>(rows, kSlab)))) return rc;
  for (uint64_t r0 = 0; r0 < rows; r0 += kSlab) {
    const uint32_t r = (uint32_t)std::min<uint64_t>(kSlab, rows - r0);
    if ((rc = forward(m, feat_dev + r0, ld, r, score_dev + r0, nullptr))) return rc;
  }
  MLPCHK(hipStreamSynchronize(m->stream));
  return PDF_OK;
}

int pdf_mlp_gradients(pdf_mlp *m, const uint16_t *feat_dev, uint64_t ld, const float *target_dev, uint64_t rows,
                      float *loss, float *grad_host) {
  int rc = check_rows(m, feat_dev, ld, rows);
  if (rc) return rc;
  if (!target_dev) {
    pdf::set_error("null argument");
    return PDF_ERR_ARG;
  }
  MLPCHK(hipSetDevice(m->device));
  if ((rc = gradients(m, feat_dev, ld, target_dev, (uint32_t)rows))) return rc;
  if (loss) MLPCHK(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  if (grad_host) MLPCHK(hipMemcpyAsync(grad_host, m->g, kNumParams * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLPCHK(hipStreamSynchronize(m->stream));
  return PDF_OK;
}

int pdf_mlp_adam(pdf_mlp *m, float lr, float beta1, float beta2, float eps) {
  if (!m) {
    pdf::set_error("null argument");
    return PDF_ERR_ARG;
  }
  MLPCHK(hipSetDevice(m->device));
  int rc = adam(m, lr, beta1, beta2, eps);
  if (rc) return rc;
  MLPCHK(hipStreamSynchronize(m->stream));
  return PDF_OK;
}

int pdf_mlp_train_step(pdf_mlp *m, const uint16_t *feat_dev, uint64_t ld, const float *target_dev, uint64_t rows,
                       float lr, float *loss) {
  int rc = check_rows(m, feat_dev, ld, rows);
  if (rc) return rc;
  if (!target_dev) {
    pdf::set_error("null argument");
    return PDF_ERR_ARG;
  }
  MLPCHK(hipSetDevice(m->device));
  if ((rc = gradients(m, feat_dev, ld, target_dev, (uint32_t)rows))) return rc;
  if ((rc = adam(m, lr, 0.9f, 0.999f, 1e-8f))) return rc;
  if (loss) {
    MLPCHK(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MLPCHK(hipStreamSynchronize(m->stream));
  }
  return PDF_OK;
}

int pdf_mlp_sync(pdf_mlp *m) {
  if (!m) {
    pdf::set_error("null argument");
    return PDF_ERR_ARG;
  }
  MLPCHK(hipSetDevice(m->device));
  MLPCHK(hipStreamSynchronize(m->stream));
  return PDF_OK;
}

}  // extern "C"
