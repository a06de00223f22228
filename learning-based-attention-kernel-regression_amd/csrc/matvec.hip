// Fused attention-kernel matvec  Y = lambda X + exp(min(s E E^T, c)) X  on gfx950.
//
// The n x n kernel is never stored.  A wave owns ITW row tiles of 32 rows and
// streams 32-row column tiles J of E:
//   1. Gram tile on MFMA, computed transposed (D[jj][ii] = e_J[jj] . e_I[ii]) so
//      that each lane owns one output row ii = lane%32 and 16 column indices
//      jj(v) = 8*(v/4) + 4*(lane/32) + v%4 (the 32x32 MFMA output layout);
//   2. exp2 in registers (the row-side operand is pre-scaled by s*log2(e), so
//      the MFMA output is already the exp2 argument);
//   3. the G.X reduction: VALU FMAs for k <= 4 columns, MFMA (FP32 or BF16)
//      with the exp tile fed straight back as the A operand for larger k.
// Column tiles are split over `jsplit` partial sums (jsplit = 8 maps one split
// to one XCD under round-robin dispatch), reduced deterministically by a
// second kernel that also adds lambda X.
//
// Reference behaviour restated: laker/kernels.py:80-130 (exp_safe clamp,
// dense matvec), :154-185 (skip-clamp predicate), :187-270 (matvec).

#include "laker_common.h"
#include "laker_internal.h"
#include <vector>

#include <algorithm>
#include <cmath>
#include <mutex>

namespace laker {

// ---------------------------------------------------------------------------
// Packing.  Row-side (B operand) arrays are pre-scaled by c = s*log2(e).
//   F32:   ka = round_up(d,2) floats per row; MFMA step s contracts dims
//          {s, s + ka/2} (lane half h takes dim s + h*ka/2).
//   BF16 split: ka = round_up(6d,16) bf16 per row from e = hi + mid + lo:
//          col side [hi|hi|mid|hi|lo|mid], row side [hi'|mid'|hi'|lo'|hi'|mid'],
//          so the K=6d contraction is the six leading products of the split
//          (about 2^-24 relative, FP32 accumulation in the MFMA).
//   FP16 split (square operator's k = 1 product): ka = round_up(3d,16) halves
//          per row from e = hi + lo: col side [hi|hi|lo], row side
//          [hi'|lo'|hi'], the three leading products (22 significant bits
//          per operand, FP32 accumulation); used only when max |e| and
//          max |c e| stay far below the FP16 range (2^14).
// Rows are padded with zeros to a multiple of 64.
// ---------------------------------------------------------------------------

template <class TE>
__global__ void pack_f32_kernel(const TE* __restrict__ E, int64_t n, int d, int64_t npad, int ka,
                                float cscale, float* __restrict__ colside, float* __restrict__ rowside) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npad * ka) return;
  int64_t i = idx / ka;
  int p = (int)(idx % ka);
  // natural order: lane half h reads positions [h*ka/2, (h+1)*ka/2), so MFMA
  // step s contracts dims s and s + ka/2.
  float v = 0.f;
  if (i < n && p < d) v = (float)E[i * d + p];
  if (colside) colside[idx] = v;
  if (rowside) rowside[idx] = v * cscale;
}

template <class TE>
__global__ void pack_bf16x4_kernel(const TE* __restrict__ E, int64_t n, int d, int64_t npad, int ka,
                                   float cscale, uint16_t* __restrict__ colside,
                                   uint16_t* __restrict__ rowside) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npad * ka) return;
  int64_t i = idx / ka;
  int p = (int)(idx % ka);
  uint16_t cv = 0, rv = 0;
  if (i < n && p < 6 * d) {
    const int seg = p / d, dim = p % d;
    // three-way bf16 split e = hi + mid + lo (24 significant bits)
    const float e = (float)E[i * d + dim];
    const uint16_t hi = f32_to_bf16_bits(e);
    const float r1 = e - bf16_bits_to_f32(hi);
    const uint16_t mid = f32_to_bf16_bits(r1);
    const uint16_t lo = f32_to_bf16_bits(r1 - bf16_bits_to_f32(mid));
    const float es = e * cscale;
    const uint16_t his = f32_to_bf16_bits(es);
    const float s1 = es - bf16_bits_to_f32(his);
    const uint16_t mids = f32_to_bf16_bits(s1);
    const uint16_t los = f32_to_bf16_bits(s1 - bf16_bits_to_f32(mids));
    // six products: hi.hi' + hi.mid' + mid.hi' + hi.lo' + lo.hi' + mid.mid'
    const uint16_t col_seg[6] = {hi, hi, mid, hi, lo, mid};
    const uint16_t row_seg[6] = {his, mids, his, los, his, mids};
    cv = col_seg[seg];
    rv = row_seg[seg];
  }
  if (colside) colside[idx] = cv;
  if (rowside) rowside[idx] = rv;
}

template <class TE>
__global__ void pack_f16x2_kernel(const TE* __restrict__ E, int64_t n, int d, int64_t npad, int ka,
                                  float cscale, uint16_t* __restrict__ colside, uint16_t* __restrict__ rowside) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npad * ka) return;
  int64_t i = idx / ka;
  int p = (int)(idx % ka);
  uint16_t cv = 0, rv = 0;
  if (i < n && p < 3 * d) {
    const int seg = p / d, dim = p % d;
    const float e = (float)E[i * d + dim];
    const _Float16 hi = (_Float16)e;
    const _Float16 lo = (_Float16)(e - (float)hi);
    const float es = e * cscale;
    const _Float16 his = (_Float16)es;
    const _Float16 los = (_Float16)(es - (float)his);
    // three products: hi.hi' + hi.lo' + lo.hi'
    const _Float16 col_seg[3] = {hi, hi, lo};
    const _Float16 row_seg[3] = {his, los, his};
    cv = __builtin_bit_cast(uint16_t, col_seg[seg]);
    rv = __builtin_bit_cast(uint16_t, row_seg[seg]);
  }
  colside[idx] = cv;
  rowside[idx] = rv;
}

template <class TE>
__global__ void pack_f64_kernel(const TE* __restrict__ E, int64_t n, int d, double* __restrict__ out) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * d) return;
  out[idx] = (double)E[idx];
}

// max_i ||e_i||^2 (FP64) -> out[0] via atomicMax on the bit pattern (non-negative doubles).
template <class TE>
__global__ void max_sqnorm_kernel(const TE* __restrict__ E, int64_t n, int d,
                                  unsigned long long* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (i < n)
    for (int k = 0; k < d; ++k) {
      double v = (double)E[i * d + k];
      s += v * v;
    }
  for (int off = 32; off > 0; off >>= 1) s = fmax(s, __shfl_xor(s, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__builtin_bit_cast(uint64_t, s));
}

// ---------------------------------------------------------------------------
// Gram tile helpers
// ---------------------------------------------------------------------------

template <int MODE>
struct GramOps;

// FP32 MFMA 32x32x2: ka floats per row, steps = ka/2.
template <>
struct GramOps<LAKER_PREC_FP32> {
  template <int NS>
  struct Operand {
    float v[NS];
  };
  template <int NS>
  __device__ static inline void load(Operand<NS>& op, const void* base, int64_t row, int ka, int h) {
    const float* p = reinterpret_cast<const float*>(base) + row * ka + h * (ka / 2);
#pragma unroll
    for (int s = 0; s < NS; ++s) op.v[s] = (s < ka / 2) ? p[s] : 0.f;
  }
  template <int NS>
  __device__ static inline floatx16 gram(const Operand<NS>& a, const Operand<NS>& b, int ka) {
    floatx16 acc = {0};
#pragma unroll
    for (int s = 0; s < NS; ++s)
      if (s < ka / 2) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.v[s], b.v[s], acc, 0, 0, 0);
    return acc;
  }
  // all NS steps, no run-time step count (NS == ka / 2)
  template <int NS>
  __device__ static inline floatx16 gram_exact(const Operand<NS>& a, const Operand<NS>& b) {
    floatx16 acc = {0};
#pragma unroll
    for (int s = 0; s < NS; ++s) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.v[s], b.v[s], acc, 0, 0, 0);
    return acc;
  }
};

// BF16 MFMA 32x32x16 on the hi/lo packed operands: chunks = ka/16.
template <>
struct GramOps<LAKER_PREC_MIXED> {
  using vec = bf16x8;
  template <int NS>
  struct Operand {
    bf16x8 v[NS];
  };
  template <int NS>
  __device__ static inline void load(Operand<NS>& op, const void* base, int64_t row, int ka, int h) {
    const uint16_t* p = reinterpret_cast<const uint16_t*>(base) + row * ka + h * 8;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      if (q < ka / 16) {
        shortx8 raw = *reinterpret_cast<const shortx8*>(p + q * 16);
        op.v[q] = __builtin_bit_cast(bf16x8, raw);
      } else {
        op.v[q] = __builtin_bit_cast(bf16x8, shortx8{0, 0, 0, 0, 0, 0, 0, 0});
      }
    }
  }
  template <int NS>
  __device__ static inline floatx16 gram(const Operand<NS>& a, const Operand<NS>& b, int ka) {
    floatx16 acc = {0};
#pragma unroll
    for (int q = 0; q < NS; ++q)
      if (q < ka / 16) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.v[q], b.v[q], acc, 0, 0, 0);
    return acc;
  }
  // all NS chunks, no run-time chunk count (NS == ka / 16)
  template <int NS>
  __device__ static inline floatx16 gram_exact(const Operand<NS>& a, const Operand<NS>& b) {
    floatx16 acc = {0};
#pragma unroll
    for (int q = 0; q < NS; ++q) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.v[q], b.v[q], acc, 0, 0, 0);
    return acc;
  }
};

// FP16 MFMA 32x32x16 on the 2-way FP16 split operands (square operator,
// k = 1): chunks = ka/16.  Internal Gram mode, not a public precision.
constexpr int GRAM_F16S = 16;
template <>
struct GramOps<GRAM_F16S> {
  using vec = halfx8;
  template <int NS>
  struct Operand {
    halfx8 v[NS];
  };
  template <int NS>
  __device__ static inline void load(Operand<NS>& op, const void* base, int64_t row, int ka, int h) {
    const uint16_t* p = reinterpret_cast<const uint16_t*>(base) + row * ka + h * 8;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      if (q < ka / 16) {
        shortx8 raw = *reinterpret_cast<const shortx8*>(p + q * 16);
        op.v[q] = __builtin_bit_cast(halfx8, raw);
      } else {
        op.v[q] = __builtin_bit_cast(halfx8, shortx8{0, 0, 0, 0, 0, 0, 0, 0});
      }
    }
  }
  template <int NS>
  __device__ static inline floatx16 gram_exact(const Operand<NS>& a, const Operand<NS>& b) {
    floatx16 acc = {0};
#pragma unroll
    for (int q = 0; q < NS; ++q) acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a.v[q], b.v[q], acc, 0, 0, 0);
    return acc;
  }
};

__device__ inline int jj_of(int v, int h) { return 8 * (v >> 2) + 4 * h + (v & 3); }

// Element transform of the exp2 argument.  DERIV: weight of the backward
// (0 where the clamp is active, i.e. the clamped entry has zero derivative).
template <bool CLAMP, bool DERIV>
__device__ inline float kexp(float t, float clampv) {
  if constexpr (DERIV) {
    if (CLAMP && t > clampv) return 0.f;
    return __builtin_amdgcn_exp2f(t);
  } else {
    if constexpr (CLAMP) t = fminf(t, clampv);
    return __builtin_amdgcn_exp2f(t);
  }
}

// ---------------------------------------------------------------------------
// Vector kernel: k <= KC columns accumulated on VALU.
// grid.x = n_row_blocks * jsplit; block = 256 (4 waves, each ITW row tiles).
// part layout: part[js][i][c], i < n_rows, c < KC.
// ---------------------------------------------------------------------------
template <int MODE, int NS, int KC, int ITW, bool CLAMP, bool DERIV>
__global__ __launch_bounds__(256) void mv_vec_kernel(
    const void* __restrict__ colpack, const void* __restrict__ rowpack, int ka, int64_t n_rows,
    int64_t n_cols, const float* __restrict__ X, int64_t ldx, int kcols,
    float* __restrict__ part, int jsplit, float clampv, const int* __restrict__ skip) {
  using G = GramOps<MODE>;
  if (skip && *skip) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, h = lane >> 5, li = lane & 31;
  const int js = blockIdx.x % jsplit;
  const int64_t ib = blockIdx.x / jsplit;
  const int64_t I0 = (ib * 4 + w) * (32 * ITW);
  if (I0 >= n_rows) return;

  typename G::template Operand<NS> B[ITW];
#pragma unroll
  for (int it = 0; it < ITW; ++it) G::load(B[it], rowpack, I0 + it * 32 + li, ka, h);

  const int64_t ntiles = (n_cols + 31) / 32;
  const int64_t t0 = ntiles * js / jsplit, t1 = ntiles * (js + 1) / jsplit;

  // two accumulators per (row tile, column) halve the FMA dependency chains
  float acc[ITW][KC][2];
#pragma unroll
  for (int it = 0; it < ITW; ++it)
#pragma unroll
    for (int c = 0; c < KC; ++c) acc[it][c][0] = acc[it][c][1] = 0.f;
  // k == 1 with a contiguous, 16-byte aligned x: the lane's 16 x values are
  // four aligned float4 runs x[J0 + 8g + 4h .. +3]
  const bool vec4 = (KC == 1) && ldx == 1 && ((reinterpret_cast<uintptr_t>(X) & 15) == 0);

  for (int64_t t = t0; t < t1; ++t) {
    const int64_t J0 = t * 32;
    typename G::template Operand<NS> A;
    G::load(A, colpack, J0 + li, ka, h);
    float xv[16][KC];
    if (J0 + 32 <= n_cols) {
      if (vec4) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const floatx4 f = *reinterpret_cast<const floatx4*>(X + J0 + 8 * g + 4 * h);
#pragma unroll
          for (int q = 0; q < 4; ++q) xv[4 * g + q][0] = f[q];
        }
      } else {
#pragma unroll
        for (int v = 0; v < 16; ++v)
#pragma unroll
          for (int c = 0; c < KC; ++c)
            xv[v][c] = (c < kcols) ? X[(J0 + jj_of(v, h)) * ldx + c] : 0.f;
      }
    } else {
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int64_t j = J0 + jj_of(v, h);
#pragma unroll
        for (int c = 0; c < KC; ++c) xv[v][c] = (j < n_cols && c < kcols) ? X[j * ldx + c] : 0.f;
      }
    }
    // all Gram tiles first (MFMA), then the exponentials (VALU): the MFMA pipe
    // works on tile it+1 while the VALU drains tile it
    floatx16 D[ITW];
#pragma unroll
    for (int it = 0; it < ITW; ++it) D[it] = G::gram(A, B[it], ka);
#pragma unroll
    for (int it = 0; it < ITW; ++it) {
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const float e = kexp<CLAMP, DERIV>(D[it][v], clampv);
#pragma unroll
        for (int c = 0; c < KC; ++c) acc[it][c][v & 1] = __builtin_fmaf(e, xv[v][c], acc[it][c][v & 1]);
      }
    }
  }
  float accs[ITW][KC];
#pragma unroll
  for (int it = 0; it < ITW; ++it)
#pragma unroll
    for (int c = 0; c < KC; ++c) accs[it][c] = acc[it][c][0] + acc[it][c][1];
#pragma unroll
  for (int it = 0; it < ITW; ++it) {
    float* accr = accs[it];
#pragma unroll
    for (int c = 0; c < KC; ++c) accr[c] += __shfl_xor(accr[c], 32, 64);
    const int64_t i = I0 + it * 32 + li;
    if (h == 0 && i < n_rows) {
#pragma unroll
      for (int c = 0; c < KC; ++c)
        if (c < kcols) part[((int64_t)js * n_rows + i) * KC + c] = accr[c];
    }
  }
}

// ---------------------------------------------------------------------------
// k = 1 kernel with the column-side operand and x staged through LDS.
// A block (4 waves x 2 row tiles = 256 rows) walks its column range in chunks
// of V1_JT column tiles; chunk c+1 is fetched from global memory into
// registers while chunk c is consumed from LDS (double buffer, one barrier
// per chunk), so the E / x reads are shared by the 4 waves and their latency
// hides behind a chunk of exp work.  LDS rows are padded to spread banks.
// NS must equal the exact step count (ka/2 for FP32, ka/16 for BF16).
// ---------------------------------------------------------------------------
constexpr int V1_JT = 4;  // column tiles per LDS chunk

template <int MODE>
struct V1Layout {
  static __host__ __device__ inline int gbytes(int ka) { return ka * (MODE == LAKER_PREC_FP32 ? 4 : 2); }
  static __host__ __device__ inline int lbytes(int ka) {
    const int g = gbytes(ka);
    return MODE == LAKER_PREC_FP32 ? ((g / 4) % 2 == 0 ? g + 4 : g) : g + 16;
  }
  static __host__ __device__ inline int chunk_bytes(int ka) { return V1_JT * 32 * (lbytes(ka) + 4); }
};

template <int MODE, int NS, bool CLAMP>
__global__ __launch_bounds__(256) void mv_vec1_kernel(
    const void* __restrict__ colpack, const void* __restrict__ rowpack, int ka, int64_t n_rows,
    int64_t n_cols, const float* __restrict__ X, int64_t ldx, double* __restrict__ part, int jsplit,
    float clampv, const int* __restrict__ skip) {
  using G = GramOps<MODE>;
  constexpr int ITW = 2;
  if (skip && *skip) return;  // uniform across the block
  // 16-byte vectors per thread per chunk: V1_JT*32 rows * row bytes / 16 / 256
  constexpr int VPT = MODE == LAKER_PREC_FP32 ? (8 * NS + 31) / 32 : NS;
  extern __shared__ __align__(16) char v1_smem[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, h = lane >> 5, li = lane & 31;
  const int js = blockIdx.x % jsplit;
  const int64_t ib = blockIdx.x / jsplit;
  const int64_t I0 = (ib * 4 + w) * (32 * ITW);
  const bool active = I0 < n_rows;
  // row bytes are fixed by NS (ka = 2*NS floats or 16*NS bf16 values)
  constexpr int gb = MODE == LAKER_PREC_FP32 ? 8 * NS : 32 * NS;
  constexpr int lb = MODE == LAKER_PREC_FP32 ? ((gb / 4) % 2 == 0 ? gb + 4 : gb) : gb + 16;
  constexpr int chunk = V1_JT * 32 * (lb + 4);
  char* const buf0 = v1_smem;
  char* const buf1 = v1_smem + chunk;

  typename G::template Operand<NS> B[ITW];
  if (active) {
#pragma unroll
    for (int it = 0; it < ITW; ++it) G::load(B[it], rowpack, I0 + it * 32 + li, ka, h);
  }
  const int64_t ntiles = (n_cols + 31) / 32;
  const int64_t t0 = ntiles * js / jsplit, t1 = ntiles * (js + 1) / jsplit;
  const char* gsrc = reinterpret_cast<const char*>(colpack);

  // The packed side holds round_up(n, 64) rows, so whole tiles are in bounds;
  // a chunk's rows start 32*t*gb bytes in, a multiple of 16 bytes.
  uint4 pre[VPT];
  float prex = 0.f;
  auto fetch = [&](int64_t t) {
    const int nt = (int)((t1 - t) < V1_JT ? (t1 - t) : V1_JT);
    const int nvec = nt * 2 * gb;  // nt*32 rows * gb bytes / 16
    (void)ka;
    const uint4* src = reinterpret_cast<const uint4*>(gsrc + t * 32 * gb);
#pragma unroll
    for (int u = 0; u < VPT; ++u) {
      const int vi = u * 256 + threadIdx.x;
      pre[u] = (vi < nvec) ? src[vi] : uint4{0u, 0u, 0u, 0u};
    }
    const int64_t j = t * 32 + threadIdx.x;
    prex = (threadIdx.x < nt * 32 && j < n_cols) ? X[j * ldx] : 0.f;
  };
  auto stash = [&](char* dst) {
    if constexpr (MODE == LAKER_PREC_FP32) {
      // rows of ka words may straddle a vector: place word by word
      constexpr int wpr = gb / 4, lwpr = lb / 4;
#pragma unroll
      for (int u = 0; u < VPT; ++u) {
        const int wi = (u * 256 + threadIdx.x) * 4;
        const uint32_t vals[4] = {pre[u].x, pre[u].y, pre[u].z, pre[u].w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int word = wi + q;
          if (word < V1_JT * 32 * wpr)
            reinterpret_cast<uint32_t*>(dst)[(word / wpr) * lwpr + word % wpr] = vals[q];
        }
      }
    } else {
      constexpr int vpr = gb / 16;  // BF16 rows are whole multiples of 32 bytes
#pragma unroll
      for (int u = 0; u < VPT; ++u) {
        const int vi = u * 256 + threadIdx.x;
        if (vi < V1_JT * 32 * vpr)
          *reinterpret_cast<uint4*>(dst + (vi / vpr) * lb + (vi % vpr) * 16) = pre[u];
      }
    }
    if (threadIdx.x < V1_JT * 32) reinterpret_cast<float*>(dst + V1_JT * 32 * lb)[threadIdx.x] = prex;
  };

  // per tile: two FP32 chains of 8 products per row; across tiles: FP64
  // (long FP32 running sums of terms that cancel would dominate the error)
  double acc64[ITW];
#pragma unroll
  for (int it = 0; it < ITW; ++it) acc64[it] = 0.0;

  // unconditional (an empty range stages zeros): the LDS stores wait for every
  // outstanding global load, the row operands included, once before the loop
  fetch(t0);
  stash(buf0);
  __syncthreads();
  bool odd = false;
  for (int64_t tc = t0; tc < t1; tc += V1_JT) {
    const bool more = tc + V1_JT < t1;
    if (more) fetch(tc + V1_JT);
    const int nt = (int)((t1 - tc) < V1_JT ? (t1 - tc) : V1_JT);
    if (active) {
      const char* cur = odd ? buf1 : buf0;
      const float* xs = reinterpret_cast<const float*>(cur + V1_JT * 32 * lb);
      for (int tt = 0; tt < nt; ++tt) {
        typename G::template Operand<NS> A;
        const char* rowp = cur + (tt * 32 + li) * lb;
        if constexpr (MODE == LAKER_PREC_FP32) {
          const float* p = reinterpret_cast<const float*>(rowp) + h * NS;
#pragma unroll
          for (int s = 0; s < NS; ++s) A.v[s] = p[s];
        } else {
#pragma unroll
          for (int q = 0; q < NS; ++q)
            A.v[q] = __builtin_bit_cast(typename G::vec, *reinterpret_cast<const shortx8*>(rowp + q * 32 + h * 16));
        }
        float xv[16];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const floatx4 f = *reinterpret_cast<const floatx4*>(xs + tt * 32 + 8 * g + 4 * h);
#pragma unroll
          for (int q = 0; q < 4; ++q) xv[4 * g + q] = f[q];
        }
        floatx16 D[ITW];
#pragma unroll
        for (int it = 0; it < ITW; ++it) D[it] = G::template gram_exact<NS>(A, B[it]);
#pragma unroll
        for (int it = 0; it < ITW; ++it) {
          float a0 = kexp<CLAMP, false>(D[it][0], clampv) * xv[0];
          float a1 = kexp<CLAMP, false>(D[it][1], clampv) * xv[1];
#pragma unroll
          for (int v = 2; v < 16; v += 2) {
            a0 = __builtin_fmaf(kexp<CLAMP, false>(D[it][v], clampv), xv[v], a0);
            a1 = __builtin_fmaf(kexp<CLAMP, false>(D[it][v + 1], clampv), xv[v + 1], a1);
          }
          acc64[it] += (double)(a0 + a1);
        }
      }
    }
    if (more) stash(odd ? buf0 : buf1);
    __syncthreads();
    odd = !odd;
  }
  if (!active) return;
#pragma unroll
  for (int it = 0; it < ITW; ++it) {
    double s = acc64[it];
    s += __shfl_xor(s, 32, 64);
    const int64_t i = I0 + it * 32 + li;
    if (h == 0 && i < n_rows) part[(int64_t)js * n_rows + i] = s;
  }
}

// ---------------------------------------------------------------------------
// Symmetric k = 1 product of the square operator.  Workgroup (b, seg) owns
// rows [SYM_RB b, SYM_RB (b + 1)) (4 waves x SYM_ITW row tiles) and the column
// tiles of segment seg, [SYM_SEG seg, SYM_SEG (seg + 1)), at or right of tile
// SYM_DT b: only tiles at or right of the diagonal SYM_RB x SYM_RB block are
// visited, so every exponential is evaluated once for the pair (i, j), (j, i).
// Each exp tile P feeds the row sums y_I += P x_J (in-lane, as in the
// non-symmetric kernel) and, right of the diagonal block, the column sums
// y_J += P^T x_I: the products of the wave's row tiles are summed in-lane,
// then a 32-lane transpose-reduce (one permlane16 swap and four DPP levels)
// runs once per column tile for all of them, a fixed-order sum over the 4
// waves in LDS follows, and one FP64 partial is stored per (row block,
// column).  Row partials: part_r[seg][i]; column partials:
// part_c[coff(b) + j - SYM_RB (b + 1)] for j >= SYM_RB (b + 1),
// coff(b) = b n - (SYM_RB / 2) b (b + 1).
// The kernel is bound by VALU issue (v_exp_f32 takes two FMA slots): per
// pair of entries one exp, one row FMA and one column FMA; the transpose-
// reduce (about 40 instructions per column tile) is what SYM_ITW amortises.
// ---------------------------------------------------------------------------
#ifndef LAKER_SYM_ITW
#define LAKER_SYM_ITW 4
#endif
constexpr int SYM_SEG = 32;                // column tiles per workgroup (1024 columns: 16..64 measured,
                                           // 32 best for load balance against row-partial traffic)
constexpr int SYM_ITW = LAKER_SYM_ITW;     // row tiles per wave
constexpr int SYM_RB = 4 * 32 * SYM_ITW;   // rows per workgroup
constexpr int SYM_DT = SYM_RB / 32;        // column tiles of the diagonal block

template <int CTRL>
__device__ inline float dpp_f(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF,
                                                               0xF, false));
}

// Sums q[v] over the 32 lanes of each half-wave.  Lane l returns the sum of
// value index v(l) = b1 + 2 b2 + 4 b3 + 8 b4 (b = bits of l & 31); lanes l and
// l ^ 1 return the same sum.  Pairing masks per level: 16 (permlane16 swap),
// 15 (row_mirror), 7 (row_half_mirror), 3 and 1 (quad_perm): together they
// span all 32 lanes, and each level keeps the half of the values selected by
// the lane bit that differs between partners.
__device__ inline float transpose_reduce16(const float (&q)[16], int lane) {
  float r[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const auto pr = __builtin_amdgcn_permlane16_swap(__builtin_bit_cast(unsigned, q[k]),
                                                     __builtin_bit_cast(unsigned, q[k + 8]), false, false);
    r[k] = __builtin_bit_cast(float, (unsigned)pr[0]) + __builtin_bit_cast(float, (unsigned)pr[1]);
  }
  const bool s3 = lane & 8, s2 = lane & 4, s1 = lane & 2;
  float s4[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float keep = s3 ? r[k + 4] : r[k], send = s3 ? r[k] : r[k + 4];
    s4[k] = keep + dpp_f<0x140>(send);  // row_mirror: lane ^ 15
  }
  float u2[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const float keep = s2 ? s4[k + 2] : s4[k], send = s2 ? s4[k] : s4[k + 2];
    u2[k] = keep + dpp_f<0x141>(send);  // row_half_mirror: lane ^ 7
  }
  const float keep = s1 ? u2[1] : u2[0], send = s1 ? u2[0] : u2[1];
  const float w = keep + dpp_f<0x1B>(send);  // quad_perm [3,2,1,0]: lane ^ 3
  return w + dpp_f<0xB1>(w);                 // quad_perm [1,0,3,2]: lane ^ 1
}

__host__ __device__ inline int64_t sym_coff(int64_t b, int64_t n) { return b * n - (SYM_RB / 2) * b * (b + 1); }

// Waves per SIMD the register allocation must allow: the FP16-split
// instance with at most two Gram steps fits 168 VGPRs without spills and runs
// three waves per SIMD instead of two; the others keep the compiler's choice.
template <int MODE, int NS>
constexpr int sym1_min_waves() { return MODE == GRAM_F16S && NS <= 2 ? 3 : 1; }

template <int MODE, int NS, bool CLAMP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sym1_min_waves<MODE, NS>()))) void mv_sym1_kernel(
    const void* __restrict__ colpack, const void* __restrict__ rowpack, int64_t n,
    const float* __restrict__ X, int64_t ldx, const int2* __restrict__ map, double* __restrict__ part_r,
    double* __restrict__ part_c, float clampv, const int* __restrict__ skip) {
  using G = GramOps<MODE>;
  constexpr int ITW = SYM_ITW;
  if (skip && *skip) return;  // uniform across the grid
  extern __shared__ __align__(16) char sym_smem[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, h = lane >> 5, li = lane & 31;
  const int2 bs = map[blockIdx.x];
  const int64_t b = bs.x;
  const int64_t I0 = b * SYM_RB + w * (32 * ITW);
  const bool active = I0 < n;
  const int64_t wrem = (n - b * SYM_RB + 32 * ITW - 1) / (32 * ITW);
  const int nact = wrem < 4 ? (int)wrem : 4;
  constexpr int gb = MODE == LAKER_PREC_FP32 ? 8 * NS : 32 * NS;
  constexpr int lb = MODE == LAKER_PREC_FP32 ? ((gb / 4) % 2 == 0 ? gb + 4 : gb) : gb + 16;
  constexpr int chunk = V1_JT * 32 * (lb + 4);
  char* const buf0 = sym_smem;
  char* const buf1 = sym_smem + chunk;
  float* const colbuf = reinterpret_cast<float*>(sym_smem + 2 * chunk);  // [2][4][V1_JT*32]

  typename G::template Operand<NS> B[ITW];
  float xr[ITW];
  if (active) {
    // packed rows exist up to round_up(n, 64): a row tile starting at or
    // beyond n (the wave's last tiles in the last row block) stays zero
#pragma unroll
    for (int it = 0; it < ITW; ++it) {
      if (I0 + it * 32 < n)
        G::load(B[it], rowpack, I0 + it * 32 + li, NS * (MODE == LAKER_PREC_FP32 ? 2 : 16), h);
      else
        B[it] = typename G::template Operand<NS>{};
      const int64_t i = I0 + it * 32 + li;
      xr[it] = i < n ? X[i * ldx] : 0.f;
    }
  }
  const int64_t ntiles = (n + 31) / 32;
  const int64_t tdiag = SYM_DT * (b + 1);  // first tile right of the diagonal block
  // segments are aligned to absolute column tiles, so the workgroups of one
  // segment (all row blocks) share its columns in L2
  const int64_t t0 = (SYM_DT * b > (int64_t)bs.y * SYM_SEG) ? SYM_DT * b : (int64_t)bs.y * SYM_SEG;
  const int64_t t1 = ((int64_t)bs.y * SYM_SEG + SYM_SEG < ntiles) ? (int64_t)bs.y * SYM_SEG + SYM_SEG : ntiles;
  const char* gsrc = reinterpret_cast<const char*>(colpack);
  constexpr int VPT = MODE == LAKER_PREC_FP32 ? (8 * NS + 31) / 32 : NS;
  uint4 pre[VPT];
  float prex = 0.f;
  auto fetch = [&](int64_t t) {
    const int nt = (int)((t1 - t) < V1_JT ? (t1 - t) : V1_JT);
    const int nvec = nt * 2 * gb;
    const uint4* src = reinterpret_cast<const uint4*>(gsrc + t * 32 * gb);
#pragma unroll
    for (int u = 0; u < VPT; ++u) {
      const int vi = u * 256 + threadIdx.x;
      pre[u] = (vi < nvec) ? src[vi] : uint4{0u, 0u, 0u, 0u};
    }
    const int64_t j = t * 32 + threadIdx.x;
    prex = (threadIdx.x < nt * 32 && j < n) ? X[j * ldx] : 0.f;
  };
  auto stash = [&](char* dst) {
    if constexpr (MODE == LAKER_PREC_FP32) {
      constexpr int wpr = gb / 4, lwpr = lb / 4;
#pragma unroll
      for (int u = 0; u < VPT; ++u) {
        const int wi = (u * 256 + threadIdx.x) * 4;
        const uint32_t vals[4] = {pre[u].x, pre[u].y, pre[u].z, pre[u].w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int word = wi + q;
          if (word < V1_JT * 32 * wpr)
            reinterpret_cast<uint32_t*>(dst)[(word / wpr) * lwpr + word % wpr] = vals[q];
        }
      }
    } else {
      constexpr int vpr = gb / 16;
#pragma unroll
      for (int u = 0; u < VPT; ++u) {
        const int vi = u * 256 + threadIdx.x;
        if (vi < V1_JT * 32 * vpr)
          *reinterpret_cast<uint4*>(dst + (vi / vpr) * lb + (vi % vpr) * 16) = pre[u];
      }
    }
    if (threadIdx.x < V1_JT * 32) reinterpret_cast<float*>(dst + V1_JT * 32 * lb)[threadIdx.x] = prex;
  };
  // lane -> column of its transpose-reduce result (see transpose_reduce16)
  const int vq = ((lane >> 1) & 1) + 2 * ((lane >> 2) & 1) + 4 * ((lane >> 3) & 1) + 8 * ((lane >> 4) & 1);
  const int jq = jj_of(vq, h);

  double acc64[ITW];
#pragma unroll
  for (int it = 0; it < ITW; ++it) acc64[it] = 0.0;
  fetch(t0);
  stash(buf0);
  __syncthreads();
  bool odd = false;
  for (int64_t tc = t0; tc < t1; tc += V1_JT) {
    const bool more = tc + V1_JT < t1;
    if (more) fetch(tc + V1_JT);
    const int nt = (int)((t1 - tc) < V1_JT ? (t1 - tc) : V1_JT);
    float* cb = colbuf + (odd ? 4 * V1_JT * 32 : 0);
    if (active) {
      const char* cur = odd ? buf1 : buf0;
      const float* xs = reinterpret_cast<const float*>(cur + V1_JT * 32 * lb);
      for (int tt = 0; tt < nt; ++tt) {
        typename G::template Operand<NS> A;
        const char* rowp = cur + (tt * 32 + li) * lb;
        if constexpr (MODE == LAKER_PREC_FP32) {
          const float* p = reinterpret_cast<const float*>(rowp) + h * NS;
#pragma unroll
          for (int s2 = 0; s2 < NS; ++s2) A.v[s2] = p[s2];
        } else {
#pragma unroll
          for (int q = 0; q < NS; ++q)
            A.v[q] = __builtin_bit_cast(typename G::vec, *reinterpret_cast<const shortx8*>(rowp + q * 32 + h * 16));
        }
        float xv[16];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const floatx4 f = *reinterpret_cast<const floatx4*>(xs + tt * 32 + 8 * g + 4 * h);
#pragma unroll
          for (int q = 0; q < 4; ++q) xv[4 * g + q] = f[q];
        }
        // row tiles in pairs: two Gram tiles live at a time
        float q[16];
#pragma unroll
        for (int p = 0; p < ITW / 2; ++p) {
          floatx16 D[2];
#pragma unroll
          for (int u = 0; u < 2; ++u) D[u] = G::template gram_exact<NS>(A, B[2 * p + u]);
#pragma unroll
          for (int u = 0; u < 2; ++u) {
#pragma unroll
            for (int v = 0; v < 16; ++v) D[u][v] = kexp<CLAMP, false>(D[u][v], clampv);
            float a0 = D[u][0] * xv[0], a1 = D[u][1] * xv[1];
#pragma unroll
            for (int v = 2; v < 16; v += 2) {
              a0 = __builtin_fmaf(D[u][v], xv[v], a0);
              a1 = __builtin_fmaf(D[u][v + 1], xv[v + 1], a1);
            }
            acc64[2 * p + u] += (double)(a0 + a1);
          }
          // column-side products of the row tiles, summed in-lane
#pragma unroll
          for (int v = 0; v < 16; ++v)
            q[v] = __builtin_fmaf(D[1][v], xr[2 * p + 1],
                                  p == 0 ? D[0][v] * xr[0] : __builtin_fmaf(D[0][v], xr[2 * p], q[v]));
        }
        if (tc + tt >= tdiag) {  // uniform across the block
          const float cs = transpose_reduce16(q, lane);
          if ((lane & 1) == 0) cb[w * (V1_JT * 32) + tt * 32 + jq] = cs;
        }
      }
    }
    if (more) stash(odd ? buf0 : buf1);
    __syncthreads();
    // column partials of this chunk: fixed-order sum over the active waves
    if (threadIdx.x < nt * 32) {
      const int tt = threadIdx.x >> 5;
      const int64_t t = tc + tt;
      const int64_t j = t * 32 + (threadIdx.x & 31);
      if (t >= tdiag && j < n) {
        double v = 0.0;
        for (int ww = 0; ww < nact; ++ww) v += (double)cb[ww * (V1_JT * 32) + threadIdx.x];
        part_c[sym_coff(b, n) + j - SYM_RB * (b + 1)] = v;
      }
    }
    odd = !odd;
  }
  if (!active) return;
#pragma unroll
  for (int it = 0; it < ITW; ++it) {
    double s = acc64[it];
    s += __shfl_xor(s, 32, 64);
    const int64_t i = I0 + it * 32 + li;
    if (h == 0 && i < n) part_r[(int64_t)bs.y * n + i] = s;
  }
}

// y[i] = lam xr[i] (+ y[i]) + sum_seg part_r[seg][i] + sum_{b < i/SYM_RB} part_c[coff(b) + i - SYM_RB (b+1)]
// dots (optional, block = 256 rows): dots[q * gridDim.x + block] = the
// block's sum of x.y (q = 0), x.x (q = 1) and y.y (q = 2) over its rows, with
// x = Xr and y the stored Y, in the order and with the block reduction of the
// PCG's column-dot kernel for 256-row blocks (bit-identical partials).
__global__ void mv_sym_reduce_kernel(const double* __restrict__ part_r, const double* __restrict__ part_c,
                                     int64_t n, double lam, const float* __restrict__ Xr, int64_t ldxr,
                                     float* __restrict__ Y, int64_t ldy, int accumulate,
                                     const int* __restrict__ skip, double* __restrict__ dots) {
  if (skip && *skip) return;  // uniform over the grid
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double xd = 0.0, yd = 0.0;
  if (i < n) {
    const int64_t ntiles = (n + 31) / 32;
    const int64_t b = i / SYM_RB;
    const int64_t nseg = (ntiles + SYM_SEG - 1) / SYM_SEG;
    // eight independent chains (fixed assignment) keep the loads in flight
    double a8[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double tail = 0.0;
    int64_t sg = (SYM_DT * b) / SYM_SEG;  // first segment of row block b
    for (; sg + 8 <= nseg; sg += 8) {
#pragma unroll
      for (int u = 0; u < 8; ++u) a8[u] += part_r[(sg + u) * n + i];
    }
    for (; sg < nseg; ++sg) tail += part_r[sg * n + i];
    int64_t bb = 0;
    for (; bb + 8 <= b; bb += 8) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int64_t q = bb + u;
        a8[u] += part_c[sym_coff(q, n) + i - SYM_RB * (q + 1)];
      }
    }
    for (; bb < b; ++bb) tail += part_c[sym_coff(bb, n) + i - SYM_RB * (bb + 1)];
    double s = (((a8[0] + a8[1]) + (a8[2] + a8[3])) + ((a8[4] + a8[5]) + (a8[6] + a8[7]))) + tail;
    if (Xr) s += lam * (double)Xr[i * ldxr];
    if (accumulate) s += (double)Y[i * ldy];
    const float y = (float)s;
    Y[i * ldy] = y;
    if (dots) {
      xd = (double)Xr[i * ldxr];
      yd = (double)y;
    }
  }
  if (dots) {  // uniform: every thread of the block takes part
    __shared__ double red[16];
    const double v[3] = {xd * yd, xd * xd, yd * yd};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double t = block_sum(v[q], red);
      if (threadIdx.x == 0) dots[(int64_t)q * gridDim.x + blockIdx.x] = t;
    }
  }
}

// ---------------------------------------------------------------------------
// Multi-RHS kernel: G.X on MFMA.  Each wave owns ITW row tiles and NCC chunks
// of 32 columns.  X tiles (32 rows x 32*NCC cols) are staged in LDS per
// column tile and shared by the 4 waves of a block.
// GX = LAKER_PREC_FP32: 16 x mfma_f32_32x32x2f32 per 32x32 block.
// GX = LAKER_PREC_BF16: P and X rounded to bf16, 2 x mfma 32x32x16.
// GX = LAKER_PREC_MIXED: P and X split hi/lo, 3 products (hi.hi + lo.hi + hi.lo).
// part layout: part[js][i][c], c < ldp (= k).
// ---------------------------------------------------------------------------
template <int MODE, int NS, int GX, int NCC, int ITW, bool CLAMP, bool DERIV>
__global__ __launch_bounds__(256) void mv_multi_kernel(
    const void* __restrict__ colpack, const void* __restrict__ rowpack, int ka, int64_t n_rows,
    int64_t n_cols, const float* __restrict__ X, int64_t ldx, int kcols, int c_base,
    float* __restrict__ part, int64_t ldp, int jsplit, float clampv) {
  using G = GramOps<MODE>;
  constexpr int KW = 32 * NCC;
  __shared__ float xs[32][KW + 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, h = lane >> 5, li = lane & 31;
  const int js = blockIdx.x % jsplit;
  const int64_t ib = blockIdx.x / jsplit;
  const int64_t I0 = (ib * 4 + w) * (32 * ITW);
  const bool active = I0 < n_rows;

  typename G::template Operand<NS> B[ITW];
  if (active) {
#pragma unroll
    for (int it = 0; it < ITW; ++it) G::load(B[it], rowpack, I0 + it * 32 + li, ka, h);
  }
  const int64_t ntiles = (n_cols + 31) / 32;
  const int64_t t0 = ntiles * js / jsplit, t1 = ntiles * (js + 1) / jsplit;

  floatx16 acc[ITW][NCC];
#pragma unroll
  for (int it = 0; it < ITW; ++it)
#pragma unroll
    for (int q = 0; q < NCC; ++q) acc[it][q] = floatx16{0};

  for (int64_t t = t0; t < t1; ++t) {
    const int64_t J0 = t * 32;
    __syncthreads();
    for (int e = threadIdx.x; e < 32 * KW; e += 256) {
      const int r = e / KW, c = e % KW;
      const int64_t j = J0 + r;
      const int cc = c_base + c;
      xs[r][c] = (j < n_cols && cc < kcols) ? X[j * ldx + cc] : 0.f;
    }
    __syncthreads();
    if (!active) continue;
    typename G::template Operand<NS> A;
    G::load(A, colpack, J0 + li, ka, h);
#pragma unroll
    for (int it = 0; it < ITW; ++it) {
      floatx16 D = G::gram(A, B[it], ka);
      float P[16];
#pragma unroll
      for (int v = 0; v < 16; ++v) P[v] = kexp<CLAMP, DERIV>(D[v], clampv);
      if constexpr (GX == LAKER_PREC_FP32) {
#pragma unroll
        for (int q = 0; q < NCC; ++q)
#pragma unroll
          for (int v = 0; v < 16; ++v)
            acc[it][q] = __builtin_amdgcn_mfma_f32_32x32x2f32(P[v], xs[jj_of(v, h)][q * 32 + li],
                                                              acc[it][q], 0, 0, 0);
      } else {
        // K ordering of chunk u (u=0,1): k = 8h + t  <->  j = jj(8u + t, h)
        bf16x8 Ph[2], Pl[2];
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
          for (int t2 = 0; t2 < 8; ++t2) {
            const float p = P[8 * u + t2];
            const uint16_t hb = f32_to_bf16_bits(p);
            Ph[u][t2] = __builtin_bit_cast(__bf16, hb);
            if constexpr (GX == LAKER_PREC_MIXED)
              Pl[u][t2] = __builtin_bit_cast(__bf16, f32_to_bf16_bits(p - bf16_bits_to_f32(hb)));
          }
#pragma unroll
        for (int q = 0; q < NCC; ++q) {
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            bf16x8 Xh, Xl;
#pragma unroll
            for (int t2 = 0; t2 < 8; ++t2) {
              const float xvv = xs[jj_of(8 * u + t2, h)][q * 32 + li];
              const uint16_t hb = f32_to_bf16_bits(xvv);
              Xh[t2] = __builtin_bit_cast(__bf16, hb);
              if constexpr (GX == LAKER_PREC_MIXED)
                Xl[t2] = __builtin_bit_cast(__bf16, f32_to_bf16_bits(xvv - bf16_bits_to_f32(hb)));
            }
            acc[it][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ph[u], Xh, acc[it][q], 0, 0, 0);
            if constexpr (GX == LAKER_PREC_MIXED) {
              acc[it][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Pl[u], Xh, acc[it][q], 0, 0, 0);
              acc[it][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ph[u], Xl, acc[it][q], 0, 0, 0);
            }
          }
        }
      }
    }
  }
  if (!active) return;
  // acc[it][q][r]: row i = I0 + it*32 + 8*(r/4) + 4*h + r%4 ; column c = c_base + q*32 + li
#pragma unroll
  for (int it = 0; it < ITW; ++it)
#pragma unroll
    for (int q = 0; q < NCC; ++q)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t i = I0 + it * 32 + jj_of(r, h);
        const int c = c_base + q * 32 + li;
        if (i < n_rows && c < kcols) part[((int64_t)js * n_rows + i) * ldp + c] = acc[it][q][r];
      }
}

// ---------------------------------------------------------------------------
// Multi-RHS kernel with BF16 right-hand sides (GX = LAKER_PREC_BF16: P and X
// rounded to bf16; GX = LAKER_PREC_MIXED: both split hi/lo, 3 products).
// X is packed beforehand (pack_x_kernel) straight into the MFMA B-operand
// layout, so a lane fetches each operand with one 16-byte load; the exp tile
// P is converted once per row tile and reused for all 32 nq columns of a pass
// (nq = 1, 2 or 4 column chunks of 32).
// Xp index: ((t * nq + q) * 2 + u) * 64 + lane, 8 bf16 values
//   X[32 t + jj(8u + v, lane/32)][c_base + 32 q + lane%32], v = 0..7.
// ---------------------------------------------------------------------------
__global__ void pack_x_kernel(const float* __restrict__ X, int64_t ldx, int64_t n_cols, int kcols,
                              int c_base, int64_t ntiles, int nq, bf16x8* __restrict__ xh,
                              bf16x8* __restrict__ xl, bf16x8* __restrict__ xm) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ntiles * 128 * nq) return;
  const int lane = (int)(idx & 63), u = (int)((idx >> 6) & 1);
  const int q = (int)((idx >> 7) % nq);
  const int64_t t = (idx >> 7) / nq;
  const int h = lane >> 5, li = lane & 31;
  const int c = c_base + 32 * q + li;
  // x = hi + lo (hi/lo mode) or x = hi + mid + lo3 (xm given: 3-way split)
  bf16x8 hi, lo, mid;
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    const int64_t j = t * 32 + jj_of(8 * u + v, h);
    const float x = (j < n_cols && c < kcols) ? X[j * ldx + c] : 0.f;
    const uint16_t hb = f32_to_bf16_bits(x);
    const float r1 = x - bf16_bits_to_f32(hb);
    const uint16_t mb = f32_to_bf16_bits(r1);
    hi[v] = __builtin_bit_cast(__bf16, hb);
    lo[v] = __builtin_bit_cast(__bf16, xm ? f32_to_bf16_bits(r1 - bf16_bits_to_f32(mb)) : mb);
    mid[v] = __builtin_bit_cast(__bf16, mb);
  }
  xh[idx] = hi;
  if (xl) xl[idx] = lo;
  if (xm) xm[idx] = mid;
}

constexpr int GX_SPLIT3 = 5;  // exp tile and X both split 3-way, 6 BF16 products (~2^-24)
constexpr int GX_PLXH = 6;    // exp tile split hi/lo, X exactly BF16: 2 products (hi.X + lo.X)

template <int GX>
struct M2Parts {
  static constexpr int value = GX == GX_SPLIT3 ? 3 : (GX == LAKER_PREC_MIXED ? 2 : 1);
};
// LDS bytes of one chunk of JT column tiles: operand rows (padded) + X operands
__host__ __device__ constexpr int m2_chunk_bytes(int ns, int nparts, int ncc, int jt) {
  return jt * 32 * (32 * ns + 16) + jt * nparts * ncc * 2 * 64 * 16;
}

// Column operand rows and packed X operands of JT column tiles are staged
// global -> registers -> LDS one chunk ahead (double buffer, one barrier per
// chunk) and shared by the block's 4 waves.  GM: Gram operands, the BF16
// 3-way split (LAKER_PREC_MIXED) or the FP16 2-way split (GRAM_F16S, square
// operator only).
template <int GM, int NS, int GX, int NCC, int JT, int ITW, bool CLAMP>
__global__ __launch_bounds__(256) void mv_multi2_kernel(
    const void* __restrict__ colpack, const void* __restrict__ rowpack, int ka, int64_t n_rows,
    int64_t n_cols, const bf16x8* __restrict__ xh, const bf16x8* __restrict__ xl,
    const bf16x8* __restrict__ xm, int kcols, int c_base, float* __restrict__ part, int64_t ldp,
    int jsplit, float clampv) {
  using G = GramOps<GM>;
  constexpr int NP = M2Parts<GX>::value;
  constexpr int gb = 32 * NS, lb = gb + 16, vpr = gb / 16;
  constexpr int VA = JT * 32 * vpr;           // operand-row vectors per chunk
  constexpr int XT = NP * NCC * 2 * 64;       // X vectors per tile
  constexpr int VX = JT * XT;
  constexpr int VPT = (VA + VX + 255) / 256;  // vectors per thread per chunk
  constexpr int CHUNK = m2_chunk_bytes(NS, NP, NCC, JT);
  extern __shared__ __align__(16) char m2_smem[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, h = lane >> 5, li = lane & 31;
  const int js = blockIdx.x % jsplit;
  const int64_t ib = blockIdx.x / jsplit;
  const int64_t I0 = (ib * 4 + w) * (32 * ITW);
  const bool active = I0 < n_rows;
  (void)ka;
  typename G::template Operand<NS> B[ITW];
  if (active) {
#pragma unroll
    for (int it = 0; it < ITW; ++it) G::load(B[it], rowpack, I0 + it * 32 + li, NS * 16, h);
  }
  const int64_t ntiles = (n_cols + 31) / 32;
  const int64_t t0 = ntiles * js / jsplit, t1 = ntiles * (js + 1) / jsplit;
  const char* gsrc = reinterpret_cast<const char*>(colpack);
  const uint4* xsrc[3] = {reinterpret_cast<const uint4*>(xh), reinterpret_cast<const uint4*>(xl),
                          reinterpret_cast<const uint4*>(xm)};
  uint4 pre[VPT];
  auto fetch = [&](int64_t t) {
    const int nt = (int)((t1 - t) < JT ? (t1 - t) : JT);
#pragma unroll
    for (int u2 = 0; u2 < VPT; ++u2) {
      const int vi = u2 * 256 + threadIdx.x;
      uint4 v = {0u, 0u, 0u, 0u};
      if (vi < VA) {
        if (vi < nt * 32 * vpr) v = reinterpret_cast<const uint4*>(gsrc + t * 32 * gb)[vi];
      } else if (vi < VA + VX) {
        const int xi = vi - VA;
        const int tt = xi / XT, rem = xi % XT;
        const int pp = rem / (NCC * 2 * 64), off = rem % (NCC * 2 * 64);  // off = (q*2+u)*64 + lane
        if (tt < nt) v = xsrc[pp][(t + tt) * (NCC * 128) + off];
      }
      pre[u2] = v;
    }
  };
  auto stash = [&](char* dst) {
#pragma unroll
    for (int u2 = 0; u2 < VPT; ++u2) {
      const int vi = u2 * 256 + threadIdx.x;
      if (vi < VA) {
        *reinterpret_cast<uint4*>(dst + (vi / vpr) * lb + (vi % vpr) * 16) = pre[u2];
      } else if (vi < VA + VX) {
        *reinterpret_cast<uint4*>(dst + JT * 32 * lb + (vi - VA) * 16) = pre[u2];
      }
    }
  };
  floatx16 acc[ITW][NCC];
#pragma unroll
  for (int it = 0; it < ITW; ++it)
#pragma unroll
    for (int q = 0; q < NCC; ++q) acc[it][q] = floatx16{0};

  fetch(t0);
  stash(m2_smem);
  __syncthreads();
  bool odd = false;
  for (int64_t tc = t0; tc < t1; tc += JT) {
    const bool more = tc + JT < t1;
    if (more) fetch(tc + JT);
    const int nt = (int)((t1 - tc) < JT ? (t1 - tc) : JT);
    if (active) {
      const char* cur = m2_smem + (odd ? CHUNK : 0);
      const uint4* xs = reinterpret_cast<const uint4*>(cur + JT * 32 * lb);
      for (int tt = 0; tt < nt; ++tt) {
        typename G::template Operand<NS> A;
        const char* rowp = cur + (tt * 32 + li) * lb;
#pragma unroll
        for (int q = 0; q < NS; ++q)
          A.v[q] = __builtin_bit_cast(typename G::vec, *reinterpret_cast<const shortx8*>(rowp + q * 32 + h * 16));
        bf16x8 Xh[NCC][2], Xl[NCC][2], Xm[NCC][2];
#pragma unroll
        for (int q = 0; q < NCC; ++q)
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            const int o = tt * XT + (q * 2 + u) * 64 + lane;
            Xh[q][u] = __builtin_bit_cast(bf16x8, xs[o]);
            if constexpr (NP >= 2) Xl[q][u] = __builtin_bit_cast(bf16x8, xs[o + NCC * 2 * 64]);
            if constexpr (NP >= 3) Xm[q][u] = __builtin_bit_cast(bf16x8, xs[o + 2 * NCC * 2 * 64]);
          }
        floatx16 D[ITW];
#pragma unroll
        for (int it = 0; it < ITW; ++it) D[it] = G::template gram_exact<NS>(A, B[it]);
#pragma unroll
        for (int it = 0; it < ITW; ++it) {
          // K ordering of chunk u: k = 8h + v  <->  j = jj(8u + v, h)
          bf16x8 Ph[2], Pl[2], Pm[2];
#pragma unroll
          for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int v = 0; v < 8; ++v) {
              const float pv = kexp<CLAMP, false>(D[it][8 * u + v], clampv);
              const uint16_t hb = f32_to_bf16_bits(pv);
              Ph[u][v] = __builtin_bit_cast(__bf16, hb);
              if constexpr (GX == LAKER_PREC_MIXED || GX == GX_PLXH)
                Pl[u][v] = __builtin_bit_cast(__bf16, f32_to_bf16_bits(pv - bf16_bits_to_f32(hb)));
              if constexpr (GX == GX_SPLIT3) {
                const float r1 = pv - bf16_bits_to_f32(hb);
                const uint16_t mb = f32_to_bf16_bits(r1);
                Pm[u][v] = __builtin_bit_cast(__bf16, mb);
                Pl[u][v] = __builtin_bit_cast(__bf16, f32_to_bf16_bits(r1 - bf16_bits_to_f32(mb)));
              }
            }
#pragma unroll
          for (int q = 0; q < NCC; ++q)
#pragma unroll
            for (int u = 0; u < 2; ++u) {
              if constexpr (GX == GX_SPLIT3) {
                // smallest terms first: lo.hi + hi.lo + mid.mid + mid.hi + hi.mid + hi.hi
                acc[it][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Pl[u], Xh[q][u], acc[it][q], 0, 0, 0);
                acc[it][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ph[u], Xl[q][u], acc[it][q], 0, 0, 0);
                acc[it][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Pm[u], Xm[q][u], acc[it][q], 0, 0, 0);
                acc[it][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Pm[u], Xh[q][u], acc[it][q], 0, 0, 0);
                acc[it][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ph[u], Xm[q][u], acc[it][q], 0, 0, 0);
              }
              if constexpr (GX == GX_PLXH)
                acc[it][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Pl[u], Xh[q][u], acc[it][q], 0, 0, 0);
              acc[it][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ph[u], Xh[q][u], acc[it][q], 0, 0, 0);
              if constexpr (GX == LAKER_PREC_MIXED) {
                acc[it][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Pl[u], Xh[q][u], acc[it][q], 0, 0, 0);
                acc[it][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ph[u], Xl[q][u], acc[it][q], 0, 0, 0);
              }
            }
        }
      }
    }
    if (more) stash(m2_smem + (odd ? 0 : CHUNK));
    __syncthreads();
    odd = !odd;
  }
  if (!active) return;
  // acc[it][q][r]: row i = I0 + it*32 + jj(r, h) ; column c = c_base + q*32 + li
#pragma unroll
  for (int it = 0; it < ITW; ++it)
#pragma unroll
    for (int q = 0; q < NCC; ++q)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t i = I0 + it * 32 + jj_of(r, h);
        const int c = c_base + q * 32 + li;
        if (i < n_rows && c < kcols) part[((int64_t)js * n_rows + i) * ldp + c] = acc[it][q][r];
      }
}

// ---------------------------------------------------------------------------
// FP64 kernel: one thread per output row, column rows staged in LDS.
// grid.x = ceil(n_rows/256) * jsplit.  XT: type of X (float or double).
// ---------------------------------------------------------------------------
template <int KC, int DMAX, bool CLAMP, class XT>
__global__ __launch_bounds__(256) void mv_f64_kernel(const double* __restrict__ Ecol,
                                                      const double* __restrict__ Erow, int d,
                                                      int64_t n_rows, int64_t n_cols, double scale,
                                                      double clampv, const XT* __restrict__ X,
                                                      int64_t ldx, int kcols, int c_base,
                                                      double* __restrict__ part, int64_t ldp,
                                                      int jsplit) {
  __shared__ double es[64][DMAX + 1];
  __shared__ double xs[64][KC];
  const int js = blockIdx.x % jsplit;
  const int64_t ib = blockIdx.x / jsplit;
  const int64_t i = ib * 256 + threadIdx.x;
  double ei[DMAX];
#pragma unroll
  for (int q = 0; q < DMAX; ++q) ei[q] = (i < n_rows && q < d) ? Erow[i * d + q] * scale : 0.0;
  double acc[KC];
#pragma unroll
  for (int c = 0; c < KC; ++c) acc[c] = 0.0;
  const int64_t ntiles = (n_cols + 63) / 64;
  const int64_t t0 = ntiles * js / jsplit, t1 = ntiles * (js + 1) / jsplit;
  for (int64_t t = t0; t < t1; ++t) {
    const int64_t J0 = t * 64;
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * DMAX; e += 256) {
      const int r = e / DMAX, q = e % DMAX;
      const int64_t j = J0 + r;
      es[r][q] = (j < n_cols && q < d) ? Ecol[j * d + q] : 0.0;
    }
    for (int e = threadIdx.x; e < 64 * KC; e += 256) {
      const int r = e / KC, c = e % KC;
      const int64_t j = J0 + r;
      xs[r][c] = (j < n_cols && c_base + c < kcols) ? (double)X[j * ldx + c_base + c] : 0.0;
    }
    __syncthreads();
    const int rmax = (n_cols - J0) < 64 ? (int)(n_cols - J0) : 64;
    for (int r = 0; r < rmax; ++r) {
      double g = 0.0;
#pragma unroll
      for (int q = 0; q < DMAX; ++q) g = fma(ei[q], es[r][q], g);
      if constexpr (CLAMP) g = fmin(g, clampv);
      const double ev = exp(g);
#pragma unroll
      for (int c = 0; c < KC; ++c) acc[c] = fma(ev, xs[r][c], acc[c]);
    }
  }
  if (i < n_rows) {
#pragma unroll
    for (int c = 0; c < KC; ++c)
      if (c_base + c < kcols) part[((int64_t)js * n_rows + i) * ldp + c_base + c] = acc[c];
  }
}

// ---------------------------------------------------------------------------
// FP64 k = 1 kernel with the Gram tiles on the FP64 MFMA (16x16x4).
// The row operand is pre-scaled by 256 * s * log2(e), so the MFMA returns
// t' = 256 t with exp(s g) = 2^t, and the exponential is a table-driven exp2
// (Tang's method): k = rint(t'), r = t' - k (exact, |r| <= 1/2),
// 2^t = 2^(k >> 8) * T[k & 255] * 2^(r/256), T[m] = 2^(m/256) in LDS and
// 2^(r/256) by its Taylor series to degree 5 (truncation < 1e-20).
// Each wave owns ITW = 4 groups of 16 rows; lane l accumulates row
// i = I0 + 16 it + l%16 over the columns j = J0 + l/16 + 4 r (r < 4) of each
// 16-column tile; the four lane groups are summed at the end in fixed order.
// ---------------------------------------------------------------------------
__device__ inline double exp2_t256(double tp, const double* __restrict__ tab) {
  constexpr double c = 0.69314718055994530942 / 256.0;
  constexpr double p1 = c, p2 = c * c / 2.0, p3 = c * c * c / 6.0, p4 = c * c * c * c / 24.0,
                   p5 = c * c * c * c * c / 120.0;
  const double kd = __builtin_rint(tp);
  const double r = tp - kd;
  const int ki = (int)kd;
  double p = __builtin_fma(p5, r, p4);
  p = __builtin_fma(p, r, p3);
  p = __builtin_fma(p, r, p2);
  p = __builtin_fma(p, r, p1);
  p = __builtin_fma(p, r, 1.0);
  return __builtin_ldexp(tab[ki & 255] * p, ki >> 8);
}

template <int STEPS, int KC, int ITW, bool CLAMP, class XT>
__global__ __launch_bounds__(256) void mv_f64m_kernel(const double* __restrict__ Ecol,
                                                       const double* __restrict__ Erow, int d,
                                                       int64_t n_rows, int64_t n_cols, double rscale,
                                                       double clampv, const XT* __restrict__ X,
                                                       int64_t ldx, int kcols, int c_base,
                                                       double* __restrict__ part, int64_t ldp,
                                                       int jsplit) {
  __shared__ double tab[256];
  tab[threadIdx.x] = exp2((double)threadIdx.x * (1.0 / 256.0));
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int js = blockIdx.x % jsplit;
  const int64_t ib = blockIdx.x / jsplit;
  const int64_t I0 = (ib * 4 + w) * (16 * ITW);
  if (I0 >= n_rows) return;
  double b[ITW][STEPS];
#pragma unroll
  for (int it = 0; it < ITW; ++it)
#pragma unroll
    for (int s = 0; s < STEPS; ++s) {
      const int64_t i = I0 + 16 * it + m;
      const int k = kq + 4 * s;
      b[it][s] = (i < n_rows && k < d) ? Erow[i * d + k] * rscale : 0.0;
    }
  const int64_t ntiles = (n_cols + 15) / 16;
  const int64_t t0 = ntiles * js / jsplit, t1 = ntiles * (js + 1) / jsplit;
  double acc[ITW][KC];
#pragma unroll
  for (int it = 0; it < ITW; ++it)
#pragma unroll
    for (int c = 0; c < KC; ++c) acc[it][c] = 0.0;
  for (int64_t t = t0; t < t1; ++t) {
    const int64_t J0 = t * 16;
    double a[STEPS];
#pragma unroll
    for (int s = 0; s < STEPS; ++s) {
      const int64_t j = J0 + m;
      const int k = kq + 4 * s;
      a[s] = (j < n_cols && k < d) ? Ecol[j * d + k] : 0.0;
    }
    double xv[4][KC];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t j = J0 + kq + 4 * r;
#pragma unroll
      for (int c = 0; c < KC; ++c)
        xv[r][c] = (j < n_cols && c_base + c < kcols) ? (double)X[j * ldx + c_base + c] : 0.0;
    }
#pragma unroll
    for (int it = 0; it < ITW; ++it) {
      doublex4 D = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < STEPS; ++s) D = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[it][s], D, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double tp = D[r];
        if constexpr (CLAMP) tp = fmin(tp, clampv);
        const double ev = exp2_t256(tp, tab);
#pragma unroll
        for (int c = 0; c < KC; ++c) acc[it][c] = __builtin_fma(ev, xv[r][c], acc[it][c]);
      }
    }
  }
#pragma unroll
  for (int it = 0; it < ITW; ++it) {
    const int64_t i = I0 + 16 * it + m;
#pragma unroll
    for (int c = 0; c < KC; ++c) {
      double v = acc[it][c];
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      if (kq == 0 && i < n_rows && c_base + c < kcols) part[((int64_t)js * n_rows + i) * ldp + c_base + c] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Symmetric FP64 k = 1 product of the square operator (the refinement
// residual).  Workgroup (b, seg) owns rows [S64_RB b, S64_RB (b + 1)) (4 waves
// x 4 groups of 16 rows) and the 16-column tiles of segment seg at or right of
// tile S64_DT b, so each exponential is evaluated once for the pair (i, j),
// (j, i).  Gram on v_mfma_f64_16x16x4f64, exponential by exp2_t256.  Lane
// (m, kq) holds P[j = J0 + kq + 4 r][i = I0 + 16 it + m]: row sums (weights
// x_j) accumulate in-lane; column sums (weights x_i) are summed over the row
// groups in-lane, over the 16 lanes m by DPP within the row, over the 4 waves
// in LDS (fixed order), and stored as one FP64 partial per (row block, column)
// right of the diagonal block: part_c[coff(b) + j - S64_RB (b + 1)],
// coff(b) = b n - (S64_RB / 2) b (b + 1).  Row partials: part_r[seg][i].
// ---------------------------------------------------------------------------
constexpr int S64_RB = 256;             // rows per workgroup
constexpr int S64_SEG = 128;            // 16-column tiles per segment
constexpr int S64_DT = S64_RB / 16;     // tiles of the diagonal block
constexpr int S64_CH = 8;               // tiles per column-sum chunk (one barrier pair)

__host__ __device__ inline int64_t s64_coff(int64_t b, int64_t n) { return b * n - (S64_RB / 2) * b * (b + 1); }

template <int CTRL>
__device__ inline double dpp_d(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Sum over the 16 lanes of a DPP row (every lane receives it): quad_perm
// [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror.
__device__ inline double row_sum16(double v) {
  v += dpp_d<0xB1>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0x141>(v);
  v += dpp_d<0x140>(v);
  return v;
}

// Transpose-reduce of four values over the 16 lanes of a DPP row: lane m ends
// with the 16-lane sum of q[a(m)], a = {0,2,1,3,3,1,2,0,0,2,1,3,3,1,2,0}[m],
// bit-identical to row_sum16(q[a(m)]): the levels pair the same lanes in the
// same order (lane ^ 1, lane ^ 2, row_half_mirror, row_mirror) and every
// addition has the same two operands, but each lane adds only the values it
// keeps (27 instructions instead of 4 x 12).  a(m) is invariant under both
// mirrors, and lanes m and m ^ 1 (m ^ 2) keep complementary values.
constexpr uint32_t TR16_A = 0u | (2u << 2) | (1u << 4) | (3u << 6) | (3u << 8) | (1u << 10) | (2u << 12) |
                            (0u << 14) | (0u << 16) | (2u << 18) | (1u << 20) | (3u << 22) | (3u << 24) |
                            (1u << 26) | (2u << 28) | (0u << 30);
__device__ inline int tr16_value(int m) { return (int)((TR16_A >> (2 * m)) & 3u); }
__device__ inline double tr_sum16x4(const double (&q)[4], int am) {
  const bool s1 = am >> 1, s2 = am & 1;
  double t[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const double keep = s1 ? q[k + 2] : q[k], send = s1 ? q[k] : q[k + 2];
    t[k] = keep + dpp_d<0xB1>(send);  // quad_perm [1,0,3,2]
  }
  const double keep = s2 ? t[1] : t[0], send = s2 ? t[0] : t[1];
  double u = keep + dpp_d<0x4E>(send);  // quad_perm [2,3,0,1]
  u += dpp_d<0x141>(u);                 // row_half_mirror
  u += dpp_d<0x140>(u);                 // row_mirror
  return u;
}

template <int STEPS, bool CLAMP, class XT>
__global__ __launch_bounds__(256) void mv_sym64_kernel(const double* __restrict__ E, int d, int64_t n,
                                                        double rscale, double clampv,
                                                        const XT* __restrict__ X, int64_t ldx,
                                                        const int2* __restrict__ map,
                                                        double* __restrict__ part_r,
                                                        double* __restrict__ part_c) {
  __shared__ double tab[256];
  __shared__ double colbuf[4][S64_CH * 16];
  tab[threadIdx.x] = exp2((double)threadIdx.x * (1.0 / 256.0));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int am = tr16_value(m);  // value index this lane's column sum ends with
  const int2 bs = map[blockIdx.x];
  const int64_t b = bs.x;
  const int64_t I0 = b * S64_RB + w * 64;
  const bool active = I0 < n;
  const int64_t wrem = (n - b * S64_RB + 63) / 64;
  const int nact = wrem < 4 ? (int)wrem : 4;
  double bo[4][STEPS], xr[4];
#pragma unroll
  for (int it = 0; it < 4; ++it) {
    const int64_t i = I0 + 16 * it + m;
#pragma unroll
    for (int s = 0; s < STEPS; ++s) {
      const int k = kq + 4 * s;
      bo[it][s] = (i < n && k < d) ? E[i * d + k] * rscale : 0.0;
    }
    xr[it] = i < n ? (double)X[i * ldx] : 0.0;
  }
  const int64_t ntiles = (n + 15) / 16;
  const int64_t tdiag = S64_DT * (b + 1);  // first tile right of the diagonal block
  const int64_t t0 = (S64_DT * b > (int64_t)bs.y * S64_SEG) ? S64_DT * b : (int64_t)bs.y * S64_SEG;
  const int64_t t1 = ((int64_t)bs.y * S64_SEG + S64_SEG < ntiles) ? (int64_t)bs.y * S64_SEG + S64_SEG : ntiles;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  __syncthreads();  // table
  for (int64_t tc = t0; tc < t1; tc += S64_CH) {
    const int nt = (int)((t1 - tc) < S64_CH ? (t1 - tc) : S64_CH);
    if (active) {
      for (int tt = 0; tt < nt; ++tt) {
        const int64_t J0 = (tc + tt) * 16;
        double a[STEPS];
#pragma unroll
        for (int s = 0; s < STEPS; ++s) {
          const int64_t j = J0 + m;
          const int k = kq + 4 * s;
          a[s] = (j < n && k < d) ? E[j * d + k] : 0.0;
        }
        double xv[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t j = J0 + kq + 4 * r;
          xv[r] = j < n ? (double)X[j * ldx] : 0.0;
        }
        double q[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int it = 0; it < 4; ++it) {
          doublex4 D = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int s = 0; s < STEPS; ++s) D = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], bo[it][s], D, 0, 0, 0);
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            double tp = D[r];
            if constexpr (CLAMP) tp = fmin(tp, clampv);
            const double ev = exp2_t256(tp, tab);
            acc[it] = __builtin_fma(ev, xv[r], acc[it]);
            q[r] = __builtin_fma(ev, xr[it], q[r]);
          }
        }
        const double cs = tr_sum16x4(q, am);
        if (m < 4) colbuf[w][tt * 16 + kq + 4 * am] = cs;
      }
    }
    __syncthreads();
    if (threadIdx.x < nt * 16) {
      const int64_t t = tc + (threadIdx.x >> 4);
      const int64_t j = t * 16 + (threadIdx.x & 15);
      if (t >= tdiag && j < n) {
        double v = 0.0;
        for (int ww = 0; ww < nact; ++ww) v += colbuf[ww][threadIdx.x];
        part_c[s64_coff(b, n) + j - S64_RB * (b + 1)] = v;
      }
    }
    __syncthreads();
  }
  if (!active) return;
#pragma unroll
  for (int it = 0; it < 4; ++it) {
    double v = acc[it];
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    const int64_t i = I0 + 16 * it + m;
    if (kq == 0 && i < n) part_r[(int64_t)bs.y * n + i] = v;
  }
}

// s_i = lam xr_i + sum_seg part_r[seg][i] + sum_{b < i / S64_RB} part_c[coff(b) + i - S64_RB (b + 1)];
// Y_i = B_i - s_i (residual) or s_i
template <class XT>
__global__ void mv_sym64_reduce_kernel(const double* __restrict__ part_r, const double* __restrict__ part_c,
                                       int64_t n, double lam, const XT* __restrict__ Xr, int64_t ldxr,
                                       const double* __restrict__ B, double* __restrict__ Y, int64_t ldy) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t ntiles = (n + 15) / 16;
  const int64_t b = i / S64_RB;
  const int64_t nseg = (ntiles + S64_SEG - 1) / S64_SEG;
  double a8[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double tail = 0.0;
  int64_t sg = (S64_DT * b) / S64_SEG;  // first segment of row block b
  for (; sg + 8 <= nseg; sg += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) a8[u] += part_r[(sg + u) * n + i];
  }
  for (; sg < nseg; ++sg) tail += part_r[sg * n + i];
  int64_t bb = 0;
  for (; bb + 8 <= b; bb += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int64_t q = bb + u;
      a8[u] += part_c[s64_coff(q, n) + i - S64_RB * (q + 1)];
    }
  }
  for (; bb < b; ++bb) tail += part_c[s64_coff(bb, n) + i - S64_RB * (bb + 1)];
  double s = (((a8[0] + a8[1]) + (a8[2] + a8[3])) + ((a8[4] + a8[5]) + (a8[6] + a8[7]))) + tail;
  if (Xr) s += lam * (double)Xr[i * ldxr];
  Y[i * ldy] = B ? B[i * ldy] - s : s;
}

// Y[i][c] = lam * Xr[i][c] + sum_js part[js][i][c]     (Xr may be null)
// MODE_RES: Y = B - lam*Xr - sum   (FP64 residual)
template <class PT, class XT, class YT, bool RES>
__global__ void mv_reduce_kernel(const PT* __restrict__ part, int64_t n_rows, int k, int64_t ldp,
                                 int jsplit, double lam, const XT* __restrict__ Xr, int64_t ldx,
                                 const YT* __restrict__ B, YT* __restrict__ Y, int64_t ldy,
                                 int accumulate, const int* __restrict__ skip) {
  if (skip && *skip) return;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n_rows * k) return;
  const int64_t i = idx / k;
  const int c = (int)(idx % k);
  double s = 0.0;
  for (int js = 0; js < jsplit; ++js) s += (double)part[((int64_t)js * n_rows + i) * ldp + c];
  if (Xr) s += lam * (double)Xr[i * ldx + c];
  if constexpr (RES) {
    Y[i * ldy + c] = (YT)((double)B[i * ldy + c] - s);
  } else {
    if (accumulate) s += (double)Y[i * ldy + c];
    Y[i * ldy + c] = (YT)s;
  }
}

template <class T>
__global__ void diag_kernel(const T* __restrict__ E, int64_t n, int d, double scale, double lam,
                            T* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int q = 0; q < d; ++q) {
    double v = (double)E[i * d + q];
    s += v * v;
  }
  out[i] = (T)(lam + exp(scale * s));
}

// K[i][j] = exp(min(s * x_i . y_j, c))  (FP32 math for FP32, FP64 for FP64)
template <class T>
__global__ void kernel_eval_kernel(const T* __restrict__ Xe, int64_t m, const T* __restrict__ Ye,
                                   int64_t p, int d, double scale, int clamp, double clampv,
                                   double diag_add, T* __restrict__ K, int64_t ldk) {
  const int64_t total = m * p;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx / p, j = idx % p;
    T g = 0;
    for (int q = 0; q < d; ++q) g += Xe[i * d + q] * Ye[j * d + q];
    g = g * (T)scale;
    if (clamp) g = g < (T)clampv ? g : (T)clampv;
    T v = exp(g);
    if (i == j) v += (T)diag_add;
    K[i * ldk + j] = v;
  }
}

static dim3 eval_grid(int64_t total) {
  return dim3((unsigned)std::min<int64_t>(ceil_div(total, 256), 65536));
}

// ---------------------------------------------------------------------------
// Operator handle
// ---------------------------------------------------------------------------

static int pick_jsplit(int64_t n_rows, int64_t n_cols, int rows_per_block) {
  const int64_t blocks = ceil_div(n_rows, rows_per_block);
  const int64_t tiles = ceil_div(n_cols, 32);
  int64_t js = 8;
  // aim for >= ~2048 blocks of 4 waves; at least 2 column tiles per split
  while (blocks * js < 2048 && js < 64) js *= 2;
  while (js > 1 && tiles / js < 2) js /= 2;
  return (int)js;
}

static int launch_vec(int mode, const void* colpack, const void* rowpack, int ka,
                      int64_t n_rows, int64_t n_cols, const float* X, int64_t ldx, int kc,
                      float* part, int jsplit, bool clamp, bool deriv, float clampv,
                      const int* skip, hipStream_t st) {
  const int rows_per_block = 4 * 64;  // ITW=2
  dim3 grid((unsigned)(ceil_div(n_rows, rows_per_block) * jsplit)), block(256);
#define LAKER_VEC_LAUNCH(MODE, NS, KC, CL, DV)                                                  \
  hipLaunchKernelGGL((mv_vec_kernel<MODE, NS, KC, 2, CL, DV>), grid, block, 0, st, colpack, rowpack, \
                     ka, n_rows, n_cols, X, ldx, kc, part, jsplit, clampv, skip)
#define LAKER_VEC_KC(MODE, NS, CL, DV)              \
  do {                                              \
    if (kc == 1) LAKER_VEC_LAUNCH(MODE, NS, 1, CL, DV); \
    else if (kc == 2) LAKER_VEC_LAUNCH(MODE, NS, 2, CL, DV); \
    else LAKER_VEC_LAUNCH(MODE, NS, 4, CL, DV);     \
  } while (0)
#define LAKER_VEC_CL(MODE, NS)                                  \
  do {                                                          \
    if (deriv) {                                                \
      if (clamp) LAKER_VEC_KC(MODE, NS, true, true);            \
      else LAKER_VEC_KC(MODE, NS, false, true);                 \
    } else {                                                    \
      if (clamp) LAKER_VEC_KC(MODE, NS, true, false);           \
      else LAKER_VEC_KC(MODE, NS, false, false);                \
    }                                                           \
  } while (0)
  if (mode == LAKER_PREC_FP32) {
    const int ns = ka / 2;
    if (ns <= 4) LAKER_VEC_CL(LAKER_PREC_FP32, 4);
    else if (ns <= 8) LAKER_VEC_CL(LAKER_PREC_FP32, 8);
    else if (ns <= 16) LAKER_VEC_CL(LAKER_PREC_FP32, 16);
    else return fail(LAKER_EINVAL, "embedding_dim > 32 not supported by the FP32 matvec");
  } else {
    const int ns = ka / 16;
    if (ns <= 2) LAKER_VEC_CL(LAKER_PREC_MIXED, 2);
    else if (ns <= 4) LAKER_VEC_CL(LAKER_PREC_MIXED, 4);
    else if (ns <= 8) LAKER_VEC_CL(LAKER_PREC_MIXED, 8);
    else if (ns <= 12) LAKER_VEC_CL(LAKER_PREC_MIXED, 12);
    else return fail(LAKER_EINVAL, "embedding_dim > 32 not supported by the mixed matvec");
  }
#undef LAKER_VEC_CL
#undef LAKER_VEC_KC
#undef LAKER_VEC_LAUNCH
  LAKER_HIP_CHECK(hipGetLastError());
  return LAKER_OK;
}

// k = 1 product through the LDS-staged kernel.  Returns 1 when the shape has
// no instance (the caller then uses mv_vec_kernel).
static int launch_vec1(int mode, const void* colpack, const void* rowpack, int ka, int64_t n_rows,
                       int64_t n_cols, const float* X, int64_t ldx, double* part, int jsplit,
                       bool clamp, float clampv, const int* skip, hipStream_t st) {
  const size_t lds = 2 * (size_t)(mode == LAKER_PREC_FP32 ? V1Layout<LAKER_PREC_FP32>::chunk_bytes(ka)
                                                          : V1Layout<LAKER_PREC_MIXED>::chunk_bytes(ka));
  if (lds > 65536) return 1;
  dim3 grid((unsigned)(ceil_div(n_rows, 256) * jsplit)), block(256);
#define LAKER_V1(MODE, NS)                                                                        \
  do {                                                                                            \
    if (clamp)                                                                                    \
      hipLaunchKernelGGL((mv_vec1_kernel<MODE, NS, true>), grid, block, lds, st, colpack, rowpack, \
                         ka, n_rows, n_cols, X, ldx, part, jsplit, clampv, skip);                 \
    else                                                                                          \
      hipLaunchKernelGGL((mv_vec1_kernel<MODE, NS, false>), grid, block, lds, st, colpack,        \
                         rowpack, ka, n_rows, n_cols, X, ldx, part, jsplit, clampv, skip);        \
  } while (0)
  if (mode == LAKER_PREC_FP32) {
    switch (ka / 2) {
      case 2: LAKER_V1(LAKER_PREC_FP32, 2); break;
      case 3: LAKER_V1(LAKER_PREC_FP32, 3); break;
      case 4: LAKER_V1(LAKER_PREC_FP32, 4); break;
      case 5: LAKER_V1(LAKER_PREC_FP32, 5); break;
      case 8: LAKER_V1(LAKER_PREC_FP32, 8); break;
      case 16: LAKER_V1(LAKER_PREC_FP32, 16); break;
      default: return 1;
    }
  } else {
    switch (ka / 16) {
      case 1: LAKER_V1(LAKER_PREC_MIXED, 1); break;
      case 2: LAKER_V1(LAKER_PREC_MIXED, 2); break;
      case 3: LAKER_V1(LAKER_PREC_MIXED, 3); break;
      case 4: LAKER_V1(LAKER_PREC_MIXED, 4); break;
      case 6: LAKER_V1(LAKER_PREC_MIXED, 6); break;
      default: return 1;
    }
  }
#undef LAKER_V1
  LAKER_KCHECK("mv_vec1_kernel");
  return LAKER_OK;
}

// Symmetric k = 1 product of the square operator (see mv_sym1_kernel) plus
// its reduction into Y.  Returns 1 when the shape has no instance.
static int launch_sym1(Operator& op, int mode, const void* colpack, const void* rowpack, int ka, int64_t n,
                       const float* X, int64_t ldx, float* Y, int64_t ldy, double lam, const float* Xr,
                       int64_t ldxr, bool clamp, float clampv, int accumulate, const int* skip,
                       hipStream_t st, double* dots) {
  // the 16-bit modes (BF16 3-way split, FP16 2-way split) share one layout
  const size_t chunk = mode == LAKER_PREC_FP32 ? V1Layout<LAKER_PREC_FP32>::chunk_bytes(ka)
                                               : V1Layout<LAKER_PREC_MIXED>::chunk_bytes(ka);
  const size_t lds = 2 * chunk + 2 * 4 * V1_JT * 32 * sizeof(float);
  if (lds > 65536 || op.sym_nwg <= 0) return 1;
  const int ns = mode == LAKER_PREC_FP32 ? ka / 2 : ka / 16;
  const bool fp32_ok = ns == 2 || ns == 3 || ns == 4 || ns == 5 || ns == 8 || ns == 16;
  const bool bf16_ok = ns == 1 || ns == 2 || ns == 3 || ns == 4 || ns == 6;
  if (mode == LAKER_PREC_FP32 ? !fp32_ok : !bf16_ok) return 1;
  const int64_t ntiles = ceil_div(n, 32);
  const int64_t nsegmax = ceil_div(ntiles, SYM_SEG);
  const int64_t ncb = ceil_div(n, SYM_RB) - 1;  // row blocks with columns right of their diagonal block
  LAKER_TRY(op.ensure_part((size_t)nsegmax * n * sizeof(double)));
  LAKER_TRY(op.part_c.ensure((size_t)std::max<int64_t>(1, sym_coff(ncb, n)) * sizeof(double)));
  double* part_r = op.part.as<double>();
  double* part_c = op.part_c.as<double>();
  const int2* map = op.sym_map.as<int2>();
  dim3 grid((unsigned)op.sym_nwg), block(256);
#define LAKER_S1(MODE, NS)                                                                            \
  do {                                                                                                \
    if (clamp)                                                                                        \
      hipLaunchKernelGGL((mv_sym1_kernel<MODE, NS, true>), grid, block, lds, st, colpack, rowpack, n, X, \
                         ldx, map, part_r, part_c, clampv, skip);                                     \
    else                                                                                              \
      hipLaunchKernelGGL((mv_sym1_kernel<MODE, NS, false>), grid, block, lds, st, colpack, rowpack, n, X, \
                         ldx, map, part_r, part_c, clampv, skip);                                     \
  } while (0)
  if (mode == LAKER_PREC_FP32) {
    switch (ns) {
      case 2: LAKER_S1(LAKER_PREC_FP32, 2); break;
      case 3: LAKER_S1(LAKER_PREC_FP32, 3); break;
      case 4: LAKER_S1(LAKER_PREC_FP32, 4); break;
      case 5: LAKER_S1(LAKER_PREC_FP32, 5); break;
      case 8: LAKER_S1(LAKER_PREC_FP32, 8); break;
      default: LAKER_S1(LAKER_PREC_FP32, 16); break;
    }
  } else if (mode == GRAM_F16S) {
    switch (ns) {
      case 1: LAKER_S1(GRAM_F16S, 1); break;
      case 2: LAKER_S1(GRAM_F16S, 2); break;
      case 3: LAKER_S1(GRAM_F16S, 3); break;
      case 4: LAKER_S1(GRAM_F16S, 4); break;
      default: LAKER_S1(GRAM_F16S, 6); break;
    }
  } else {
    switch (ns) {
      case 1: LAKER_S1(LAKER_PREC_MIXED, 1); break;
      case 2: LAKER_S1(LAKER_PREC_MIXED, 2); break;
      case 3: LAKER_S1(LAKER_PREC_MIXED, 3); break;
      case 4: LAKER_S1(LAKER_PREC_MIXED, 4); break;
      default: LAKER_S1(LAKER_PREC_MIXED, 6); break;
    }
  }
#undef LAKER_S1
  LAKER_KCHECK("mv_sym1_kernel");
  hipLaunchKernelGGL(mv_sym_reduce_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st,
                     (const double*)part_r, (const double*)part_c, n, lam, Xr, ldxr, Y, ldy, accumulate, skip,
                     Xr ? dots : nullptr);
  LAKER_KCHECK("mv_sym_reduce_kernel");
  return LAKER_OK;
}

static int launch_multi(int mode, int gx, const void* colpack, const void* rowpack, int ka,
                        int64_t n_rows, int64_t n_cols, const float* X, int64_t ldx, int kcols,
                        int c_base, int ncc, float* part, int64_t ldp, int jsplit, bool clamp,
                        bool deriv, float clampv, hipStream_t st) {
  const int rows_per_block = 4 * 64;  // ITW=2
  dim3 grid((unsigned)(ceil_div(n_rows, rows_per_block) * jsplit)), block(256);
#define LAKER_MULTI_LAUNCH(MODE, NS, GX, NCC, CL, DV)                                             \
  hipLaunchKernelGGL((mv_multi_kernel<MODE, NS, GX, NCC, 2, CL, DV>), grid, block, 0, st, colpack, \
                     rowpack, ka, n_rows, n_cols, X, ldx, kcols, c_base, part, ldp, jsplit, clampv)
#define LAKER_MULTI_NCC(MODE, NS, GX, CL, DV)                   \
  do {                                                          \
    if (ncc == 1) LAKER_MULTI_LAUNCH(MODE, NS, GX, 1, CL, DV);  \
    else LAKER_MULTI_LAUNCH(MODE, NS, GX, 2, CL, DV);           \
  } while (0)
#define LAKER_MULTI_CL(MODE, NS, GX)                                        \
  do {                                                                      \
    if (deriv) {                                                            \
      if (clamp) LAKER_MULTI_NCC(MODE, NS, GX, true, true);                 \
      else LAKER_MULTI_NCC(MODE, NS, GX, false, true);                      \
    } else {                                                                \
      if (clamp) LAKER_MULTI_NCC(MODE, NS, GX, true, false);                \
      else LAKER_MULTI_NCC(MODE, NS, GX, false, false);                     \
    }                                                                       \
  } while (0)
#define LAKER_MULTI_GX(MODE, NS)                                                  \
  do {                                                                            \
    if (gx == LAKER_PREC_FP32) LAKER_MULTI_CL(MODE, NS, LAKER_PREC_FP32);         \
    else if (gx == LAKER_PREC_BF16) LAKER_MULTI_CL(MODE, NS, LAKER_PREC_BF16);    \
    else LAKER_MULTI_CL(MODE, NS, LAKER_PREC_MIXED);                              \
  } while (0)
  if (mode == LAKER_PREC_FP32) {
    const int ns = ka / 2;
    if (ns <= 4) LAKER_MULTI_GX(LAKER_PREC_FP32, 4);
    else if (ns <= 8) LAKER_MULTI_GX(LAKER_PREC_FP32, 8);
    else if (ns <= 16) LAKER_MULTI_GX(LAKER_PREC_FP32, 16);
    else return fail(LAKER_EINVAL, "embedding_dim > 32 not supported");
  } else {
    const int ns = ka / 16;
    if (ns <= 2) LAKER_MULTI_GX(LAKER_PREC_MIXED, 2);
    else if (ns <= 4) LAKER_MULTI_GX(LAKER_PREC_MIXED, 4);
    else if (ns <= 8) LAKER_MULTI_GX(LAKER_PREC_MIXED, 8);
    else if (ns <= 12) LAKER_MULTI_GX(LAKER_PREC_MIXED, 12);
    else return fail(LAKER_EINVAL, "embedding_dim > 32 not supported");
  }
#undef LAKER_MULTI_GX
#undef LAKER_MULTI_CL
#undef LAKER_MULTI_NCC
#undef LAKER_MULTI_LAUNCH
  LAKER_HIP_CHECK(hipGetLastError());
  return LAKER_OK;
}

template <class XT>
static int launch_f64(const double* Ecol, const double* Erow, int d, int64_t n_rows, int64_t n_cols,
                      double scale, bool clamp, double clampv, const XT* X, int64_t ldx, int kcols,
                      double* part, int64_t ldp, int jsplit, hipStream_t st) {
  dim3 grid((unsigned)(ceil_div(n_rows, 256) * jsplit)), block(256);
  for (int c0 = 0; c0 < kcols; c0 += 4) {
    const int rem = kcols - c0;
#define LAKER_F64_LAUNCH(KC, DM, CL)                                                          \
  hipLaunchKernelGGL((mv_f64_kernel<KC, DM, CL, XT>), grid, block, 0, st, Ecol, Erow, d, n_rows, \
                     n_cols, scale, clampv, X, ldx, kcols, c0, part, ldp, jsplit)
#define LAKER_F64_DM(KC, CL)                                  \
  do {                                                        \
    if (d <= 8) LAKER_F64_LAUNCH(KC, 8, CL);                  \
    else if (d <= 16) LAKER_F64_LAUNCH(KC, 16, CL);           \
    else LAKER_F64_LAUNCH(KC, 32, CL);                        \
  } while (0)
#define LAKER_F64_KC(CL)                                      \
  do {                                                        \
    if (rem == 1) LAKER_F64_DM(1, CL);                        \
    else if (rem == 2) LAKER_F64_DM(2, CL);                   \
    else LAKER_F64_DM(4, CL);                                 \
  } while (0)
    if (clamp) LAKER_F64_KC(true);
    else LAKER_F64_KC(false);
#undef LAKER_F64_KC
#undef LAKER_F64_DM
#undef LAKER_F64_LAUNCH
    LAKER_HIP_CHECK(hipGetLastError());
  }
  return LAKER_OK;
}

// Core dispatcher:  Y (n_rows x k) = [lam*Xr] + exp(min(s*Erow.Ecol^T, c)) X
// Works for the square operator (rows = cols = E) and the rectangular
// cross products.  fp32 X/Y for FP32/MIXED/BF16 precisions.
int fused_product_f32(Operator& op, int precision, const Side& rows, const Side& cols,
                      const float* X, int k, int64_t ldx, float* Y, int64_t ldy, double lam,
                      const float* Xr, int64_t ldxr, bool clamp, bool deriv, int accumulate,
                      hipStream_t st, const int* skip, double* dots3, bool* dots3_done) {
  if (dots3_done) *dots3_done = false;
  LAKER_CHECK_ARG(k >= 1, "k must be >= 1");
  const float clampv = (float)(op.clamp_limit * kLog2eD);
  const bool mixed = precision == LAKER_PREC_MIXED || precision == LAKER_PREC_BF16 || precision == LAKER_PREC_HILO ||
                     precision == PREC_HILO_XBF16;
  const void* colpack = mixed ? cols.pack16 : cols.pack32;
  const void* rowpack = mixed ? rows.pack16 : rows.pack32;
  const int ka = mixed ? op.ka16 : op.ka32;
  const int gmode = mixed ? LAKER_PREC_MIXED : LAKER_PREC_FP32;
  const int64_t n_rows = rows.n, n_cols = cols.n;
  const int jsplit = pick_jsplit(n_rows, n_cols, 256);
  if (k <= 4) {
    LAKER_TRY(op.ensure_part((size_t)jsplit * n_rows * 4 * sizeof(float)));
    // the square operator's own product: symmetric kernel (each exp once per pair)
    const bool self_sq = n_rows == n_cols && rows.pack16 == op.self.row16.ptr && cols.pack16 == op.self.col16.ptr &&
                         rows.pack32 == op.self.row32.ptr && cols.pack32 == op.self.col32.ptr;
    if (k == 1 && !deriv && self_sq && op.sym_enabled) {
      // split precisions: the FP16 2-way split (3 products) when the scaled
      // embedding fits FP16, else the BF16 3-way split (6 products)
      const bool f16s = mixed && op.f16s_ok;
      double* d3 = (dots3 && Xr) ? dots3 : nullptr;
      const int rs = f16s ? launch_sym1(op, GRAM_F16S, op.self.col16h.ptr, op.self.row16h.ptr, op.ka16h, n_rows, X,
                                        ldx, Y, ldy, lam, Xr, ldxr, clamp, clampv, accumulate, skip, st, d3)
                          : launch_sym1(op, gmode, colpack, rowpack, ka, n_rows, X, ldx, Y, ldy, lam, Xr, ldxr,
                                        clamp, clampv, accumulate, skip, st, d3);
      if (rs == 0 && d3 && dots3_done) *dots3_done = true;
      if (rs <= 0) return rs;
    }
    float* part = op.part.as<float>();
    int rc = 1;
    if (k == 1 && !deriv)
      rc = launch_vec1(gmode, colpack, rowpack, ka, n_rows, n_cols, X, ldx, op.part.as<double>(), jsplit,
                       clamp, clampv, skip, st);
    if (rc < 0) return rc;
    int64_t tot = n_rows * k;
    if (rc == 0) {
      // FP64 partial sums of the LDS-staged k = 1 kernel
      hipLaunchKernelGGL((mv_reduce_kernel<double, float, float, false>),
                         dim3((unsigned)ceil_div(tot, 256)), dim3(256), 0, st, op.part.as<double>(),
                         n_rows, 1, (int64_t)1, jsplit, lam, Xr, ldxr, (const float*)nullptr, Y, ldy,
                         accumulate, skip);
      LAKER_KCHECK("mv_reduce_kernel");
      return LAKER_OK;
    }
    LAKER_TRY(launch_vec(gmode, colpack, rowpack, ka, n_rows, n_cols, X, ldx, k, part, jsplit,
                         clamp, deriv, clampv, skip, st));
    const int kcp = k == 1 ? 1 : (k == 2 ? 2 : 4);
    hipLaunchKernelGGL((mv_reduce_kernel<float, float, float, false>),
                       dim3((unsigned)ceil_div(tot, 256)), dim3(256), 0, st, part, n_rows, k,
                       (int64_t)kcp, jsplit, lam, Xr, ldxr, (const float*)nullptr, Y, ldy,
                       accumulate, skip);
    LAKER_KCHECK("mv_reduce_kernel");
    LAKER_HIP_CHECK(hipGetLastError());
    return LAKER_OK;
  }
  // multi-RHS on MFMA.  GX precision: FP32 path -> FP32 MFMA; MIXED -> hi/lo split; BF16 -> bf16.
  // the right-hand-side product: FP32 MFMA; BF16 (1 product); hi/lo (3 products);
  // 3-way split (6 products, the solver's mixed operator)
  const int gx = precision == LAKER_PREC_FP32 ? LAKER_PREC_FP32
                 : precision == LAKER_PREC_BF16 ? LAKER_PREC_BF16
                 : precision == LAKER_PREC_HILO ? LAKER_PREC_MIXED
                 : precision == PREC_HILO_XBF16 ? GX_PLXH
                                                : GX_SPLIT3;
  // columns per launch: up to 64 (NCC=2) per pass; passes recompute the Gram tile.
  const int64_t kp = round_up(k, 32);
  LAKER_TRY(op.ensure_part((size_t)jsplit * n_rows * kp * sizeof(float)));
  float* part = op.part.as<float>();
  // the square operator's Gram from the FP16 2-way split when E fits FP16 (d <= 16)
  const bool self_sq = n_rows == n_cols && rows.pack16 == op.self.row16.ptr && cols.pack16 == op.self.col16.ptr;
  const int ns16h = op.ka16h / 16;
  const bool f16s = mixed && !deriv && gx != LAKER_PREC_FP32 && self_sq && op.f16s_ok && (ns16h == 2 || ns16h == 3);
  const int ns16 = op.ka16 / 16;
  const bool packed_x =
      mixed && !deriv && gx != LAKER_PREC_FP32 && (f16s || ns16 == 2 || ns16 == 4 || ns16 == 6);
  if (packed_x) {
    // X parts packed per pass, each region sized for the widest pass its
    // mode takes (pack_x_kernel writes ntiles * 128 * nq vectors per part):
    // 3 x 2 chunks (3-way split), 2 x 4 chunks (hi/lo), 1 x 4 (BF16) or
    // 1 x 8 (BF16-exact X)
    const int64_t ntiles = ceil_div(n_cols, 32);
    const int nq_max = gx == GX_SPLIT3 ? 2 : (gx == GX_PLXH ? 8 : 4);
    const size_t part_bytes = (size_t)ntiles * 128 * nq_max * sizeof(bf16x8);
    // parts written: hi only (BF16, BF16-exact X), hi + lo (hi/lo), hi + lo + mid (3-way)
    const int nparts = gx == GX_SPLIT3 ? 3 : (gx == LAKER_PREC_MIXED ? 2 : 1);
    LAKER_TRY(op.xpack.ensure(nparts * part_bytes));
    bf16x8* xh = op.xpack.as<bf16x8>();
    bf16x8* xl = nparts >= 2 ? reinterpret_cast<bf16x8*>(op.xpack.as<char>() + part_bytes) : nullptr;
    bf16x8* xm = nparts >= 3 ? reinterpret_cast<bf16x8*>(op.xpack.as<char>() + 2 * part_bytes) : nullptr;
    const int jsplit2 = jsplit;
    const dim3 block(256);
    for (int c0 = 0; c0 < k;) {
      // column chunks of 32 per pass: the 3-way split keeps 2 for its register budget;
      // hi/lo and BF16 take 4 when more than 64 columns remain (half the Gram passes);
      // BF16-exact X with the hi/lo exp tile (two products) takes 8 when more than
      // 128 remain (248 VGPRs, one column tile per LDS chunk)
      const int rem = k - c0;
      const int nq = (gx == GX_PLXH && rem > 128) ? 8 : (gx != GX_SPLIT3 && rem > 64) ? 4 : (rem > 32 ? 2 : 1);
      hipLaunchKernelGGL(pack_x_kernel, dim3((unsigned)ceil_div(ntiles * 128 * nq, 256)), dim3(256), 0, st, X,
                         ldx, n_cols, k, c0, ntiles, nq, xh, xl, xm);
      LAKER_KCHECK("pack_x_kernel");
#define LAKER_M2J(NS, GXM, NC, JTV)                                                                   \
  do {                                                                                                \
    constexpr size_t lds2 = 2 * (size_t)m2_chunk_bytes(NS, M2Parts<GXM>::value, NC, JTV);            \
    /* one row tile per wave for the register-heavy 3-way split and 4-chunk passes */                \
    constexpr int ITWV = (GXM == GX_SPLIT3 || NC >= 4) ? 1 : 2;                                       \
    dim3 grid2((unsigned)(ceil_div(n_rows, 128 * ITWV) * jsplit2));                                   \
    if (f16s) {                                                                                       \
      if (clamp)                                                                                      \
        hipLaunchKernelGGL((mv_multi2_kernel<GRAM_F16S, NS, GXM, NC, JTV, ITWV, true>), grid2, block,  \
                           lds2, st, op.self.col16h.ptr, op.self.row16h.ptr, op.ka16h, n_rows, n_cols, \
                           (const bf16x8*)xh, (const bf16x8*)xl, (const bf16x8*)xm, k, c0, part, kp,   \
                           jsplit2, clampv);                                                          \
      else                                                                                            \
        hipLaunchKernelGGL((mv_multi2_kernel<GRAM_F16S, NS, GXM, NC, JTV, ITWV, false>), grid2, block, \
                           lds2, st, op.self.col16h.ptr, op.self.row16h.ptr, op.ka16h, n_rows, n_cols, \
                           (const bf16x8*)xh, (const bf16x8*)xl, (const bf16x8*)xm, k, c0, part, kp,   \
                           jsplit2, clampv);                                                          \
    } else if (clamp)                                                                                 \
      hipLaunchKernelGGL((mv_multi2_kernel<LAKER_PREC_MIXED, NS, GXM, NC, JTV, ITWV, true>), grid2,    \
                         block, lds2, st, colpack, rowpack, ka, n_rows, n_cols, (const bf16x8*)xh,    \
                         (const bf16x8*)xl, (const bf16x8*)xm, k, c0, part, kp, jsplit2, clampv);     \
    else                                                                                              \
      hipLaunchKernelGGL((mv_multi2_kernel<LAKER_PREC_MIXED, NS, GXM, NC, JTV, ITWV, false>), grid2,   \
                         block, lds2, st, colpack, rowpack, ka, n_rows, n_cols, (const bf16x8*)xh,    \
                         (const bf16x8*)xl, (const bf16x8*)xm, k, c0, part, kp, jsplit2, clampv);     \
  } while (0)
#define LAKER_M2C(NS, GXM, NC)                                                                        \
  do {                                                                                                \
    if (2 * m2_chunk_bytes(NS, M2Parts<GXM>::value, NC, 2) <= 65536) LAKER_M2J(NS, GXM, NC, 2);        \
    else LAKER_M2J(NS, GXM, NC, 1);                                                                   \
  } while (0)
#define LAKER_M2(NS, GXM)                              \
  do {                                                 \
    if (nq == 4) LAKER_M2C(NS, GXM, 4);                \
    else if (nq == 2) LAKER_M2C(NS, GXM, 2);           \
    else LAKER_M2C(NS, GXM, 1);                        \
  } while (0)
#define LAKER_M2_2(NS, GXM)                            \
  do {                                                 \
    if (nq == 2) LAKER_M2C(NS, GXM, 2);                \
    else LAKER_M2C(NS, GXM, 1);                        \
  } while (0)
#define LAKER_M2_8(NS, GXM)                            \
  do {                                                 \
    if (nq == 8) LAKER_M2C(NS, GXM, 8);                \
    else LAKER_M2(NS, GXM);                            \
  } while (0)
#define LAKER_M2_NS8(GXM)                    \
  do {                                       \
    if (gns == 2) LAKER_M2_8(2, GXM);        \
    else if (gns == 3) LAKER_M2_8(3, GXM);   \
    else if (gns == 4) LAKER_M2_8(4, GXM);   \
    else LAKER_M2_8(6, GXM);                 \
  } while (0)
#define LAKER_M2_NS(GXM)                     \
  do {                                       \
    if (gns == 2) LAKER_M2(2, GXM);          \
    else if (gns == 3) LAKER_M2(3, GXM);     \
    else if (gns == 4) LAKER_M2(4, GXM);     \
    else LAKER_M2(6, GXM);                   \
  } while (0)
#define LAKER_M2_NS3(GXM)                    \
  do {                                       \
    if (gns == 2) LAKER_M2_2(2, GXM);        \
    else if (gns == 3) LAKER_M2_2(3, GXM);   \
    else if (gns == 4) LAKER_M2_2(4, GXM);   \
    else LAKER_M2_2(6, GXM);                 \
  } while (0)
      // Gram steps: ka16h / 16 (FP16 split) or ka16 / 16 (BF16 split)
      const int gns = f16s ? ns16h : ns16;
      if (gx == GX_SPLIT3) LAKER_M2_NS3(GX_SPLIT3);
      else if (gx == LAKER_PREC_MIXED) LAKER_M2_NS(LAKER_PREC_MIXED);
      else if (gx == GX_PLXH) LAKER_M2_NS8(GX_PLXH);
      else LAKER_M2_NS(LAKER_PREC_BF16);
#undef LAKER_M2_NS
#undef LAKER_M2_NS8
#undef LAKER_M2_8
#undef LAKER_M2_NS3
#undef LAKER_M2
#undef LAKER_M2_2
#undef LAKER_M2C
#undef LAKER_M2J
      LAKER_KCHECK("mv_multi2_kernel");
      c0 += 32 * nq;
    }
  } else {
    for (int c0 = 0; c0 < k; c0 += 64) {
      const int ncc = (k - c0) > 32 ? 2 : 1;
      LAKER_TRY(launch_multi(gmode, gx, colpack, rowpack, ka, n_rows, n_cols, X, ldx, k, c0, ncc,
                             part, kp, jsplit, clamp, deriv, clampv, st));
    }
  }
  int64_t tot = n_rows * k;
  hipLaunchKernelGGL((mv_reduce_kernel<float, float, float, false>),
                     dim3((unsigned)ceil_div(tot, 256)), dim3(256), 0, st, part, n_rows, k, kp,
                     jsplit, lam, Xr, ldxr, (const float*)nullptr, Y, ldy, accumulate,
                     (const int*)nullptr);
  LAKER_KCHECK("mv_reduce_kernel");
  LAKER_HIP_CHECK(hipGetLastError());
  return LAKER_OK;
}

// FP64 product: Y = [lam*Xr] + exp(min(s*g, c)) X  (or residual Y = B - ...).
template <class XT>
int fused_product_f64(Operator& op, const Side& rows, const Side& cols, const XT* X, int k,
                      int64_t ldx, double* Y, int64_t ldy, double lam, const XT* Xr, int64_t ldxr,
                      bool clamp, const double* B, hipStream_t st) {
  LAKER_CHECK_ARG(k >= 1, "k must be >= 1");
  const int64_t n_rows = rows.n, n_cols = cols.n;
  int jsplit = 8;
  while (ceil_div(n_rows, 256) * jsplit < 1024 && jsplit < 64) jsplit *= 2;
  while (jsplit > 1 && ceil_div(n_cols, 64) / jsplit < 2) jsplit /= 2;
  const int64_t ldp = k;
  // t' = 256 s log2(e) g must stay inside int range for the table exp2
  const double tmax = op.max_sq_norm * std::fabs(op.scale) * kLog2eD * 256.0;
  if (k == 1 && op.sym_enabled && op.sym64_nwg > 0 && rows.e64 == cols.e64 && n_rows == n_cols &&
      op.d <= 16 && (clamp || tmax < 1e9)) {
    // square operator, one right-hand side: each exponential once per pair
    const int steps = (op.d + 3) / 4;
    const double rscale = op.scale * kLog2eD * 256.0;
    const double clampv = op.clamp_limit * kLog2eD * 256.0;
    const int64_t n = n_rows;
    const int64_t nseg = ceil_div(ceil_div(n, 16), S64_SEG);
    const int64_t ncb = ceil_div(n, S64_RB) - 1;
    LAKER_TRY(op.part64.ensure((size_t)nseg * n * sizeof(double)));
    LAKER_TRY(op.part_c64.ensure((size_t)std::max<int64_t>(1, s64_coff(ncb, n)) * sizeof(double)));
    double* part_r = op.part64.as<double>();
    double* part_c = op.part_c64.as<double>();
    const int2* map = op.sym64_map.as<int2>();
    dim3 grid((unsigned)op.sym64_nwg), block(256);
#define LAKER_S64(ST, CL)                                                                              \
  hipLaunchKernelGGL((mv_sym64_kernel<ST, CL, XT>), grid, block, 0, st, rows.e64, op.d, n, rscale, clampv, \
                     X, ldx, map, part_r, part_c)
#define LAKER_S64_ST(CL)                     \
  do {                                       \
    if (steps == 1) LAKER_S64(1, CL);        \
    else if (steps == 2) LAKER_S64(2, CL);   \
    else if (steps == 3) LAKER_S64(3, CL);   \
    else LAKER_S64(4, CL);                   \
  } while (0)
    if (clamp) LAKER_S64_ST(true);
    else LAKER_S64_ST(false);
#undef LAKER_S64_ST
#undef LAKER_S64
    LAKER_KCHECK("mv_sym64_kernel");
    hipLaunchKernelGGL((mv_sym64_reduce_kernel<XT>), dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st,
                       (const double*)part_r, (const double*)part_c, n, lam, Xr, ldxr, B, Y, ldy);
    LAKER_KCHECK("mv_sym64_reduce_kernel");
    return LAKER_OK;
  }
  LAKER_TRY(op.part64.ensure((size_t)jsplit * n_rows * k * sizeof(double)));
  double* part = op.part64.as<double>();
  if (op.d <= 16 && (clamp || tmax < 1e9)) {
    const int steps = (op.d + 3) / 4;
    const double rscale = op.scale * kLog2eD * 256.0;
    const double clampv = op.clamp_limit * kLog2eD * 256.0;
    // k = 1: 4 row groups of 16 per wave; k > 1: 2 row groups, up to 8 columns per pass
    const int itw = k == 1 ? 4 : 2;
    dim3 grid((unsigned)(ceil_div(n_rows, 64 * itw) * jsplit)), block(256);
    for (int c0 = 0; c0 < k; c0 += 8) {
      const int kc = k - c0;
#define LAKER_F64M(ST, KCT, ITWT, CL)                                                                     \
  hipLaunchKernelGGL((mv_f64m_kernel<ST, KCT, ITWT, CL, XT>), grid, block, 0, st, cols.e64, rows.e64, op.d, \
                     n_rows, n_cols, rscale, clampv, X, ldx, k, c0, part, ldp, jsplit)
#define LAKER_F64M_KC(ST, CL)                                       \
  do {                                                              \
    if (k == 1) LAKER_F64M(ST, 1, 4, CL);                           \
    else if (kc <= 2) LAKER_F64M(ST, 2, 2, CL);                     \
    else if (kc <= 4) LAKER_F64M(ST, 4, 2, CL);                     \
    else LAKER_F64M(ST, 8, 2, CL);                                  \
  } while (0)
#define LAKER_F64M_ST(CL)                        \
  do {                                           \
    if (steps == 1) LAKER_F64M_KC(1, CL);        \
    else if (steps == 2) LAKER_F64M_KC(2, CL);   \
    else if (steps == 3) LAKER_F64M_KC(3, CL);   \
    else LAKER_F64M_KC(4, CL);                   \
  } while (0)
      if (clamp) LAKER_F64M_ST(true);
      else LAKER_F64M_ST(false);
#undef LAKER_F64M_ST
#undef LAKER_F64M_KC
#undef LAKER_F64M
      LAKER_KCHECK("mv_f64m_kernel");
    }
  } else {
    LAKER_TRY(launch_f64<XT>(cols.e64, rows.e64, op.d, n_rows, n_cols, op.scale, clamp,
                             op.clamp_limit, X, ldx, k, part, ldp, jsplit, st));
  }
  int64_t tot = n_rows * k;
  if (B) {
    hipLaunchKernelGGL((mv_reduce_kernel<double, XT, double, true>),
                       dim3((unsigned)ceil_div(tot, 256)), dim3(256), 0, st, part, n_rows, k, ldp,
                       jsplit, lam, Xr, ldxr, B, Y, ldy, 0, (const int*)nullptr);
    LAKER_KCHECK("mv_reduce_kernel");
  } else {
    hipLaunchKernelGGL((mv_reduce_kernel<double, XT, double, false>),
                       dim3((unsigned)ceil_div(tot, 256)), dim3(256), 0, st, part, n_rows, k, ldp,
                       jsplit, lam, Xr, ldxr, (const double*)nullptr, Y, ldy, 0, (const int*)nullptr);
    LAKER_KCHECK("mv_reduce_kernel");
  }
  LAKER_HIP_CHECK(hipGetLastError());
  return LAKER_OK;
}
template int fused_product_f64<float>(Operator&, const Side&, const Side&, const float*, int,
                                      int64_t, double*, int64_t, double, const float*, int64_t,
                                      bool, const double*, hipStream_t);
template int fused_product_f64<double>(Operator&, const Side&, const Side&, const double*, int,
                                       int64_t, double*, int64_t, double, const double*, int64_t,
                                       bool, const double*, hipStream_t);

// Build the packed operand arrays of one embedding set.
template <class TE>
static int pack_store_t(SideStore& s, const TE* E, int64_t n, int d, int ka32, int ka16,
                        double cscale, bool need_col, bool need_row, hipStream_t st) {
  s.n = n;
  s.npad = round_up(n, 64);
  const int64_t n32 = s.npad * ka32, n16 = s.npad * ka16;
  if (need_col) {
    LAKER_TRY(s.col32.ensure(n32 * 4));
    LAKER_TRY(s.col16.ensure(n16 * 2));
  }
  if (need_row) {
    LAKER_TRY(s.row32.ensure(n32 * 4));
    LAKER_TRY(s.row16.ensure(n16 * 2));
  }
  LAKER_TRY(s.e64.ensure(n * d * 8));
  LAKER_TRY(s.user.ensure(n * d * sizeof(TE)));
  LAKER_HIP_CHECK(hipMemcpyAsync(s.user.ptr, E, n * d * sizeof(TE), hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL((pack_f32_kernel<TE>), dim3((unsigned)ceil_div(n32, 256)), dim3(256), 0, st, E,
                     n, d, s.npad, ka32, (float)cscale, need_col ? s.col32.as<float>() : nullptr,
                     need_row ? s.row32.as<float>() : nullptr);
  LAKER_KCHECK("pack_f32_kernel");
  hipLaunchKernelGGL((pack_bf16x4_kernel<TE>), dim3((unsigned)ceil_div(n16, 256)), dim3(256), 0, st,
                     E, n, d, s.npad, ka16, (float)cscale,
                     need_col ? s.col16.as<uint16_t>() : nullptr,
                     need_row ? s.row16.as<uint16_t>() : nullptr);
  LAKER_KCHECK("pack_bf16x4_kernel");
  hipLaunchKernelGGL((pack_f64_kernel<TE>), dim3((unsigned)ceil_div(n * d, 256)), dim3(256), 0, st,
                     E, n, d, s.e64.as<double>());
  LAKER_KCHECK("pack_f64_kernel");
  LAKER_HIP_CHECK(hipGetLastError());
  return LAKER_OK;
}

// FP16 2-way split operands of the square operator (k = 1 products).
template <class TE>
static int pack_f16_split_t(SideStore& s, const TE* E, int64_t n, int d, int ka, double cscale, hipStream_t st) {
  const int64_t nh = round_up(n, 64) * ka;
  LAKER_TRY(s.col16h.ensure(nh * 2));
  LAKER_TRY(s.row16h.ensure(nh * 2));
  hipLaunchKernelGGL((pack_f16x2_kernel<TE>), dim3((unsigned)ceil_div(nh, 256)), dim3(256), 0, st, E, n, d,
                     round_up(n, 64), ka, (float)cscale, s.col16h.as<uint16_t>(), s.row16h.as<uint16_t>());
  LAKER_KCHECK("pack_f16x2_kernel");
  return LAKER_OK;
}

int pack_f16_split(SideStore& s, const void* E, int dtype, int64_t n, int d, int ka, double scale, hipStream_t st) {
  const double cscale = scale * kLog2eD;
  if (dtype == LAKER_F32) return pack_f16_split_t<float>(s, (const float*)E, n, d, ka, cscale, st);
  return pack_f16_split_t<double>(s, (const double*)E, n, d, ka, cscale, st);
}

int pack_store(SideStore& s, const void* E, int dtype, int64_t n, int d, int ka32, int ka16,
               double scale, bool need_col, bool need_row, hipStream_t st) {
  const double cscale = scale * kLog2eD;
  if (dtype == LAKER_F32)
    return pack_store_t<float>(s, (const float*)E, n, d, ka32, ka16, cscale, need_col, need_row, st);
  return pack_store_t<double>(s, (const double*)E, n, d, ka32, ka16, cscale, need_col, need_row, st);
}

// ---------------------------------------------------------------------------
// Allocation-free products (laker_matvec_noalloc).  Their split-K partial sums
// live in `scratch`, a file-scope device array that every GPU the library
// runs on holds until the process ends.  Host code never passes `scratch`
// itself (its host-side name is not a device address): the device address
// comes from hipGetSymbolAddress, once per device with that device current,
// and is kept in scratch_dev[device].
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;  // per GPU
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex scratch_mu;
static void* scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device (resolved on first use).
static int scratch_for_current_device(void** out) {
  int dev = -1;
  LAKER_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(LAKER_EINVAL, "device ordinal " + std::to_string(dev) + " is beyond the scratch table");
  std::lock_guard<std::mutex> g(scratch_mu);
  if (scratch_dev[dev] == nullptr) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr) {
      (void)hipGetLastError();
      return fail(LAKER_EHIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    }
    scratch_dev[dev] = ptr;
  }
  *out = scratch_dev[dev];
  return LAKER_OK;
}

// Partial-sum bytes the allocation-free product needs for n rows and k <= 4
// columns: FP64 partials of the LDS-staged k = 1 kernel, FP32 partials with
// stride 1, 2 or 4 for k = 2..4.
static size_t noalloc_part_bytes(int64_t n, int k) {
  const int jsplit = pick_jsplit(n, n, 256);
  const int kcp = k == 1 ? 1 : (k == 2 ? 2 : 4);
  return k == 1 ? (size_t)jsplit * n * sizeof(double) : (size_t)jsplit * n * kcp * sizeof(float);
}

// Y = lambda X + exp(min(s E E^T, c)) X for an FP32 operator and k <= 4,
// with the partial sums in the caller-provided device buffer `part`
// (noalloc_part_bytes(n, k) bytes).  Every partial the reduction reads is
// written by the product kernel first, so `part` needs no clearing.  No
// device memory is allocated.
static int product_vec_into(Operator& op, int precision, const float* X, int k, int64_t ldx, float* Y,
                            int64_t ldy, void* part, hipStream_t st) {
  const bool clamp = op.clamp_for(LAKER_CLAMP_AUTO);
  const float clampv = (float)(op.clamp_limit * kLog2eD);
  const bool mixed = precision == LAKER_PREC_MIXED;
  const Side rows = op.rows(), cols = op.cols();
  const void* colpack = mixed ? cols.pack16 : cols.pack32;
  const void* rowpack = mixed ? rows.pack16 : rows.pack32;
  const int ka = mixed ? op.ka16 : op.ka32;
  const int gmode = mixed ? LAKER_PREC_MIXED : LAKER_PREC_FP32;
  const int64_t n = op.n;
  const int jsplit = pick_jsplit(n, n, 256);
  const int64_t tot = n * k;
  if (k == 1) {
    const int rc = launch_vec1(gmode, colpack, rowpack, ka, n, n, X, ldx, (double*)part, jsplit, clamp, clampv,
                               nullptr, st);
    if (rc < 0) return rc;
    if (rc == 0) {
      hipLaunchKernelGGL((mv_reduce_kernel<double, float, float, false>), dim3((unsigned)ceil_div(tot, 256)),
                         dim3(256), 0, st, (const double*)part, n, 1, (int64_t)1, jsplit, op.lambda_reg, X, ldx,
                         (const float*)nullptr, Y, ldy, 0, (const int*)nullptr);
      LAKER_KCHECK("mv_reduce_kernel");
      return LAKER_OK;
    }
    // no LDS-staged instance for this width: FP32 partials (half the bytes)
  }
  LAKER_TRY(launch_vec(gmode, colpack, rowpack, ka, n, n, X, ldx, k, (float*)part, jsplit, clamp, false, clampv,
                       nullptr, st));
  const int kcp = k == 1 ? 1 : (k == 2 ? 2 : 4);
  hipLaunchKernelGGL((mv_reduce_kernel<float, float, float, false>), dim3((unsigned)ceil_div(tot, 256)),
                     dim3(256), 0, st, (const float*)part, n, k, (int64_t)kcp, jsplit, op.lambda_reg, X, ldx,
                     (const float*)nullptr, Y, ldy, 0, (const int*)nullptr);
  LAKER_KCHECK("mv_reduce_kernel");
  return LAKER_OK;
}

}  // namespace laker

using namespace laker;

extern "C" {

int laker_operator_create(const void* E, int64_t n, int32_t d, int32_t dtype, double lambda_reg,
                          double scale, int32_t clamp_policy, hipStream_t stream,
                          laker_operator_t* out) {
  LAKER_CHECK_ARG(out != nullptr, "out handle is NULL");
  LAKER_CHECK_ARG(E != nullptr || n == 0, "embeddings pointer is NULL");
  LAKER_CHECK_ARG(n >= 1, "embeddings must have at least one row");
  LAKER_CHECK_ARG(d >= 1 && d <= 32, "embedding_dim must be in [1, 32]");
  LAKER_CHECK_ARG(dtype == LAKER_F32 || dtype == LAKER_F64, "dtype must be LAKER_F32 or LAKER_F64");
  LAKER_CHECK_ARG(clamp_policy >= 0 && clamp_policy <= 2, "bad clamp policy");
  Operator* op = new Operator();
  op->n = n;
  op->d = d;
  op->dtype = dtype;
  op->lambda_reg = lambda_reg;
  op->scale = scale;
  op->clamp_policy = clamp_policy;
  op->clamp_limit = dtype == LAKER_F32 ? 80.0 : 700.0;
  op->ka32 = (int)round_up(d, 2);
  op->ka16 = (int)round_up(6 * d, 16);
  op->ka16h = (int)round_up(3 * d, 16);
  int rc = pack_store(op->self, E, dtype, n, d, op->ka32, op->ka16, scale, true, true, stream);
  if (rc == LAKER_OK) rc = pack_f16_split(op->self, E, dtype, n, d, op->ka16h, scale, stream);
  if (rc != LAKER_OK) {
    delete op;
    return rc;
  }
  // skip-clamp predicate: max ||e||^2 (times scale) < limit
  DevBuf mx;
  if ((rc = mx.ensure(8)) != LAKER_OK) {
    delete op;
    return rc;
  }
  hipError_t e = hipMemsetAsync(mx.ptr, 0, 8, stream);
  if (e == hipSuccess) {
    if (dtype == LAKER_F32)
      hipLaunchKernelGGL((max_sqnorm_kernel<float>), dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                         stream, (const float*)E, n, d, mx.as<unsigned long long>());
    else
      hipLaunchKernelGGL((max_sqnorm_kernel<double>), dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                         stream, (const double*)E, n, d, mx.as<unsigned long long>());
    e = hipGetLastError();
  }
  // workgroup map of the symmetric k = 1 product (row block, segment)
  std::vector<int2> smap;
  {
    // segment-major order: consecutive workgroups (spread over the XCDs)
    // work on the same column segment
    const int64_t ntiles = ceil_div(n, 32), nb = ceil_div(n, SYM_RB), nseg = ceil_div(ntiles, SYM_SEG);
    for (int64_t sg = 0; sg < nseg; ++sg)
      for (int64_t b = 0; b < nb && SYM_DT * b < SYM_SEG * (sg + 1); ++b) smap.push_back(int2{(int)b, (int)sg});
  }
  if (e == hipSuccess && (rc = op->sym_map.ensure(smap.size() * sizeof(int2))) != LAKER_OK) {
    delete op;
    return rc;
  }
  op->sym_nwg = (int)smap.size();
  if (e == hipSuccess)
    e = hipMemcpyAsync(op->sym_map.ptr, smap.data(), smap.size() * sizeof(int2), hipMemcpyHostToDevice, stream);
  // workgroup map of the symmetric FP64 k = 1 product (256-row blocks, 2048-column segments)
  std::vector<int2> smap64;
  {
    const int64_t ntiles = ceil_div(n, 16), nb = ceil_div(n, S64_RB), nseg = ceil_div(ntiles, S64_SEG);
    for (int64_t sg = 0; sg < nseg; ++sg)
      for (int64_t b = 0; b < nb && S64_DT * b < S64_SEG * (sg + 1); ++b) smap64.push_back(int2{(int)b, (int)sg});
  }
  if (e == hipSuccess && (rc = op->sym64_map.ensure(smap64.size() * sizeof(int2))) != LAKER_OK) {
    delete op;
    return rc;
  }
  op->sym64_nwg = (int)smap64.size();
  if (e == hipSuccess)
    e = hipMemcpyAsync(op->sym64_map.ptr, smap64.data(), smap64.size() * sizeof(int2), hipMemcpyHostToDevice,
                       stream);
  uint64_t bits = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&bits, mx.ptr, 8, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) {
    delete op;
    return fail(LAKER_EHIP, std::string("HIP error in operator create: ") + hipGetErrorString(e));
  }
  op->max_sq_norm = __builtin_bit_cast(double, bits);
  op->skip_clamp = (op->max_sq_norm * scale) < op->clamp_limit;
  // set-up of the allocation-free products on this device (once per device)
  void* scratch_ptr = nullptr;
  if ((rc = scratch_for_current_device(&scratch_ptr)) != LAKER_OK) {
    delete op;
    return rc;
  }
  // max |e_ij| <= max ||e_i||: both split sides stay below 2^14, far from FP16 overflow
  op->f16s_ok = std::sqrt(op->max_sq_norm) * std::max(1.0, std::fabs(scale * kLog2eD)) < 16384.0;
  *out = reinterpret_cast<laker_operator_t>(op);
  return LAKER_OK;
}

int laker_operator_destroy(laker_operator_t h) {
  delete reinterpret_cast<Operator*>(h);
  return LAKER_OK;
}

int laker_operator_info(laker_operator_t h, int32_t* skip_clamp, double* max_sq_norm) {
  LAKER_CHECK_ARG(h != nullptr, "operator handle is NULL");
  Operator* op = reinterpret_cast<Operator*>(h);
  if (skip_clamp) *skip_clamp = op->skip_clamp ? 1 : 0;
  if (max_sq_norm) *max_sq_norm = op->max_sq_norm;
  return LAKER_OK;
}

int laker_operator_set_lambda(laker_operator_t h, double lambda_reg) {
  LAKER_CHECK_ARG(h != nullptr, "operator handle is NULL");
  reinterpret_cast<Operator*>(h)->lambda_reg = lambda_reg;
  return LAKER_OK;
}

int laker_matvec(laker_operator_t h, const void* X, int32_t k, int64_t ldx, void* Y, int64_t ldy,
                 int32_t precision, hipStream_t stream) {
  LAKER_CHECK_ARG(h != nullptr, "operator handle is NULL");
  LAKER_CHECK_ARG(X != nullptr && Y != nullptr, "X/Y pointer is NULL");
  LAKER_CHECK_ARG(k >= 1 && ldx >= k && ldy >= k, "bad k / leading dimension");
  Operator* op = reinterpret_cast<Operator*>(h);
  const bool clamp = op->clamp_for(LAKER_CLAMP_AUTO);
  Side rows = op->rows(), cols = op->cols();
  if (op->dtype == LAKER_F64 || precision == LAKER_PREC_FP64) {
    LAKER_CHECK_ARG(op->dtype == LAKER_F64, "LAKER_PREC_FP64 matvec needs an FP64 operator; use laker_residual_fp64");
    return fused_product_f64<double>(*op, rows, cols, (const double*)X, k, ldx, (double*)Y, ldy,
                                     op->lambda_reg, (const double*)X, ldx, clamp, nullptr, stream);
  }
  LAKER_CHECK_ARG(precision == LAKER_PREC_FP32 || precision == LAKER_PREC_MIXED ||
                      precision == LAKER_PREC_BF16 || precision == LAKER_PREC_HILO,
                  "bad precision");
  return fused_product_f32(*op, precision, rows, cols, (const float*)X, k, ldx, (float*)Y, ldy,
                           op->lambda_reg, (const float*)X, ldx, clamp, false, 0, stream);
}

int laker_matvec_noalloc(laker_operator_t h, const float* X, int32_t k, int64_t ldx, float* Y, int64_t ldy,
                         int32_t precision, hipStream_t stream) {
  LAKER_CHECK_ARG(h != nullptr, "operator handle is NULL");
  LAKER_CHECK_ARG(X != nullptr && Y != nullptr, "X/Y pointer is NULL");
  LAKER_CHECK_ARG(k >= 1 && k <= 4, "laker_matvec_noalloc takes 1 <= k <= 4 right-hand sides");
  LAKER_CHECK_ARG(ldx >= k && ldy >= k, "bad leading dimension");
  LAKER_CHECK_ARG(precision == LAKER_PREC_FP32 || precision == LAKER_PREC_MIXED,
                  "laker_matvec_noalloc takes LAKER_PREC_FP32 or LAKER_PREC_MIXED");
  Operator* op = reinterpret_cast<Operator*>(h);
  LAKER_CHECK_ARG(op->dtype == LAKER_F32, "laker_matvec_noalloc needs an FP32 operator");
  const size_t need = noalloc_part_bytes(op->n, k);
  if (need > kNoAllocScratchBytes)
    return fail(LAKER_EINVAL, "laker_matvec_noalloc: n = " + std::to_string(op->n) + " with k = " +
                                  std::to_string(k) + " needs " + std::to_string(need) +
                                  " bytes of scratch, above its maximum of " +
                                  std::to_string(kNoAllocScratchBytes) +
                                  " (n <= 262144 for k <= 2, n <= 131072 for k = 3, 4)");
  void* part = nullptr;
  LAKER_TRY(scratch_for_current_device(&part));
  return product_vec_into(*op, precision, (const float*)X, k, ldx, (float*)Y, ldy, part, stream);
}

int laker_residual_fp64(laker_operator_t h, const double* B, const double* X, int32_t k, double* R,
                        hipStream_t stream) {
  LAKER_CHECK_ARG(h != nullptr, "operator handle is NULL");
  LAKER_CHECK_ARG(B && X && R && k >= 1, "bad residual arguments");
  Operator* op = reinterpret_cast<Operator*>(h);
  const bool clamp = op->clamp_for(LAKER_CLAMP_AUTO);
  Side rows = op->rows(), cols = op->cols();
  return fused_product_f64<double>(*op, rows, cols, X, k, k, R, k, op->lambda_reg, X, k, clamp, B,
                                   stream);
}

int laker_cross_matvec(laker_operator_t h, const void* Eq, int64_t m, const void* X, int32_t k,
                       int64_t ldx, void* Y, int64_t ldy, int32_t clamp_policy, int32_t precision,
                       hipStream_t stream) {
  LAKER_CHECK_ARG(h != nullptr, "operator handle is NULL");
  LAKER_CHECK_ARG(Eq && X && Y && m >= 1 && k >= 1, "bad cross_matvec arguments");
  Operator* op = reinterpret_cast<Operator*>(h);
  const bool clamp = op->clamp_for(clamp_policy);
  std::lock_guard<std::mutex> g(op->query_mu);
  LAKER_TRY(pack_store(op->query, Eq, op->dtype, m, op->d, op->ka32, op->ka16, op->scale, false,
                       true, stream));
  Side qrows = Operator::view(op->query, true);
  Side cols = op->cols();
  if (op->dtype == LAKER_F64 || precision == LAKER_PREC_FP64) {
    // FP64 operators, or an FP32 operator's FP64 copy of E with FP64 X and Y
    // (predictions from an FP64-refined solution; the clamp limit stays the
    // operator's)
    return fused_product_f64<double>(*op, qrows, cols, (const double*)X, k, ldx, (double*)Y,
                                     ldy, 0.0, nullptr, 0, clamp, nullptr, stream);
  }
  return fused_product_f32(*op, precision, qrows, cols, (const float*)X, k, ldx, (float*)Y, ldy,
                           0.0, nullptr, 0, clamp, false, 0, stream);
}

int laker_cross_matvec_t(laker_operator_t h, const void* Eq, int64_t m, const void* Xq, int32_t k,
                         int64_t ldx, void* Y, int64_t ldy, int32_t clamp_policy,
                         hipStream_t stream) {
  LAKER_CHECK_ARG(h != nullptr, "operator handle is NULL");
  LAKER_CHECK_ARG(Eq && Xq && Y && m >= 1 && k >= 1, "bad cross_matvec_t arguments");
  Operator* op = reinterpret_cast<Operator*>(h);
  const bool clamp = op->clamp_for(clamp_policy);
  std::lock_guard<std::mutex> g(op->query_mu);
  LAKER_TRY(pack_store(op->query, Eq, op->dtype, m, op->d, op->ka32, op->ka16, op->scale, true,
                       false, stream));
  Side qcols = Operator::view(op->query, false);
  Side rows = op->rows();
  if (op->dtype == LAKER_F64)
    return fused_product_f64<double>(*op, rows, qcols, (const double*)Xq, k, ldx, (double*)Y,
                                     ldy, 0.0, nullptr, 0, clamp, nullptr, stream);
  return fused_product_f32(*op, LAKER_PREC_FP32, rows, qcols, (const float*)Xq, k, ldx,
                           (float*)Y, ldy, 0.0, nullptr, 0, clamp, false, 0, stream);
}

int laker_kernel_eval(const void* X, int64_t m, const void* Y, int64_t p, int32_t d, int32_t dtype,
                      double scale, int32_t clamp, void* K, int64_t ldk, hipStream_t stream) {
  LAKER_CHECK_ARG(X && Y && K && m >= 1 && p >= 1 && d >= 1 && ldk >= p, "bad kernel_eval arguments");
  dim3 grid = eval_grid(m * p);
  if (dtype == LAKER_F32)
    hipLaunchKernelGGL((kernel_eval_kernel<float>), grid, dim3(256), 0, stream, (const float*)X, m,
                       (const float*)Y, p, d, scale, clamp, 80.0, 0.0, (float*)K, ldk);
  else
    hipLaunchKernelGGL((kernel_eval_kernel<double>), grid, dim3(256), 0, stream, (const double*)X, m,
                       (const double*)Y, p, d, scale, clamp, 700.0, 0.0, (double*)K, ldk);
  LAKER_KCHECK("kernel_eval_kernel");
  return LAKER_OK;
}

int laker_diagonal(laker_operator_t h, void* out, hipStream_t stream) {
  LAKER_CHECK_ARG(h != nullptr && out != nullptr, "bad diagonal arguments");
  Operator* op = reinterpret_cast<Operator*>(h);
  if (op->dtype == LAKER_F32)
    hipLaunchKernelGGL((diag_kernel<float>), dim3((unsigned)ceil_div(op->n, 256)), dim3(256), 0,
                       stream, op->e_user32(), op->n, op->d, op->scale, op->lambda_reg, (float*)out);
  else
    hipLaunchKernelGGL((diag_kernel<double>), dim3((unsigned)ceil_div(op->n, 256)), dim3(256), 0,
                       stream, op->self.e64.as<double>(), op->n, op->d, op->scale, op->lambda_reg,
                       (double*)out);
  LAKER_KCHECK("diag_kernel");
  return LAKER_OK;
}

int laker_to_dense(laker_operator_t h, void* out, hipStream_t stream) {
  LAKER_CHECK_ARG(h != nullptr && out != nullptr, "bad to_dense arguments");
  Operator* op = reinterpret_cast<Operator*>(h);
  dim3 grid = eval_grid(op->n * op->n);
  if (op->dtype == LAKER_F32)
    hipLaunchKernelGGL((kernel_eval_kernel<float>), grid, dim3(256), 0, stream, op->e_user32(),
                       op->n, op->e_user32(), op->n, op->d, op->scale, 0, 80.0, op->lambda_reg,
                       (float*)out, op->n);
  else
    hipLaunchKernelGGL((kernel_eval_kernel<double>), grid, dim3(256), 0, stream,
                       op->self.e64.as<double>(), op->n, op->self.e64.as<double>(), op->n, op->d,
                       op->scale, 0, 700.0, op->lambda_reg, (double*)out, op->n);
  LAKER_KCHECK("kernel_eval_kernel");
  LAKER_HIP_CHECK(hipGetLastError());
  return LAKER_OK;
}

}  // extern "C"
