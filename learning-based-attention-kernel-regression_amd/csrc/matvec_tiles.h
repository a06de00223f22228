// Device building blocks shared by the fused-product translation units
// (matvec.hip, matvec_sym.hip, matvec_multi.hip, matvec_f64.hip): MFMA Gram
// tiles, the exp2 element transform, tile and partial-sum layouts, the
// split-K reduction, and the launchers each unit exports.
#pragma once
#include "laker_common.h"
#include "laker_internal.h"

namespace laker {

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

// Layout constants of the symmetric k = 1 product (matvec_sym.hip); the
// operator handle builds its workgroup map from them.
constexpr int SYM_SEG = 32;     // column tiles per workgroup (1024 columns; 512 and 2048 measured
                                // equal or slower: load balance against row-partial traffic)
constexpr int SYM_WAVES = 8;    // waves per workgroup
constexpr int SYM_ITW = 2;      // row tiles per wave
constexpr int SYM_JT = 4;       // column tiles per LDS chunk
constexpr int SYM_RB = SYM_WAVES * 32 * SYM_ITW;  // rows per workgroup (512)
constexpr int SYM_DT = SYM_RB / 32;               // column tiles of the diagonal block

__host__ __device__ inline int64_t sym_coff(int64_t b, int64_t n) { return b * n - (SYM_RB / 2) * b * (b + 1); }

// Layout constants of the symmetric FP64 product (matvec_f64.hip).
constexpr int S64_RB = 256;             // rows per workgroup
constexpr int S64_SEG = 128;            // 16-column tiles per segment
constexpr int S64_DT = S64_RB / 16;     // tiles of the diagonal block
constexpr int S64_CH = 8;               // tiles per column-sum chunk (one barrier pair)

__host__ __device__ inline int64_t s64_coff(int64_t b, int64_t n) { return b * n - (S64_RB / 2) * b * (b + 1); }

// MODE_RES: Y = B - lam*Xr - sum   (FP64 residual)
template <class PT, class XT, class YT, bool RES>
__device__ __forceinline__ void mv_reduce_kernel_body(const PT* __restrict__ part, int64_t n_rows, int k, int64_t ldp,
                                 int jsplit, double lam, const XT* __restrict__ Xr, int64_t ldx,
                                 const YT* __restrict__ B, YT* __restrict__ Y, int64_t ldy,
                                 int accumulate, const int* __restrict__ skip, int bx_, int by_) {
  if (skip && *skip) return;
  const int64_t idx = (int64_t)bx_ * blockDim.x + threadIdx.x;
  const int64_t tot = n_rows * k;
  if (idx >= tot) return;
  int64_t i;
  int c;
  if (tot <= 0x7fffffff) {  // 32-bit index arithmetic
    const uint32_t u = (uint32_t)idx, q = u / (uint32_t)k;
    i = q;
    c = (int)(u - q * (uint32_t)k);
  } else {
    i = idx / k;
    c = (int)(idx % k);
  }
  // sum over the splits in order, eight loads in flight
  const PT* p = part + i * ldp + c;
  const int64_t pstride = n_rows * ldp;
  double s = 0.0;
  int js = 0;
  for (; js + 8 <= jsplit; js += 8) {
    PT v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = p[(int64_t)(js + u) * pstride];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += (double)v[u];
  }
  for (; js < jsplit; ++js) s += (double)p[(int64_t)js * pstride];
  if (Xr) s += lam * (double)Xr[i * ldx + c];
  if constexpr (RES) {
    Y[i * ldy + c] = (YT)((double)B[i * ldy + c] - s);
  } else {
    if (accumulate) s += (double)Y[i * ldy + c];
    Y[i * ldy + c] = (YT)s;
  }
}

template <class PT, class XT, class YT, bool RES>
__global__ void mv_reduce_kernel(const PT* __restrict__ part, int64_t n_rows, int k, int64_t ldp,
                                 int jsplit, double lam, const XT* __restrict__ Xr, int64_t ldx,
                                 const YT* __restrict__ B, YT* __restrict__ Y, int64_t ldy,
                                 int accumulate, const int* __restrict__ skip) {
  mv_reduce_kernel_body<PT, XT, YT, RES>(part, n_rows, k, ldp, jsplit, lam, Xr, ldx, B, Y, ldy, accumulate, skip, (int)blockIdx.x, (int)blockIdx.y);
}

// Column splits of a split-K product: enough workgroups to fill the chip.
inline int pick_jsplit(int64_t n_rows, int64_t n_cols, int rows_per_block) {
  const int64_t blocks = ceil_div(n_rows, rows_per_block);
  const int64_t tiles = ceil_div(n_cols, 32);
  int64_t js = 8;
  // aim for >= ~2048 blocks of 4 waves; at least 2 column tiles per split
  while (blocks * js < 2048 && js < 64) js *= 2;
  while (js > 1 && tiles / js < 2) js /= 2;
  return (int)js;
}

// Symmetric k = 1 product of the square operator plus its reduction into Y
// (matvec_sym.hip).  Returns 1 when the shape has no instance.
int launch_sym1(Operator& op, int mode, const void* colpack, const void* rowpack, int ka, int64_t n,
                const float* X, int64_t ldx, float* Y, int64_t ldy, double lam, const float* Xr,
                int64_t ldxr, bool clamp, float clampv, int accumulate, const int* skip,
                hipStream_t st, double* dots);

// Multi-RHS (k > 4) part of fused_product_f32 on MFMA (matvec_multi.hip).
int product_multi_f32(Operator& op, int precision, const Side& rows, const Side& cols, const float* X, int k,
                      int64_t ldx, float* Y, int64_t ldy, double lam, const float* Xr, int64_t ldxr, bool clamp,
                      bool deriv, int accumulate, hipStream_t st);

}  // namespace laker
