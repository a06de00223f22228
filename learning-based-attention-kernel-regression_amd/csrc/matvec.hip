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

#include "matvec_tiles.h"
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
  return product_multi_f32(op, precision, rows, cols, X, k, ldx, Y, ldy, lam, Xr, ldxr, clamp, deriv,
                           accumulate, st);
}

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
  if ((rc = mx.ensure(8, stream)) != LAKER_OK) {
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
    // work on the same column segment (each XCD taking a contiguous run of
    // the list measured 1 % slower)
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
