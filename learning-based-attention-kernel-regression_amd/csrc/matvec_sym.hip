// Symmetric k = 1 product of the square operator (mv_sym1_kernel) and its
// reduction (mv_sym_reduce_kernel).
#include "matvec_tiles.h"

#include <algorithm>
#include <cmath>

namespace laker {

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

#ifndef LAKER_SYM_RACC
#define LAKER_SYM_RACC 0
#endif
#ifndef LAKER_SYM_ALL
#define LAKER_SYM_ALL 0
#endif
#ifndef LAKER_SYM_UNROLLTT
#define LAKER_SYM_UNROLLTT 0
#endif

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
#if LAKER_SYM_RACC
      // row sums of the chunk in FP32 (two chains per row tile), FP64 per chunk
      float racc[ITW][2];
#pragma unroll
      for (int it = 0; it < ITW; ++it) racc[it][0] = racc[it][1] = 0.f;
#endif
      auto tile = [&](int tt) __attribute__((always_inline)) {
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
        float q[16];
#if LAKER_SYM_ALL
        // all row tiles' Gram tiles first: later MFMAs run under earlier exps
        floatx16 D[ITW];
#pragma unroll
        for (int it = 0; it < ITW; ++it) D[it] = G::template gram_exact<NS>(A, B[it]);
#pragma unroll
        for (int it = 0; it < ITW; ++it) {
#pragma unroll
          for (int v = 0; v < 16; ++v) D[it][v] = kexp<CLAMP, false>(D[it][v], clampv);
#if LAKER_SYM_RACC
          float a0 = racc[it][0], a1 = racc[it][1];
#pragma unroll
          for (int v = 0; v < 16; v += 2) {
            a0 = __builtin_fmaf(D[it][v], xv[v], a0);
            a1 = __builtin_fmaf(D[it][v + 1], xv[v + 1], a1);
          }
          racc[it][0] = a0;
          racc[it][1] = a1;
#else
          float a0 = D[it][0] * xv[0], a1 = D[it][1] * xv[1];
#pragma unroll
          for (int v = 2; v < 16; v += 2) {
            a0 = __builtin_fmaf(D[it][v], xv[v], a0);
            a1 = __builtin_fmaf(D[it][v + 1], xv[v + 1], a1);
          }
          acc64[it] += (double)(a0 + a1);
#endif
#pragma unroll
          for (int v = 0; v < 16; ++v) q[v] = it == 0 ? D[0][v] * xr[0] : __builtin_fmaf(D[it][v], xr[it], q[v]);
        }
#else
        // row tiles in pairs: two Gram tiles live at a time
#pragma unroll
        for (int p = 0; p < ITW / 2; ++p) {
          floatx16 D[2];
#pragma unroll
          for (int u = 0; u < 2; ++u) D[u] = G::template gram_exact<NS>(A, B[2 * p + u]);
#pragma unroll
          for (int u = 0; u < 2; ++u) {
#pragma unroll
            for (int v = 0; v < 16; ++v) D[u][v] = kexp<CLAMP, false>(D[u][v], clampv);
#if LAKER_SYM_RACC
            float a0 = racc[2 * p + u][0], a1 = racc[2 * p + u][1];
#pragma unroll
            for (int v = 0; v < 16; v += 2) {
              a0 = __builtin_fmaf(D[u][v], xv[v], a0);
              a1 = __builtin_fmaf(D[u][v + 1], xv[v + 1], a1);
            }
            racc[2 * p + u][0] = a0;
            racc[2 * p + u][1] = a1;
#else
            float a0 = D[u][0] * xv[0], a1 = D[u][1] * xv[1];
#pragma unroll
            for (int v = 2; v < 16; v += 2) {
              a0 = __builtin_fmaf(D[u][v], xv[v], a0);
              a1 = __builtin_fmaf(D[u][v + 1], xv[v + 1], a1);
            }
            acc64[2 * p + u] += (double)(a0 + a1);
#endif
          }
          // column-side products of the row tiles, summed in-lane
#pragma unroll
          for (int v = 0; v < 16; ++v)
            q[v] = __builtin_fmaf(D[1][v], xr[2 * p + 1],
                                  p == 0 ? D[0][v] * xr[0] : __builtin_fmaf(D[0][v], xr[2 * p], q[v]));
        }
#endif
        if (tc + tt >= tdiag) {  // uniform across the block
          const float cs = transpose_reduce16(q, lane);
          if ((lane & 1) == 0) cb[w * (V1_JT * 32) + tt * 32 + jq] = cs;
        }
      };
#if LAKER_SYM_UNROLLTT
      if (nt == V1_JT) {
#pragma unroll
        for (int tt = 0; tt < V1_JT; ++tt) tile(tt);
      } else {
        for (int tt = 0; tt < nt; ++tt) tile(tt);
      }
#else
      for (int tt = 0; tt < nt; ++tt) tile(tt);
#endif
#if LAKER_SYM_RACC
#pragma unroll
      for (int it = 0; it < ITW; ++it) acc64[it] += (double)(racc[it][0] + racc[it][1]);
#endif
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

// Symmetric k = 1 product of the square operator (see mv_sym1_kernel) plus
// its reduction into Y.  Returns 1 when the shape has no instance.
int launch_sym1(Operator& op, int mode, const void* colpack, const void* rowpack, int ka, int64_t n,
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

}  // namespace laker
