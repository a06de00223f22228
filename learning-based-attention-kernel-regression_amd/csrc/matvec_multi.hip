// Multi-RHS products on MFMA: mv_multi_kernel (FP32 right-hand sides),
// pack_x_kernel + mv_multi2_kernel (BF16-split right-hand sides).
#include "matvec_tiles.h"
#include "batch.h"

#include <algorithm>
#include <cmath>

namespace laker {

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
__device__ __forceinline__ void pack_x_kernel_body(const float* __restrict__ X, int64_t ldx, int64_t n_cols, int kcols,
                              int c_base, int64_t ntiles, int nq, bf16x8* __restrict__ xh,
                              bf16x8* __restrict__ xl, bf16x8* __restrict__ xm, int bx_, int by_) {
  const int64_t idx = (int64_t)bx_ * blockDim.x + threadIdx.x;
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

__global__ void pack_x_kernel(const float* __restrict__ X, int64_t ldx, int64_t n_cols, int kcols,
                              int c_base, int64_t ntiles, int nq, bf16x8* __restrict__ xh,
                              bf16x8* __restrict__ xl, bf16x8* __restrict__ xm) {
  pack_x_kernel_body(X, ldx, n_cols, kcols, c_base, ntiles, nq, xh, xl, xm, (int)blockIdx.x, (int)blockIdx.y);
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
// chunk) and shared by the block's NW waves.  GM: Gram operands, the BF16
// 3-way split (LAKER_PREC_MIXED) or the FP16 2-way split (GRAM_F16S, square
// operator only).
template <int GM, int NS, int GX, int NCC, int JT, int ITW, bool CLAMP, int NW = 4>
__device__ __forceinline__ void mv_multi2_kernel_body(
    const void* __restrict__ colpack, const void* __restrict__ rowpack, int ka, int64_t n_rows,
    int64_t n_cols, const bf16x8* __restrict__ xh, const bf16x8* __restrict__ xl,
    const bf16x8* __restrict__ xm, int kcols, int c_base, float* __restrict__ part, int64_t ldp,
    int jsplit, float clampv, int bx_, int by_) {
  using G = GramOps<GM>;
  constexpr int NP = M2Parts<GX>::value;
  constexpr int gb = 32 * NS, lb = gb + 16, vpr = gb / 16;
  constexpr int VA = JT * 32 * vpr;           // operand-row vectors per chunk
  constexpr int XT = NP * NCC * 2 * 64;       // X vectors per tile
  constexpr int VX = JT * XT;
  constexpr int NT = 64 * NW;                 // threads of the block
  constexpr int VPT = (VA + VX + NT - 1) / NT;  // vectors per thread per chunk
  constexpr int CHUNK = m2_chunk_bytes(NS, NP, NCC, JT);
  extern __shared__ __align__(16) char m2_smem[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, h = lane >> 5, li = lane & 31;
  const int js = bx_ % jsplit;
  const int64_t ib = bx_ / jsplit;
  const int64_t I0 = (ib * NW + w) * (32 * ITW);
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
  auto fetch = [&](int64_t t, uint4* dst) __attribute__((always_inline)) {
    const int nt = (int)((t1 - t) < JT ? (t1 - t) : JT);
#pragma unroll
    for (int u2 = 0; u2 < VPT; ++u2) {
      const int vi = u2 * NT + threadIdx.x;
      uint4 v = {0u, 0u, 0u, 0u};
      if (vi < VA) {
        if (vi < nt * 32 * vpr) v = reinterpret_cast<const uint4*>(gsrc + t * 32 * gb)[vi];
      } else if (vi < VA + VX) {
        const int xi = vi - VA;
        const int tt = xi / XT, rem = xi % XT;
        const int pp = rem / (NCC * 2 * 64), off = rem % (NCC * 2 * 64);  // off = (q*2+u)*64 + lane
        if (tt < nt) v = xsrc[pp][(t + tt) * (NCC * 128) + off];
      }
      dst[u2] = v;
    }
  };
  auto stash = [&](char* dst) __attribute__((always_inline)) {
#pragma unroll
    for (int u2 = 0; u2 < VPT; ++u2) {
      const int vi = u2 * NT + threadIdx.x;
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

  fetch(t0, pre);
  stash(m2_smem);
  __syncthreads();
  bool odd = false;
  for (int64_t tc = t0; tc < t1; tc += JT) {
    const bool more = tc + JT < t1;
    if (more) fetch(tc + JT, pre);
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
        if constexpr (GX == GX_PLXH && ITW == 1) {
          // BF16-exact X, exp tile split hi/lo: 4 NCC products per tile.  The
          // X operands of the first K half are read before the Gram, the exps
          // and conversions of the second half run in the shadow of the first
          // half's MFMAs, and the second half's X operands are read meanwhile,
          // so neither the VALU work nor the LDS latency stalls the MFMA pipe.
          // Per accumulator the products keep the order lo.X, hi.X of K half
          // 0, then of K half 1.
          bf16x8 X0[NCC], X1[NCC];
#pragma unroll
          for (int q = 0; q < NCC; ++q) X0[q] = __builtin_bit_cast(bf16x8, xs[tt * XT + (q * 2) * 64 + lane]);
          __builtin_amdgcn_sched_barrier(0);
          const floatx16 D = G::template gram_exact<NS>(A, B[0]);
          bf16x8 Ph[2], Pl[2];
          auto split = [&](int u) __attribute__((always_inline)) {
#pragma unroll
            for (int v = 0; v < 8; ++v) {
              const float pv = kexp<CLAMP, false>(D[8 * u + v], clampv);
              const uint16_t hb = f32_to_bf16_bits(pv);
              Ph[u][v] = __builtin_bit_cast(__bf16, hb);
              Pl[u][v] = __builtin_bit_cast(__bf16, f32_to_bf16_bits(pv - bf16_bits_to_f32(hb)));
            }
          };
          split(0);
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int q = 0; q < NCC; ++q) {
            acc[0][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Pl[0], X0[q], acc[0][q], 0, 0, 0);
            acc[0][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ph[0], X0[q], acc[0][q], 0, 0, 0);
          }
#pragma unroll
          for (int q = 0; q < NCC; ++q) X1[q] = __builtin_bit_cast(bf16x8, xs[tt * XT + (q * 2 + 1) * 64 + lane]);
          split(1);
#pragma unroll
          for (int q = 0; q < 2 * NCC; ++q) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
            __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);  // two VALU
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // one LDS read
          }
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int q = 0; q < NCC; ++q) {
            acc[0][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Pl[1], X1[q], acc[0][q], 0, 0, 0);
            acc[0][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ph[1], X1[q], acc[0][q], 0, 0, 0);
          }
          continue;
        }
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

template <int GM, int NS, int GX, int NCC, int JT, int ITW, bool CLAMP, int NW = 4>
__global__ __launch_bounds__(64 * NW) void mv_multi2_kernel(
    const void* __restrict__ colpack, const void* __restrict__ rowpack, int ka, int64_t n_rows,
    int64_t n_cols, const bf16x8* __restrict__ xh, const bf16x8* __restrict__ xl,
    const bf16x8* __restrict__ xm, int kcols, int c_base, float* __restrict__ part, int64_t ldp,
    int jsplit, float clampv) {
  mv_multi2_kernel_body<GM, NS, GX, NCC, JT, ITW, CLAMP, NW>(colpack, rowpack, ka, n_rows, n_cols, xh, xl, xm,
                                                             kcols, c_base, part, ldp, jsplit, clampv,
                                                             (int)blockIdx.x, (int)blockIdx.y);
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

// Y (n_rows x k, k > 4) = [lam*Xr] + exp(min(s*Erow.Ecol^T, c)) X on MFMA.
int product_multi_f32(Operator& op, int precision, const Side& rows, const Side& cols, const float* X, int k,
                      int64_t ldx, float* Y, int64_t ldy, double lam, const float* Xr, int64_t ldxr, bool clamp,
                      bool deriv, int accumulate, hipStream_t st) {
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
    /* 8-chunk passes: 8-wave blocks, so each block streams the packed X once for 256 rows */        \
    constexpr int NWV = NC >= 8 ? 8 : 4;                                                              \
    dim3 grid2((unsigned)(ceil_div(n_rows, 32 * NWV * ITWV) * jsplit2)), block2(64 * NWV);            \
    if (f16s) {                                                                                       \
      if (clamp)                                                                                      \
        hipLaunchKernelGGL((mv_multi2_kernel<GRAM_F16S, NS, GXM, NC, JTV, ITWV, true, NWV>), grid2,    \
                           block2, lds2, st, op.self.col16h.ptr, op.self.row16h.ptr, op.ka16h, n_rows, \
                           n_cols, (const bf16x8*)xh, (const bf16x8*)xl, (const bf16x8*)xm, k, c0,    \
                           part, kp, jsplit2, clampv);                                                \
      else                                                                                            \
        hipLaunchKernelGGL((mv_multi2_kernel<GRAM_F16S, NS, GXM, NC, JTV, ITWV, false, NWV>), grid2,   \
                           block2, lds2, st, op.self.col16h.ptr, op.self.row16h.ptr, op.ka16h, n_rows, \
                           n_cols, (const bf16x8*)xh, (const bf16x8*)xl, (const bf16x8*)xm, k, c0,    \
                           part, kp, jsplit2, clampv);                                                \
    } else if (clamp)                                                                                 \
      hipLaunchKernelGGL((mv_multi2_kernel<LAKER_PREC_MIXED, NS, GXM, NC, JTV, ITWV, true, NWV>), grid2, \
                         block2, lds2, st, colpack, rowpack, ka, n_rows, n_cols, (const bf16x8*)xh,   \
                         (const bf16x8*)xl, (const bf16x8*)xm, k, c0, part, kp, jsplit2, clampv);     \
    else                                                                                              \
      hipLaunchKernelGGL((mv_multi2_kernel<LAKER_PREC_MIXED, NS, GXM, NC, JTV, ITWV, false, NWV>),     \
                         grid2, block2, lds2, st, colpack, rowpack, ka, n_rows, n_cols,               \
                         (const bf16x8*)xh, (const bf16x8*)xl, (const bf16x8*)xm, k, c0, part, kp,    \
                         jsplit2, clampv);                                                            \
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

// ---------------------------------------------------------------------------
// Batched launches (several systems per launch, see batch.h): the pack, the
// GX_SPLIT3 multi-RHS product and its reduction, with each system's block
// decomposition exactly that of product_multi_f32 for one 32-column chunk.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void pack_x_batched(const BPackX* __restrict__ desc, const int* __restrict__ off,
                                                      int nsys) {
  const int s = batch_system(off, nsys, (int)blockIdx.x);
  const BPackX d = desc[s];
  pack_x_kernel_body(d.X, d.ldx, d.n_cols, d.kcols, d.c_base, d.ntiles, d.nq, d.xh, d.xl, d.xm,
                     (int)blockIdx.x - off[s], 0);
}

template <int GM, int NS, int JT, bool CLAMP>
__global__ __launch_bounds__(256) void mv_multi2_batched(const BMulti2* __restrict__ desc, const int* __restrict__ off,
                                                         int nsys) {
  const int s = batch_system(off, nsys, (int)blockIdx.x);
  const BMulti2 d = desc[s];
  mv_multi2_kernel_body<GM, NS, GX_SPLIT3, 1, JT, 1, CLAMP>(d.colpack, d.rowpack, d.ka, d.n_rows, d.n_cols, d.xh,
                                                             d.xl, d.xm, d.kcols, d.c_base, d.part, d.ldp, d.jsplit,
                                                             d.clampv, (int)blockIdx.x - off[s], 0);
}

__global__ void reduce_batched(const BReduce* __restrict__ desc, const int* __restrict__ off, int nsys) {
  const int s = batch_system(off, nsys, (int)blockIdx.x);
  const BReduce d = desc[s];
  mv_reduce_kernel_body<float, float, float, false>(d.part, d.n_rows, d.k, d.ldp, d.jsplit, d.lam, d.Xr, d.ldx,
                                                    (const float*)nullptr, d.Y, d.ldy, 0, (const int*)nullptr,
                                                    (int)blockIdx.x - off[s], 0);
}

int multi2_split3_jt(int /*gm*/, int ns) { return 2 * m2_chunk_bytes(ns, 3, 1, 2) <= 65536 ? 2 : 1; }
size_t multi2_split3_lds(int ns, int jt) { return 2 * (size_t)m2_chunk_bytes(ns, 3, 1, jt); }

int batched_pack_x(const BatchLaunch<BPackX>& L, hipStream_t st) {
  if (L.nsys <= 0) return LAKER_OK;
  hipLaunchKernelGGL(pack_x_batched, dim3((unsigned)L.blocks), dim3(256), 0, st, L.desc, L.off, L.nsys);
  LAKER_KCHECK("pack_x_batched");
  return LAKER_OK;
}

int batched_multi2(const BatchLaunch<BMulti2>& L, int gm, int ns, bool clamp, hipStream_t st) {
  if (L.nsys <= 0) return LAKER_OK;
  const int jt = multi2_split3_jt(gm, ns);
  const size_t lds = multi2_split3_lds(ns, jt);
  const dim3 grid((unsigned)L.blocks), block(256);
#define LAKER_B2(GM_, NS_, JT_)                                                                          \
  do {                                                                                                 \
    if (clamp)                                                                                         \
      hipLaunchKernelGGL((mv_multi2_batched<GM_, NS_, JT_, true>), grid, block, lds, st, L.desc, L.off, L.nsys); \
    else                                                                                               \
      hipLaunchKernelGGL((mv_multi2_batched<GM_, NS_, JT_, false>), grid, block, lds, st, L.desc, L.off, L.nsys); \
  } while (0)
#define LAKER_B2J(GM_, NS_)                 \
  do {                                      \
    if (jt == 2) LAKER_B2(GM_, NS_, 2);     \
    else LAKER_B2(GM_, NS_, 1);             \
  } while (0)
  if (gm == GRAM_F16S) {
    if (ns == 2) LAKER_B2J(GRAM_F16S, 2);
    else if (ns == 3) LAKER_B2J(GRAM_F16S, 3);
    else return fail(LAKER_EINVAL, "batched product: no FP16-split instance for this embedding width");
  } else {
    if (ns == 2) LAKER_B2J(LAKER_PREC_MIXED, 2);
    else if (ns == 3) LAKER_B2J(LAKER_PREC_MIXED, 3);
    else if (ns == 4) LAKER_B2J(LAKER_PREC_MIXED, 4);
    else if (ns == 6) LAKER_B2J(LAKER_PREC_MIXED, 6);
    else return fail(LAKER_EINVAL, "batched product: no BF16-split instance for this embedding width");
  }
#undef LAKER_B2J
#undef LAKER_B2
  LAKER_KCHECK("mv_multi2_batched");
  return LAKER_OK;
}

int batched_reduce(const BatchLaunch<BReduce>& L, hipStream_t st) {
  if (L.nsys <= 0) return LAKER_OK;
  hipLaunchKernelGGL(reduce_batched, dim3((unsigned)L.blocks), dim3(256), 0, st, L.desc, L.off, L.nsys);
  LAKER_KCHECK("reduce_batched");
  return LAKER_OK;
}

}  // namespace laker

