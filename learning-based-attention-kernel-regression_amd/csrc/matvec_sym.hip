// Symmetric k = 1 product of the square operator (mv_sym1_kernel) and its
// reduction (mv_sym_reduce_kernel).
#include "matvec_tiles.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <tuple>
#include <vector>

namespace laker {

// ---------------------------------------------------------------------------
// Symmetric k = 1 product of the square operator.  Workgroup (b, seg) owns
// rows [SYM_RB b, SYM_RB (b + 1)) (SYM_WAVES waves x SYM_ITW row tiles of 32)
// and the column tiles of segment seg, [SYM_SEG seg, SYM_SEG (seg + 1)), at
// or right of tile SYM_DT b: only tiles at or right of the diagonal
// SYM_RB x SYM_RB block are visited, so every exponential is evaluated once
// for the pair (i, j), (j, i).  Each exp tile P feeds the row sums
// y_I += P x_J and, right of the diagonal block, the column sums
// y_J += P^T x_I.  The Gram tile is formed as D = B_row A_col^T, so a lane
// holds one column (li) and 16 rows (jj_of(v, h)) of each 32 x 32 tile:
// * column sums are in-lane FMAs over the wave's rows; both half-waves (same
//   column, other rows) store their sums in LDS, a fixed-order sum over the
//   waves (each wave's two halves first) gives one FP32 partial per (row
//   block, column);
// * row sums accumulate per (row, lane column) in FP32 registers over the
//   whole segment (one FMA per entry, 32 independent chains per lane), and
//   one 32-lane transpose-reduce per row tile at the end of the segment
//   gives one FP32 partial per (segment, row).
// Per pair of entries that is one v_exp_f32, one row FMA and one column
// FMA, the VALU-issue floor of the scheme, plus a few instructions per
// 32 x 32 tile.  Row partials: part_r[seg][i]; column partials:
// part_c[coff(b) + j - SYM_RB (b + 1)] for j >= SYM_RB (b + 1),
// coff(b) = b n - (SYM_RB / 2) b (b + 1).
// ---------------------------------------------------------------------------

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


using SymPart = float;  // partial sums of the symmetric product

// LDS of a workgroup of the symmetric product (bytes) for chunks of jt column
// tiles: the double-buffered column chunk (operand rows of pitch lb, then x),
// the per-tile column sums of the waves (double-buffered; `halves` = 2 keeps
// the two half-waves' sums apart, 1 joins them in registers first) and x of
// the block's rows.
__host__ __device__ constexpr int sym1_lds(int lb, int jt, int halves) {
  return 2 * jt * 32 * (lb + 4) + (2 * SYM_WAVES * halves * jt * 32 + SYM_RB) * 4;
}
// Two workgroups (four waves per SIMD) share a CU's 160 KB of LDS.  Chunks of
// 8 column tiles with the half-wave sums apart where that fits, else chunks of
// 4 with the sums apart, else chunks of 4 with the sums joined.
constexpr int SYM_LDS_WG = 81920;
__host__ __device__ constexpr int sym1_jt(int lb) { return sym1_lds(lb, 8, 2) <= SYM_LDS_WG ? 8 : 4; }
__host__ __device__ constexpr int sym1_halves(int lb) {
  return sym1_lds(lb, sym1_jt(lb), 2) <= SYM_LDS_WG ? 2 : 1;
}
// Waves per SIMD the register allocation must allow: 8-wave workgroups
// occupy two waves per SIMD each, so a second workgroup per CU needs the
// kernel within 128 VGPRs (four waves per SIMD; measured 0.77 ms at
// n = 100,000 against 0.94 ms at two); the instances whose operands do not
// fit keep the compiler's choice.
template <int MODE, int NS>
constexpr int sym1_min_waves() {
  return (MODE == GRAM_F16S && NS <= 2) || (MODE == LAKER_PREC_FP32 && NS <= 5) ? 4 : 1;
}

template <int MODE, int NS, bool CLAMP>
__global__ __launch_bounds__(64 * SYM_WAVES) __attribute__((amdgpu_waves_per_eu(sym1_min_waves<MODE, NS>()))) void
mv_sym1_kernel(const void* __restrict__ colpack, const void* __restrict__ rowpack, int64_t n,
               const float* __restrict__ X, int64_t ldx, const int2* __restrict__ map, float* __restrict__ part_r,
               float* __restrict__ part_c, float clampv, const int* __restrict__ skip) {
  using G = GramOps<MODE>;
  constexpr int ITW = SYM_ITW, NW = SYM_WAVES, NT = 64 * SYM_WAVES;
  if (skip && *skip) return;  // uniform across the grid
  extern __shared__ __align__(16) char sym_smem[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, h = lane >> 5, li = lane & 31;
  const int2 bs = map[blockIdx.x];
  const int64_t b = bs.x;
  const int64_t R0 = b * SYM_RB;
  const int64_t I0 = R0 + w * (32 * ITW);
  const bool active = I0 < n;
  constexpr int gb = MODE == LAKER_PREC_FP32 ? 8 * NS : 32 * NS;  // bytes of one packed row
  constexpr int lb = MODE == LAKER_PREC_FP32 ? ((gb / 4) % 2 == 0 ? gb + 4 : gb) : gb + 16;
  constexpr int JT = sym1_jt(lb);
  constexpr int chunk = JT * 32 * (lb + 4);
  constexpr int CBH = sym1_halves(lb);
  char* const buf0 = sym_smem;
  char* const buf1 = sym_smem + chunk;
  float* const colbuf = reinterpret_cast<float*>(sym_smem + 2 * chunk);  // [2][NW][CBH][JT*32]
  float* const xrow = colbuf + 2 * NW * CBH * JT * 32;                   // x of the block's rows

  for (int t = threadIdx.x; t < SYM_RB; t += NT) xrow[t] = R0 + t < n ? X[(R0 + t) * ldx] : 0.f;
  typename G::template Operand<NS> B[ITW];
  if (active) {
    // packed rows exist up to round_up(n, 64): a row tile starting at or
    // beyond n (the wave's last tiles in the last row block) stays zero
#pragma unroll
    for (int it = 0; it < ITW; ++it) {
      if (I0 + it * 32 < n)
        G::load(B[it], rowpack, I0 + it * 32 + li, NS * (MODE == LAKER_PREC_FP32 ? 2 : 16), h);
      else
        B[it] = typename G::template Operand<NS>{};
    }
  }
  const int64_t ntiles = (n + 31) / 32;
  const int64_t tdiag = SYM_DT * (b + 1);  // first tile right of the diagonal block
  // segments are aligned to absolute column tiles, so the workgroups of one
  // segment (all row blocks) share its columns in L2
  const int64_t t0 = (SYM_DT * b > (int64_t)bs.y * SYM_SEG) ? SYM_DT * b : (int64_t)bs.y * SYM_SEG;
  const int64_t t1 = ((int64_t)bs.y * SYM_SEG + SYM_SEG < ntiles) ? (int64_t)bs.y * SYM_SEG + SYM_SEG : ntiles;
  const char* gsrc = reinterpret_cast<const char*>(colpack);
  constexpr int NVEC = JT * 32 * gb / 16;  // 16-byte vectors of a full chunk
  constexpr int VPT = (NVEC + NT - 1) / NT;
  uint4 pre[VPT];
  float prex = 0.f;
  auto fetch = [&](int64_t t) {
    const int nt = (int)((t1 - t) < JT ? (t1 - t) : JT);
    const int nvec = nt * 2 * gb;
    const uint4* src = reinterpret_cast<const uint4*>(gsrc + t * 32 * gb);
#pragma unroll
    for (int u = 0; u < VPT; ++u) {
      const int vi = u * NT + threadIdx.x;
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
        const int wi = (u * NT + threadIdx.x) * 4;
        const uint32_t vals[4] = {pre[u].x, pre[u].y, pre[u].z, pre[u].w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int word = wi + q;
          if (word < JT * 32 * wpr) reinterpret_cast<uint32_t*>(dst)[(word / wpr) * lwpr + word % wpr] = vals[q];
        }
      }
    } else {
      constexpr int vpr = gb / 16;
#pragma unroll
      for (int u = 0; u < VPT; ++u) {
        const int vi = u * NT + threadIdx.x;
        if (vi < NVEC) *reinterpret_cast<uint4*>(dst + (vi / vpr) * lb + (vi % vpr) * 16) = pre[u];
      }
    }
    if (threadIdx.x < JT * 32) reinterpret_cast<float*>(dst + JT * 32 * lb)[threadIdx.x] = prex;
  };

  float racc[ITW][16];
#pragma unroll
  for (int it = 0; it < ITW; ++it)
#pragma unroll
    for (int v = 0; v < 16; ++v) racc[it][v] = 0.f;
  fetch(t0);
  stash(buf0);
  __syncthreads();
  bool odd = false;
  for (int64_t tc = t0; tc < t1; tc += JT) {
    const bool more = tc + JT < t1;
    if (more) fetch(tc + JT);
    const int nt = (int)((t1 - tc) < JT ? (t1 - tc) : JT);
    float* cb = colbuf + (odd ? NW * CBH * JT * 32 : 0);
    if (active) {
      const char* cur = odd ? buf1 : buf0;
      const float* xs = reinterpret_cast<const float*>(cur + JT * 32 * lb);
      // Every tile takes row and column sums: the column sums of tiles inside
      // the diagonal block (t < tdiag) and of the zero-filled tiles of a short
      // last chunk are never read.  Zero-filled tiles contribute exp(0) * 0
      // to the row sums.
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
        const float xc = xs[tt * 32 + li];
        float q0 = 0.f, q1 = 0.f;
#pragma unroll
        for (int it = 0; it < ITW; ++it) {
          const floatx16 D = G::template gram_exact<NS>(B[it], A);
          float xv[16];
#pragma unroll
          for (int g = 0; g < 4; ++g) {
            const floatx4 f = *reinterpret_cast<const floatx4*>(xrow + w * (32 * ITW) + it * 32 + 8 * g + 4 * h);
#pragma unroll
            for (int q = 0; q < 4; ++q) xv[4 * g + q] = f[q];
          }
#pragma unroll
          for (int v = 0; v < 16; v += 2) {
            const float e0 = kexp<CLAMP, false>(D[v], clampv), e1 = kexp<CLAMP, false>(D[v + 1], clampv);
            racc[it][v] = __builtin_fmaf(e0, xc, racc[it][v]);
            racc[it][v + 1] = __builtin_fmaf(e1, xc, racc[it][v + 1]);
            q0 = __builtin_fmaf(e0, xv[v], q0);
            q1 = __builtin_fmaf(e1, xv[v + 1], q1);
          }
        }
        const float cs = q0 + q1;
        // lanes l and l ^ 32 hold the same column (other rows)
        if constexpr (CBH == 2) {
          // both half-waves store their sums; the reduction below adds the
          // pair (h = 0 first) before it adds the wave's sum to the others
          cb[(w * 2 + h) * (JT * 32) + tt * 32 + li] = cs;
        } else {
          const auto pr = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(unsigned, cs),
                                                           __builtin_bit_cast(unsigned, cs), false, false);
          const float tot =
              __builtin_bit_cast(float, (unsigned)pr[0]) + __builtin_bit_cast(float, (unsigned)pr[1]);
          if (h == 0) cb[w * (JT * 32) + tt * 32 + li] = tot;
        }
      };
      for (int tt = 0; tt < nt; ++tt) tile(tt);
    }
    if (more) stash(odd ? buf0 : buf1);
    __syncthreads();
    // column partials of this chunk: fixed-order sum over the active waves
    if (threadIdx.x < nt * 32) {
      const int tt = threadIdx.x >> 5;
      const int64_t t = tc + tt;
      const int64_t j = t * 32 + (threadIdx.x & 31);
      if (t >= tdiag && j < n) {
        // only the last row block has idle waves, and it has no tile right of its diagonal block
        constexpr int S = JT * 32;
        float v;
        if constexpr (CBH == 2) {
          v = cb[threadIdx.x] + cb[S + threadIdx.x];
#pragma unroll
          for (int ww = 1; ww < NW; ++ww) v += (cb[2 * ww * S + threadIdx.x] + cb[(2 * ww + 1) * S + threadIdx.x]);
        } else {
          v = cb[threadIdx.x];
#pragma unroll
          for (int ww = 1; ww < NW; ++ww) v += cb[ww * S + threadIdx.x];
        }
        part_c[sym_coff(b, n) + j - SYM_RB * (b + 1)] = v;
      }
    }
    odd = !odd;
  }
  if (!active) return;
  // row partials: the 32 lanes of a half-wave hold the same 16 rows of a row tile
  const int vq = ((lane >> 1) & 1) + 2 * ((lane >> 2) & 1) + 4 * ((lane >> 3) & 1) + 8 * ((lane >> 4) & 1);
  const int jq = jj_of(vq, h);
#pragma unroll
  for (int it = 0; it < ITW; ++it) {
    const float s = transpose_reduce16(racc[it], lane);
    const int64_t i = I0 + it * 32 + jq;
    if ((lane & 1) == 0 && i < n) part_r[(int64_t)bs.y * n + i] = s;
  }
}

// y[i] = lam xr[i] (+ y[i]) + sum_seg part_r[seg][i] + sum_{b < i/SYM_RB} part_c[coff(b) + i - SYM_RB (b+1)]
// Block = 256 rows x 4 threads per row (1024 threads).  The partials of row i
// are summed in eight chains u = 0..7 (chain u takes the partials 8 m + u of
// the row list, then of the column list) plus one tail chain for what is
// left after the last full group of eight; thread c of the row runs chains c
// and c + 4, thread 3 the tail as well, and thread 0 adds the nine chain sums
// in a fixed tree.  Spreading the chains over four threads keeps four times
// more loads in flight than one thread per row would, at the same sum order.
// dots (optional): dots[q * gridDim.x + block] = the block's sum of x.y
// (q = 0), x.x (q = 1) and y.y (q = 2) over its 256 rows, with x = Xr and y
// the stored Y, in the order and with the block reduction of the PCG's
// column-dot kernel for 256-row blocks (bit-identical partials).
constexpr int SRED_ROWS = 256;
constexpr int SRED_TPR = 4;  // threads per row

__global__ __launch_bounds__(SRED_ROWS * SRED_TPR) void mv_sym_reduce_kernel(
    const SymPart* __restrict__ part_r, const SymPart* __restrict__ part_c, int64_t n, double lam,
    const float* __restrict__ Xr, int64_t ldxr, float* __restrict__ Y, int64_t ldy, int accumulate,
    const int* __restrict__ skip, double* __restrict__ dots) {
  if (skip && *skip) return;  // uniform over the grid
  __shared__ double csum[9][SRED_ROWS];
  const int r = threadIdx.x & (SRED_ROWS - 1), c = threadIdx.x / SRED_ROWS;
  const int64_t i = (int64_t)blockIdx.x * SRED_ROWS + r;
  if (i < n) {
    const int64_t ntiles = (n + 31) / 32;
    const int64_t b = i / SYM_RB;
    const int64_t nseg = (ntiles + SYM_SEG - 1) / SYM_SEG;
    const int64_t sg0 = (SYM_DT * b) / SYM_SEG;  // first segment of row block b
    const int64_t gr = (nseg - sg0) / 8, gc = b / 8;  // full groups of eight
    double a0 = 0.0, a1 = 0.0;
    const SymPart* pr = part_r + (sg0 + c) * n + i;
#pragma unroll 4
    for (int64_t m = 0; m < gr; ++m) {
      a0 += pr[(8 * m) * n];
      a1 += pr[(8 * m + 4) * n];
    }
#pragma unroll 4
    for (int64_t m = 0; m < gc; ++m) {
      const int64_t q0 = 8 * m + c, q1 = q0 + 4;
      a0 += part_c[sym_coff(q0, n) + i - SYM_RB * (q0 + 1)];
      a1 += part_c[sym_coff(q1, n) + i - SYM_RB * (q1 + 1)];
    }
    csum[c][r] = a0;
    csum[c + 4][r] = a1;
    if (c == SRED_TPR - 1) {
      double tail = 0.0;
      for (int64_t sg = sg0 + 8 * gr; sg < nseg; ++sg) tail += part_r[sg * n + i];
      for (int64_t bb = 8 * gc; bb < b; ++bb) tail += part_c[sym_coff(bb, n) + i - SYM_RB * (bb + 1)];
      csum[8][r] = tail;
    }
  }
  __syncthreads();
  double xd = 0.0, yd = 0.0;
  if (c == 0 && i < n) {
    double s = (((csum[0][r] + csum[1][r]) + (csum[2][r] + csum[3][r])) +
                ((csum[4][r] + csum[5][r]) + (csum[6][r] + csum[7][r]))) +
               csum[8][r];
    if (Xr) s += lam * (double)Xr[i * ldxr];
    if (accumulate) s += (double)Y[i * ldy];
    const float y = (float)s;
    Y[i * ldy] = y;
    if (dots) {
      xd = (double)Xr[i * ldxr];
      yd = (double)y;
    }
  }
  if (dots) {  // uniform; the block sum of the first 256 threads, as block_sum forms it for 4 waves
    __shared__ double red[4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const double v[3] = {xd * yd, xd * xd, yd * yd};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double w = wave_sum(v[q]);
      __syncthreads();
      if (lane == 0 && wid < 4) red[wid] = w;
      __syncthreads();
      if (wid == 0) {
        const double t = wave_sum(lane < 4 ? red[lane] : 0.0);
        if (lane == 0) dots[(int64_t)q * gridDim.x + blockIdx.x] = t;
      }
    }
  }
}

// Lets `kernel` take `bytes` of dynamic LDS above the 64 KB default (once per
// device and kernel).  False when the runtime refuses; the caller then takes
// another product path.
static bool allow_lds(const void* kernel, size_t bytes) {
  static std::mutex mu;
  static std::vector<std::tuple<int, const void*, size_t>> done;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return false;
  std::lock_guard<std::mutex> lock(mu);
  for (const auto& d : done)
    if (std::get<0>(d) == dev && std::get<1>(d) == kernel && std::get<2>(d) >= bytes) return true;
  if (hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  done.emplace_back(dev, kernel, bytes);
  return true;
}

// Symmetric k = 1 product of the square operator (see mv_sym1_kernel) plus
// its reduction into Y.  Returns 1 when the shape has no instance.
int launch_sym1(Operator& op, int mode, const void* colpack, const void* rowpack, int ka, int64_t n,
                       const float* X, int64_t ldx, float* Y, int64_t ldy, double lam, const float* Xr,
                       int64_t ldxr, bool clamp, float clampv, int accumulate, const int* skip,
                       hipStream_t st, double* dots) {
  // the 16-bit modes (BF16 3-way split, FP16 2-way split) share one layout
  const size_t lbytes = mode == LAKER_PREC_FP32 ? V1Layout<LAKER_PREC_FP32>::lbytes(ka)
                                                : V1Layout<LAKER_PREC_MIXED>::lbytes(ka);
  const int lbi = (int)lbytes;
  const size_t lds = (size_t)sym1_lds(lbi, sym1_jt(lbi), sym1_halves(lbi));
  if (lds > SYM_LDS_WG || op.sym_nwg <= 0) return 1;
  const int ns = mode == LAKER_PREC_FP32 ? ka / 2 : ka / 16;
  const bool fp32_ok = ns == 2 || ns == 3 || ns == 4 || ns == 5 || ns == 8 || ns == 16;
  const bool bf16_ok = ns == 1 || ns == 2 || ns == 3 || ns == 4 || ns == 6;
  if (mode == LAKER_PREC_FP32 ? !fp32_ok : !bf16_ok) return 1;
  const int64_t ntiles = ceil_div(n, 32);
  const int64_t nsegmax = ceil_div(ntiles, SYM_SEG);
  const int64_t ncb = ceil_div(n, SYM_RB) - 1;  // row blocks with columns right of their diagonal block
  LAKER_TRY(op.ensure_part((size_t)nsegmax * n * sizeof(SymPart)));
  LAKER_TRY(op.part_c.ensure((size_t)std::max<int64_t>(1, sym_coff(ncb, n)) * sizeof(SymPart)));
  SymPart* part_r = op.part.as<SymPart>();
  SymPart* part_c = op.part_c.as<SymPart>();
  const int2* map = op.sym_map.as<int2>();
  dim3 grid((unsigned)op.sym_nwg), block(64 * SYM_WAVES);
#define LAKER_S1(MODE, NS)                                                                            \
  do {                                                                                                \
    if (lds > 65536 && !allow_lds(clamp ? (const void*)mv_sym1_kernel<MODE, NS, true>                \
                                        : (const void*)mv_sym1_kernel<MODE, NS, false>,               \
                                  lds))                                                               \
      return 1; /* no room granted: the caller takes the split-K product */                           \
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
  hipLaunchKernelGGL(mv_sym_reduce_kernel, dim3((unsigned)ceil_div(n, SRED_ROWS)), dim3(SRED_ROWS * SRED_TPR), 0, st,
                     (const SymPart*)part_r, (const SymPart*)part_c, n, lam, Xr, ldxr, Y, ldy, accumulate, skip,
                     Xr ? dots : nullptr);
  LAKER_KCHECK("mv_sym_reduce_kernel");
  return LAKER_OK;
}

}  // namespace laker
