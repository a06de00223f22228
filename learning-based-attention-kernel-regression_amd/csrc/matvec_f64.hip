// FP64 products: the refinement residual (mv_sym64_kernel) and the FP64
// operators' products (mv_f64m_kernel, mv_f64_kernel).
#include "matvec_tiles.h"

#include <algorithm>
#include <cmath>

namespace laker {

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

}  // namespace laker
