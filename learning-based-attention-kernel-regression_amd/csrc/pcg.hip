// Device-resident preconditioned conjugate gradient and the mixed-precision
// solve (BF16-split Gram PCG with an FP64 solution and FP64 residual
// replacement).
//
// Restates laker/solvers.py:130-394: zero-RHS early return, warm start with
// r = b - A x0, breakdown test p^T A p <= -eps ||p|| ||Ap|| (RuntimeError in
// the reference, LAKER_EBREAKDOWN here), alpha/beta recurrences (per column for
// k > 1), optional residual replacement every restart_freq iterations, and the
// relative-residual stop ||r||_F / ||b||_F <= tol with iterations = it + 1.
// Scalars live on device in FP64; vectors stay resident; the host only polls a
// status word every `check_every` iterations, and every kernel after the
// stopping point is a no-op.
#include <atomic>
#include <cstring>
#include <memory>
#include <cmath>
#include <vector>

#include "dense64.h"
#include "laker_common.h"
#include "laker_internal.h"
#include "batch.h"
#include "matvec_tiles.h"

namespace laker {

int64_t precond_rows(const struct Precond& pc);
int precond_dtype(const struct Precond& pc);
int precond_apply(struct Precond& pc, const void* X, int k, int64_t ldx, void* Z, int64_t ldz,
                  hipStream_t st);
int precond_prepare_half(struct Precond& pc, hipStream_t st);
int precond_apply_half(struct Precond& pc, const float* X, float* Z, hipStream_t st);

enum { PCG_RUNNING = 0, PCG_CONVERGED = 1, PCG_BREAKDOWN = 2, PCG_MAXITER = 3 };

struct PcgState {
  int32_t done;
  int32_t iterations;  // final count once done
  int32_t it;          // iterations completed so far (device-side, so batches replay unchanged)
  int32_t max_iter;
  double res_norm;
  double rhs_norm;
  double tol;
  double eps;
  double xscale;       // x += xscale * alpha p in the refinement steps
};

// HIP-graph replay of batches of PCG iterations (on by default).
static std::atomic<int> g_use_graphs{1};

constexpr int DOT_TC = 64;   // columns per thread group (k > 1)
constexpr int DOT_NB = 512;  // max partial blocks

// part[q][blk][c] = sum over the block's rows of A_q[i][c] * B_q[i][c]
template <class T, int NQ>
__device__ __forceinline__ void coldots_kernel_body(const T* __restrict__ A0, const T* __restrict__ B0,
                                                      const T* __restrict__ A1, const T* __restrict__ B1,
                                                      const T* __restrict__ A2, const T* __restrict__ B2,
                                                      int64_t n, int k, int64_t rpb,
                                                      double* __restrict__ part, int nblk,
                                                      const PcgState* __restrict__ st, int bx_, int by_) {
  if (st && st->done) return;
  __shared__ double sm[NQ][256];
  const int64_t r0 = (int64_t)bx_ * rpb;
  const int64_t r1 = (r0 + rpb < n) ? r0 + rpb : n;
  if (k == 1) {
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += 256) {
      acc[0] += (double)A0[i] * (double)B0[i];
      if (NQ > 1) acc[1] += (double)A1[i] * (double)B1[i];
      if (NQ > 2) acc[2] += (double)A2[i] * (double)B2[i];
    }
    __shared__ double red[16];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const double s = block_sum(acc[q], red);
      if (threadIdx.x == 0) part[(int64_t)q * nblk + bx_] = s;
    }
    return;
  }
  const int tc = threadIdx.x % DOT_TC, tr = threadIdx.x / DOT_TC;  // 4 row lanes
  for (int c0 = 0; c0 < k; c0 += DOT_TC) {
    const int c = c0 + tc;
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
    if (c < k) {
      for (int64_t i = r0 + tr; i < r1; i += 4) {
        const int64_t o = i * k + c;
        acc[0] += (double)A0[o] * (double)B0[o];
        if (NQ > 1) acc[1] += (double)A1[o] * (double)B1[o];
        if (NQ > 2) acc[2] += (double)A2[o] * (double)B2[o];
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) sm[q][threadIdx.x] = acc[q];
    __syncthreads();
    if (tr == 0 && c < k) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const double s = sm[q][tc] + sm[q][tc + 64] + sm[q][tc + 128] + sm[q][tc + 192];
        part[((int64_t)q * nblk + bx_) * k + c] = s;
      }
    }
  }
}

template <class T, int NQ>
__global__ __launch_bounds__(256) void coldots_kernel(const T* __restrict__ A0, const T* __restrict__ B0,
                                                      const T* __restrict__ A1, const T* __restrict__ B1,
                                                      const T* __restrict__ A2, const T* __restrict__ B2,
                                                      int64_t n, int k, int64_t rpb,
                                                      double* __restrict__ part, int nblk,
                                                      const PcgState* __restrict__ st) {
  coldots_kernel_body<T, NQ>(A0, B0, A1, B1, A2, B2, n, k, rpb, part, nblk, st, (int)blockIdx.x, (int)blockIdx.y);
}

// out[q][c] = sum_blk part[q][blk][c], one wave per column in a fixed order:
// lane l adds blocks l, l+64, ..., then a butterfly across the wave (every
// lane ends with the same bits).
__device__ inline double wave_sum_part(const double* part, int q, int nblk, int k, int c, int lane) {
  double s = 0.0;
#pragma unroll 8  // loads in flight; the additions keep their order
  for (int b = lane; b < nblk; b += 64) s += part[((int64_t)q * nblk + b) * k + c];
  return wave_sum(s);
}

// alpha = rz / pAp with the breakdown test.  One block, one wave per column.
__device__ __forceinline__ void pcg_alpha_kernel_body(const double* __restrict__ part, int nblk, int k,
                                 const double* __restrict__ rz, double* __restrict__ alpha,
                                 PcgState* __restrict__ st, int bx_, int by_) {
  if (st->done) return;
  __shared__ int brk;
  if (threadIdx.x == 0) brk = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  for (int c = threadIdx.x >> 6; c < k; c += nw) {
    const double pap = wave_sum_part(part, 0, nblk, k, c, lane);
    const double pp = wave_sum_part(part, 1, nblk, k, c, lane);
    const double aa = wave_sum_part(part, 2, nblk, k, c, lane);
    if (lane == 0) {
      if (pap <= -st->eps * sqrt(pp) * sqrt(aa)) atomicOr(&brk, 1);
      alpha[c] = rz[c] / pap;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && brk) {
    st->done = PCG_BREAKDOWN;
    st->iterations = st->it;  // iterations completed before the breakdown
  }
}

__global__ void pcg_alpha_kernel(const double* __restrict__ part, int nblk, int k,
                                 const double* __restrict__ rz, double* __restrict__ alpha,
                                 PcgState* __restrict__ st) {
  pcg_alpha_kernel_body(part, nblk, k, rz, alpha, st, (int)blockIdx.x, (int)blockIdx.y);
}

// x += alpha p ; r -= alpha Ap
template <class T, class TX>
__global__ void pcg_update_xr_kernel(TX* __restrict__ x, T* __restrict__ r, const T* __restrict__ p,
                                     const T* __restrict__ ap, const double* __restrict__ alpha,
                                     int64_t n, int k, const PcgState* __restrict__ st) {
  if (st->done) return;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * k) return;
  const int c = (int)(idx % k);
  const double a = alpha[c];
  x[idx] = (TX)((double)x[idx] + a * (double)p[idx]);
  r[idx] = (T)((double)r[idx] - a * (double)ap[idx]);
}

// k = 1: x += alpha p ; r -= alpha Ap, and part[blk] = sum over the block's
// 256 rows of r_new^2 (the partials of ||r||^2 for pcg_resid_kernel)
template <class T>
__global__ __launch_bounds__(256) void pcg_update_xr_rr_kernel(T* __restrict__ x, T* __restrict__ r,
                                                               const T* __restrict__ p, const T* __restrict__ ap,
                                                               const double* __restrict__ alpha, int64_t n,
                                                               double* __restrict__ part,
                                                               const PcgState* __restrict__ st) {
  if (st->done) return;  // written by an earlier kernel: uniform over the grid
  __shared__ double red[16];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rr = 0.0;
  if (i < n) {
    const double a = alpha[0];
    x[i] = (T)((double)x[i] + a * (double)p[i]);
    const T rn = (T)((double)r[i] - a * (double)ap[i]);
    r[i] = rn;
    rr = (double)rn * (double)rn;
  }
  const double s = block_sum(rr, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// alpha = rz / pAp and the breakdown test of pcg_alpha_kernel (k = 1),
// evaluated by every block from the same partials in the same order, so all
// blocks agree; on breakdown block 0 records it and no block updates.
// Returns false when the block must not update.
__device__ inline bool block_alpha1(const double* __restrict__ part, int nblk, const double* __restrict__ rz,
                                    PcgState* __restrict__ st, double* alpha_out) {
  __shared__ double sums[3], sh[3];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (w < 3) {  // wave q sums quantity q: p.Ap, p.p, Ap.Ap
    const double v = wave_sum_part(part, w, nblk, 1, 0, lane);
    if (lane == 0) sums[w] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double pap = sums[0], pp = sums[1], aa = sums[2];
    sh[0] = pap;
    sh[1] = (pap <= -st->eps * sqrt(pp) * sqrt(aa)) ? 1.0 : 0.0;
    sh[2] = rz[0] / pap;
  }
  __syncthreads();
  if (sh[1] != 0.0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->done = PCG_BREAKDOWN;
      st->iterations = st->it;  // iterations completed before the breakdown
    }
    return false;
  }
  *alpha_out = sh[2];
  return true;
}

// k = 1, alpha folded in (the column-dot partials of p.Ap, p.p, Ap.Ap in
// dots[0 .. 3 nblk)): x += alpha p ; r -= alpha Ap, and rr[blk] = the block's
// sum of r_new^2.  rr must not overlap dots (other blocks still read them).
// Block 0 keeps this iteration's r.z in rz[1] (the next beta's divisor) and,
// with bump_it, counts the iteration as completed (the fused k = 1 loop: the
// residual test that follows reads the count, no kernel after it writes it).
template <class T>
__global__ __launch_bounds__(256) void pcg_update_xr_rr_alpha_kernel(T* __restrict__ x, T* __restrict__ r,
                                                                     const T* __restrict__ p,
                                                                     const T* __restrict__ ap,
                                                                     const double* __restrict__ dots, int nblk,
                                                                     double* __restrict__ rz, int64_t n,
                                                                     double* __restrict__ rr,
                                                                     PcgState* __restrict__ st, int bump_it) {
  if (st->done) return;  // written by an earlier kernel: uniform over the grid
  double a;
  if (!block_alpha1(dots, nblk, rz, st, &a)) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    rz[1] = rz[0];  // r.z of this iteration, for the next beta
    if (bump_it) st->it += 1;
  }
  __shared__ double red[16];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double r2 = 0.0;
  if (i < n) {
    x[i] = (T)((double)x[i] + a * (double)p[i]);
    const T rn = (T)((double)r[i] - a * (double)ap[i]);
    r[i] = rn;
    r2 = (double)rn * (double)rn;
  }
  const double s = block_sum(r2, red);
  if (threadIdx.x == 0) rr[blockIdx.x] = s;
}

// residual norm, convergence and max-iteration bookkeeping.  One block.
__device__ __forceinline__ void pcg_resid_kernel_body(const double* __restrict__ part, int nblk, int k,
                                 PcgState* __restrict__ st, int bx_, int by_) {
  if (st->done) return;
  const int it = st->it, max_iter = st->max_iter;
  __shared__ double red[16];
  const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  double s = 0.0;
  for (int c = threadIdx.x >> 6; c < k; c += nw) s += wave_sum_part(part, 0, nblk, k, c, lane);
  s = block_sum(lane == 0 ? s : 0.0, red);
  if (threadIdx.x == 0) {
    const double res = sqrt(s);
    st->res_norm = res;
    if (res / st->rhs_norm <= st->tol) {
      st->done = PCG_CONVERGED;
      st->iterations = it + 1;
    } else if (it + 1 >= max_iter) {
      st->done = PCG_MAXITER;
      st->iterations = max_iter;
    }
    st->it = it + 1;
  }
}

__global__ void pcg_resid_kernel(const double* __restrict__ part, int nblk, int k,
                                 PcgState* __restrict__ st) {
  pcg_resid_kernel_body(part, nblk, k, st, (int)blockIdx.x, (int)blockIdx.y);
}

// The residual test of the fused k = 1 loop, from the ||r||^2 partials rr of
// the update kernel (which has already counted the iteration, and no kernel
// of this iteration after it changes the count): every block evaluates it
// identically, block 0 records it.  Returns true (uniformly over the block)
// when the run stops here.
__device__ inline bool block_resid_test1(const double* __restrict__ rr, int nrr, PcgState* __restrict__ st) {
  __shared__ int stop;
  const int lane = threadIdx.x & 63;
  if (threadIdx.x < 64) {
    const double s = wave_sum_part(rr, 0, nrr, 1, 0, lane);
    if (lane == 0) {
      const int it = st->it, max_iter = st->max_iter;
      const double res = sqrt(s);
      const bool conv = res / st->rhs_norm <= st->tol;
      const bool last = !conv && it >= max_iter;
      stop = (conv || last) ? 1 : 0;
      if (blockIdx.x == 0) {
        st->res_norm = res;
        if (conv) {
          st->done = PCG_CONVERGED;
          st->iterations = it;
        } else if (last) {
          st->done = PCG_MAXITER;
          st->iterations = max_iter;
        }
      }
    }
  }
  __syncthreads();
  return stop != 0;
}

// k = 1: the residual test (block_resid_test1) followed, unless the run just
// stopped, by the r.z column-dot partials of coldots_kernel<T, 1> (blocks of
// rpb rows).  The preconditioner kernels that run between the update and this
// kernel do not read the state.
template <class T>
__global__ __launch_bounds__(256) void pcg_resid_rz_kernel(const double* __restrict__ rr, int nrr,
                                                           const T* __restrict__ r, const T* __restrict__ z,
                                                           int64_t n, int64_t rpb, double* __restrict__ part,
                                                           PcgState* __restrict__ st) {
  if (st->done) return;  // written by an earlier kernel: uniform over the grid
  if (block_resid_test1(rr, nrr, st)) return;
  const int64_t r0 = (int64_t)blockIdx.x * rpb;
  const int64_t r1 = (r0 + rpb < n) ? r0 + rpb : n;
  double acc = 0.0;
  for (int64_t i = r0 + threadIdx.x; i < r1; i += 256) acc += (double)r[i] * (double)z[i];
  __shared__ double red[16];
  const double t = block_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// beta = rz_new / rz_old ; rz_old = rz_new.  One block.
// bump_it: also advance the iteration counter (after pcg_resid_rz_kernel).
__device__ __forceinline__ void pcg_beta_kernel_body(const double* __restrict__ part, int nblk, int k, double* __restrict__ rz,
                                double* __restrict__ beta, PcgState* __restrict__ st, int bump_it, int bx_, int by_) {
  if (st->done) return;
  if (bump_it && threadIdx.x == 0) st->it += 1;
  const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  for (int c = threadIdx.x >> 6; c < k; c += nw) {
    const double v = wave_sum_part(part, 0, nblk, k, c, lane);
    if (lane == 0) {
      beta[c] = v / rz[c];
      rz[c] = v;
    }
  }
}

__global__ void pcg_beta_kernel(const double* __restrict__ part, int nblk, int k, double* __restrict__ rz,
                                double* __restrict__ beta, PcgState* __restrict__ st, int bump_it) {
  pcg_beta_kernel_body(part, nblk, k, rz, beta, st, bump_it, (int)blockIdx.x, (int)blockIdx.y);
}

// p = z + beta p
template <class T>
__device__ __forceinline__ void pcg_update_p_kernel_body(T* __restrict__ p, const T* __restrict__ z,
                                    const double* __restrict__ beta, int64_t n, int k,
                                    const PcgState* __restrict__ st, int bx_, int by_) {
  if (st->done) return;
  const int64_t idx = (int64_t)bx_ * blockDim.x + threadIdx.x;
  if (idx >= n * k) return;
  const int c = (int)(idx % k);
  p[idx] = (T)((double)z[idx] + beta[c] * (double)p[idx]);
}

template <class T>
__global__ void pcg_update_p_kernel(T* __restrict__ p, const T* __restrict__ z,
                                    const double* __restrict__ beta, int64_t n, int k,
                                    const PcgState* __restrict__ st) {
  pcg_update_p_kernel_body<T>(p, z, beta, n, k, st, (int)blockIdx.x, (int)blockIdx.y);
}

// k = 1: beta = r.z / rz_old and p = z + beta p in one kernel.  Every block
// sums the r.z partials in the order of pcg_beta_kernel and divides by
// rz_old = rz[1] (saved by the alpha-update kernel before rz[0] moves on), so
// beta and p are bit-identical to the two-kernel form; block 0 stores the new
// rz[0] (no other block reads it).
template <class T>
__global__ __launch_bounds__(256) void pcg_beta_update_p1_kernel(T* __restrict__ p, const T* __restrict__ z,
                                                                 const double* __restrict__ part, int nblk,
                                                                 double* __restrict__ rz, int64_t n,
                                                                 PcgState* __restrict__ st) {
  if (st->done) return;  // written by an earlier kernel: uniform over the grid
  __shared__ double bsh;
  if (threadIdx.x < 64) {
    const double v = wave_sum_part(part, 0, nblk, 1, 0, threadIdx.x);
    if (threadIdx.x == 0) {
      bsh = v / rz[1];
      if (blockIdx.x == 0) rz[0] = v;
    }
  }
  __syncthreads();
  const double beta = bsh;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (T)((double)z[i] + beta * (double)p[i]);
}

// set-up scalars: rz = colsum(r z), rhs_norm = ||b||_F.  One block.
__global__ void pcg_setup_kernel(const double* __restrict__ part, int nblk, int k, double* __restrict__ rz,
                                 PcgState* __restrict__ st, double tol, double eps, int it0, int max_iter,
                                 double xscale) {
  __shared__ double red[16];
  const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  double s = 0.0;
  for (int c = threadIdx.x >> 6; c < k; c += nw) {
    const double v = wave_sum_part(part, 0, nblk, k, c, lane);
    if (lane == 0) rz[c] = v;
    s += wave_sum_part(part, 1, nblk, k, c, lane);
  }
  s = block_sum(lane == 0 ? s : 0.0, red);
  if (threadIdx.x == 0) {
    st->done = PCG_RUNNING;
    st->iterations = it0;
    st->it = it0;
    st->max_iter = max_iter;
    st->rhs_norm = sqrt(s);
    st->res_norm = INFINITY;
    st->tol = tol;
    st->eps = eps;
    st->xscale = xscale;
  }
}

template <class T>
__global__ void axpby_kernel(const T* __restrict__ a, double fa, const T* __restrict__ b, double fb,
                             T* __restrict__ out, int64_t count, const PcgState* __restrict__ st) {
  if (st && st->done) return;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= count) return;
  double v = fa * (double)a[idx];
  if (b) v += fb * (double)b[idx];
  out[idx] = (T)v;
}

// ---------------------------------------------------------------------------
template <class T>
struct PcgBuffers {
  T *x, *r, *z, *p, *ap, *tmp;
  double *part, *rz, *alpha, *beta;
  PcgState* st;
  int nblk;
  int64_t rpb;
};

static int pcg_nblk(int64_t n, int64_t* rpb) {
  int64_t nb = std::min<int64_t>(DOT_NB, ceil_div(n, 256));
  *rpb = ceil_div(n, nb);
  return (int)ceil_div(n, *rpb);
}

template <class T>
static int coldots(const PcgBuffers<T>& b, int nq, const T* a0, const T* b0, const T* a1, const T* b1,
                   const T* a2, const T* b2, int64_t n, int k, bool check, hipStream_t st) {
  const PcgState* s = check ? b.st : nullptr;
  if (nq == 1)
    hipLaunchKernelGGL((coldots_kernel<T, 1>), dim3(b.nblk), dim3(256), 0, st, a0, b0, a1, b1, a2, b2, n,
                       k, b.rpb, b.part, b.nblk, s);
  else if (nq == 2)
    hipLaunchKernelGGL((coldots_kernel<T, 2>), dim3(b.nblk), dim3(256), 0, st, a0, b0, a1, b1, a2, b2, n,
                       k, b.rpb, b.part, b.nblk, s);
  else
    hipLaunchKernelGGL((coldots_kernel<T, 3>), dim3(b.nblk), dim3(256), 0, st, a0, b0, a1, b1, a2, b2, n,
                       k, b.rpb, b.part, b.nblk, s);
  LAKER_KCHECK("coldots_kernel");
  return LAKER_OK;
}

// A p for the operator in its dtype and the requested precision.
// dots3 (optional): see fused_product_f32; *dots3_done tells whether the
// p.Ap, p.p and Ap.Ap partials were produced on the way.
template <class T>
static int op_apply(Operator& op, int precision, const T* X, int k, T* Y, const int* skip,
                    hipStream_t st, double* dots3 = nullptr, bool* dots3_done = nullptr) {
  const bool clamp = op.clamp_for(LAKER_CLAMP_AUTO);
  if (dots3_done) *dots3_done = false;
  if constexpr (sizeof(T) == 8) {
    return fused_product_f64<double>(op, op.rows(), op.cols(), (const double*)X, k, k, (double*)Y, k,
                                     op.lambda_reg, (const double*)X, k, clamp, nullptr, st);
  } else {
    return fused_product_f32(op, precision, op.rows(), op.cols(), (const float*)X, k, k, (float*)Y, k,
                             op.lambda_reg, (const float*)X, k, clamp, false, 0, st, skip, dots3, dots3_done);
  }
}

// The product's reduction can emit the column-dot partials itself when the
// dot blocks are its own 256-row blocks (k = 1).
template <class T>
static bool dots_fusable(const PcgBuffers<T>& b, int64_t n, int k) {
  return k == 1 && b.rpb == 256 && (int64_t)b.nblk == ceil_div(n, 256);
}

template <class T>
__global__ void copy_kernel(const T* __restrict__ a, T* __restrict__ out, int64_t count) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < count) out[idx] = a[idx];
}

template <class T>
static int pcg_buffers(Operator& op, int k, PcgBuffers<T>& b) {
  const int64_t nk = op.n * k;
  int64_t rpb;
  const int nblk = pcg_nblk(op.n, &rpb);
  const size_t vec_bytes = round_up(nk * sizeof(T), 256);
  // coldots partials (3 x nblk x k) or, for k = 1, the per-256-row partials of
  // the fused update kernels
  // (3 nblk k), then, for k = 1, the per-256-row partials of ||r||^2 written by
  // the fused update kernels while other blocks still read the first part
  const size_t part_bytes = round_up(((size_t)3 * nblk * k + (size_t)ceil_div(op.n, 256)) * 8, 256);
  const size_t col_bytes = round_up((size_t)k * 8, 256);
  const size_t need = 5 * vec_bytes + part_bytes + 3 * col_bytes + 256;
  LAKER_TRY(op.pcg_ws.ensure(need));
  char* base = op.pcg_ws.as<char>();
  b.x = nullptr;
  b.r = (T*)base;
  b.z = (T*)(base + vec_bytes);
  b.p = (T*)(base + 2 * vec_bytes);
  b.ap = (T*)(base + 3 * vec_bytes);
  b.tmp = (T*)(base + 4 * vec_bytes);
  b.part = (double*)(base + 5 * vec_bytes);
  b.rz = (double*)(base + 5 * vec_bytes + part_bytes);
  b.alpha = (double*)(base + 5 * vec_bytes + part_bytes + col_bytes);
  b.beta = (double*)(base + 5 * vec_bytes + part_bytes + 2 * col_bytes);
  b.st = (PcgState*)(base + 5 * vec_bytes + part_bytes + 3 * col_bytes);
  b.nblk = nblk;
  b.rpb = rpb;
  return LAKER_OK;
}

// Runs `count` PCG iterations given by `body`: eagerly, or by replaying a HIP
// graph captured from one batch of `batch` iterations.  All per-iteration
// state lives on the device (PcgState), so a captured batch replays
// unchanged; launches after the stopping point are no-ops either way, so both
// modes give identical results.  Capture problems fall back to eager launches.
struct IterGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool failed = false;
  IterGraph() = default;
  IterGraph(const IterGraph&) = delete;
  IterGraph& operator=(const IterGraph&) = delete;
  ~IterGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
  template <class F>
  int run(F&& body, int count, int batch, bool allow, hipStream_t st) {
    if (allow && !failed && count == batch) {
      if (!exec && hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        int rc = LAKER_OK;
        for (int q = 0; q < batch && rc == LAKER_OK; ++q) rc = body();
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (rc == LAKER_OK && e == hipSuccess && g != nullptr &&
            hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) == hipSuccess) {
          graph = g;
        } else {
          if (g) (void)hipGraphDestroy(g);
          exec = nullptr;
          failed = true;
          (void)hipGetLastError();
        }
      } else if (!exec) {
        failed = true;
        (void)hipGetLastError();
      }
      if (exec) {
        LAKER_HIP_CHECK(hipGraphLaunch(exec, st));
        return LAKER_OK;
      }
    }
    for (int q = 0; q < count; ++q) LAKER_TRY(body());
    return LAKER_OK;
  }
};

template <class T>
static int pcg_run(Operator& op, Precond* pc, const T* B, int k, T* X, const T* X0, double tol,
                   int max_iter, double eps, int restart_freq, int precision, int check_every,
                   hipStream_t st, laker_pcg_result* res) {
  const int64_t n = op.n;
  const int64_t nk = n * k;
  PcgBuffers<T> b;
  LAKER_TRY(pcg_buffers<T>(op, k, b));
  const int nblk = b.nblk;
  b.x = X;
  const dim3 eg((unsigned)ceil_div(nk, 256)), eb(256);

  // set-up: x = x0 or 0 ; r = b - A x0 or b ; z = P r ; p = z
  if (X0) {
    hipLaunchKernelGGL((copy_kernel<T>), eg, eb, 0, st, X0, b.x, nk);
    LAKER_KCHECK("copy_kernel");
    LAKER_TRY(op_apply<T>(op, precision, b.x, k, b.tmp, nullptr, st));
    hipLaunchKernelGGL((axpby_kernel<T>), eg, eb, 0, st, B, 1.0, (const T*)b.tmp, -1.0, b.r, nk,
                       (const PcgState*)nullptr);
    LAKER_KCHECK("axpby_kernel");
  } else {
    LAKER_HIP_CHECK(hipMemsetAsync(b.x, 0, nk * sizeof(T), st));
    hipLaunchKernelGGL((copy_kernel<T>), eg, eb, 0, st, B, b.r, nk);
    LAKER_KCHECK("copy_kernel");
  }
  LAKER_HIP_CHECK(hipGetLastError());
  if (pc) LAKER_TRY(precond_apply(*pc, b.r, k, k, b.z, k, st));
  else hipLaunchKernelGGL((copy_kernel<T>), eg, eb, 0, st, (const T*)b.r, b.z, nk);
  hipLaunchKernelGGL((copy_kernel<T>), eg, eb, 0, st, (const T*)b.z, b.p, nk);
  LAKER_KCHECK("copy_kernel");
  LAKER_TRY(coldots<T>(b, 2, b.r, b.z, B, B, nullptr, nullptr, n, k, false, st));
  hipLaunchKernelGGL(pcg_setup_kernel, dim3(1), dim3(256), 0, st, b.part, nblk, k, b.rz, b.st, tol, eps, 0,
                     max_iter, 1.0);
  LAKER_KCHECK("pcg_setup_kernel");
  LAKER_HIP_CHECK(hipGetLastError());
  PcgState hs;
  LAKER_HIP_CHECK(hipMemcpyAsync(&hs, b.st, sizeof(hs), hipMemcpyDeviceToHost, st));
  LAKER_HIP_CHECK(hipStreamSynchronize(st));
  res->rhs_norm = hs.rhs_norm;
  res->zero_rhs = 0;
  if (hs.rhs_norm == 0.0) {
    res->iterations = 0;
    res->status = LAKER_OK;
    res->converged = 1;
    res->zero_rhs = 1;
    res->residual_norm = INFINITY;
    return LAKER_OK;
  }
  if (max_iter <= 0) {
    res->iterations = 0;
    res->status = LAKER_OK;
    res->converged = 0;
    res->residual_norm = INFINITY;
    return LAKER_OK;
  }
  if (check_every < 1) check_every = 1;
  // one iteration: Ap, p.Ap (+ breakdown test), x/r update, ||r||, z = P r, beta, p
  double* rr_part = b.part + (size_t)3 * nblk * k;  // k = 1: ||r||^2 partials
  const bool fuse_dots = dots_fusable(b, n, k);
  auto body = [&]() -> int {
    bool dots_done = false;
    LAKER_TRY(op_apply<T>(op, precision, b.p, k, b.ap, &b.st->done, st, fuse_dots ? b.part : nullptr,
                          &dots_done));
    if (!dots_done) LAKER_TRY(coldots<T>(b, 3, b.p, b.ap, b.p, b.p, b.ap, b.ap, n, k, true, st));
    if (k == 1) {
      // alpha and the ||r||^2 partials come out of the update itself
      // (restarted runs redo the partials after the residual replacement in tail)
      hipLaunchKernelGGL((pcg_update_xr_rr_alpha_kernel<T>), eg, eb, 0, st, b.x, b.r, (const T*)b.p,
                         (const T*)b.ap, (const double*)b.part, nblk, b.rz, n, rr_part, b.st,
                         restart_freq <= 0 ? 1 : 0);
      LAKER_KCHECK("pcg_update_xr_rr_alpha_kernel");
    } else {
      hipLaunchKernelGGL(pcg_alpha_kernel, dim3(1), dim3(256), 0, st, b.part, nblk, k, b.rz, b.alpha, b.st);
      LAKER_KCHECK("pcg_alpha_kernel");
      hipLaunchKernelGGL((pcg_update_xr_kernel<T, T>), eg, eb, 0, st, b.x, b.r, b.p, b.ap, b.alpha, n, k,
                         b.st);
      LAKER_KCHECK("pcg_update_xr_kernel");
    }
    return LAKER_OK;
  };
  // k = 1 without restarts: the residual test rides on the r.z dot kernel
  const bool fuse_resid = k == 1 && restart_freq <= 0;
  auto tail = [&]() -> int {
    if (!fuse_resid) {
      LAKER_TRY(coldots<T>(b, 1, b.r, b.r, nullptr, nullptr, nullptr, nullptr, n, k, true, st));
      hipLaunchKernelGGL(pcg_resid_kernel, dim3(1), dim3(256), 0, st, b.part, nblk, k, b.st);
      LAKER_KCHECK("pcg_resid_kernel");
    }
    if (pc) LAKER_TRY(precond_apply(*pc, b.r, k, k, b.z, k, st));
    else hipLaunchKernelGGL((copy_kernel<T>), eg, eb, 0, st, (const T*)b.r, b.z, nk);
    if (fuse_resid) {
      hipLaunchKernelGGL((pcg_resid_rz_kernel<T>), dim3(nblk), dim3(256), 0, st, (const double*)rr_part, (int)eg.x,
                         (const T*)b.r, (const T*)b.z, n, b.rpb, b.part, b.st);
      LAKER_KCHECK("pcg_resid_rz_kernel");
    } else {
      LAKER_TRY(coldots<T>(b, 1, b.r, b.z, nullptr, nullptr, nullptr, nullptr, n, k, true, st));
    }
    if (fuse_resid) {
      hipLaunchKernelGGL((pcg_beta_update_p1_kernel<T>), eg, eb, 0, st, b.p, (const T*)b.z, (const double*)b.part,
                         nblk, b.rz, n, b.st);
      LAKER_KCHECK("pcg_beta_update_p1_kernel");
      return LAKER_OK;
    }
    hipLaunchKernelGGL(pcg_beta_kernel, dim3(1), dim3(256), 0, st, b.part, nblk, k, b.rz, b.beta, b.st, 0);
    LAKER_KCHECK("pcg_beta_kernel");
    hipLaunchKernelGGL((pcg_update_p_kernel<T>), eg, eb, 0, st, b.p, b.z, b.beta, n, k, b.st);
    LAKER_KCHECK("pcg_update_p_kernel");
    return LAKER_OK;
  };
  auto full = [&]() -> int {
    LAKER_TRY(body());
    return tail();
  };
  IterGraph ig;
  const bool graphs = g_use_graphs.load() != 0 && restart_freq <= 0;
  int launched = 0;
  while (launched < max_iter) {
    const int nb = std::min(check_every, max_iter - launched);
    if (restart_freq > 0) {
      for (int q = 0; q < nb; ++q) {
        const int it = launched + q;
        LAKER_TRY(body());
        if ((it + 1) % restart_freq == 0) {
          // residual replacement r = b - A x
          LAKER_TRY(op_apply<T>(op, precision, b.x, k, b.tmp, &b.st->done, st));
          hipLaunchKernelGGL((axpby_kernel<T>), eg, eb, 0, st, B, 1.0, (const T*)b.tmp, -1.0, b.r, nk,
                             (const PcgState*)b.st);
          LAKER_KCHECK("axpby_kernel");
        }
        LAKER_TRY(tail());
      }
    } else {
      // the first batch runs eagerly (it sizes every work buffer), later ones replay
      LAKER_TRY(ig.run(full, nb, check_every, graphs && launched > 0, st));
    }
    launched += nb;
    LAKER_HIP_CHECK(hipMemcpyAsync(&hs, b.st, sizeof(hs), hipMemcpyDeviceToHost, st));
    LAKER_HIP_CHECK(hipStreamSynchronize(st));
    if (hs.done) break;
  }
  LAKER_HIP_CHECK(hipMemcpyAsync(&hs, b.st, sizeof(hs), hipMemcpyDeviceToHost, st));
  LAKER_HIP_CHECK(hipStreamSynchronize(st));
  res->iterations = hs.done ? hs.iterations : max_iter;
  res->residual_norm = hs.res_norm;
  res->converged = hs.done == PCG_CONVERGED;
  res->status = hs.done == PCG_BREAKDOWN ? LAKER_EBREAKDOWN : LAKER_OK;
  if (hs.done == PCG_BREAKDOWN) {
    res->iterations = hs.iterations;
    return fail(LAKER_EBREAKDOWN,
                "PCG breakdown: non-positive curvature detected (p^T A p <= 0). The operator may be "
                "indefinite or the preconditioner may be unsuitable.");
  }
  return LAKER_OK;
}

// ---------------------------------------------------------------------------
// FP64 helpers of the refinement loop
// ---------------------------------------------------------------------------
__global__ void frob2_part_kernel(const double* __restrict__ a, int64_t count, double* __restrict__ part) {
  __shared__ double red[16];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    s += a[i] * a[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

static int frob_norm(const double* a, int64_t count, DevBuf& tmp, double* out, hipStream_t st) {
  const int nb = (int)std::min<int64_t>(256, ceil_div(count, 256));
  LAKER_TRY(tmp.ensure((size_t)nb * 8, st));
  hipLaunchKernelGGL(frob2_part_kernel, dim3(nb), dim3(256), 0, st, a, count, tmp.as<double>());
  LAKER_KCHECK("frob2_part_kernel");
  LAKER_HIP_CHECK(hipGetLastError());
  std::vector<double> h(nb);
  LAKER_HIP_CHECK(hipMemcpyAsync(h.data(), tmp.ptr, nb * 8, hipMemcpyDeviceToHost, st));
  LAKER_HIP_CHECK(hipStreamSynchronize(st));
  double s = 0.0;
  for (double v : h) s += v;
  *out = std::sqrt(s);
  return LAKER_OK;
}

__global__ void scale_to_f32_kernel(const double* __restrict__ a, float* __restrict__ out, int64_t count,
                                    double f) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = (float)(a[i] * f);
}
// k = 1: x (FP64) += f alpha p ; r -= alpha Ap, and part[blk] = sum over the
// block's 256 rows of r_new^2
__global__ __launch_bounds__(256) void pcg_update_x64_rr_kernel(double* __restrict__ x, float* __restrict__ r,
                                                                const float* __restrict__ p,
                                                                const float* __restrict__ ap,
                                                                const double* __restrict__ alpha, int64_t n,
                                                                double* __restrict__ part,
                                                                const PcgState* __restrict__ st) {
  if (st->done) return;  // written by an earlier kernel: uniform over the grid
  __shared__ double red[16];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rr = 0.0;
  if (i < n) {
    const double a = alpha[0];
    x[i] += st->xscale * a * (double)p[i];
    const float rn = (float)((double)r[i] - a * (double)ap[i]);
    r[i] = rn;
    rr = (double)rn * (double)rn;
  }
  const double s = block_sum(rr, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// as pcg_update_x64_rr_kernel with alpha folded in (see
// pcg_update_xr_rr_alpha_kernel); rr must not overlap dots
__global__ __launch_bounds__(256) void pcg_update_x64_rr_alpha_kernel(double* __restrict__ x, float* __restrict__ r,
                                                                      const float* __restrict__ p,
                                                                      const float* __restrict__ ap,
                                                                      const double* __restrict__ dots, int nblk,
                                                                      double* __restrict__ rz, int64_t n,
                                                                      double* __restrict__ rr,
                                                                      PcgState* __restrict__ st) {
  if (st->done) return;  // written by an earlier kernel: uniform over the grid
  double a;
  if (!block_alpha1(dots, nblk, rz, st, &a)) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    rz[1] = rz[0];  // r.z of this iteration, for the next beta
    st->it += 1;    // this iteration is complete
  }
  __shared__ double red[16];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double r2 = 0.0;
  if (i < n) {
    x[i] += st->xscale * a * (double)p[i];
    const float rn = (float)((double)r[i] - a * (double)ap[i]);
    r[i] = rn;
    r2 = (double)rn * (double)rn;
  }
  const double s = block_sum(r2, red);
  if (threadIdx.x == 0) rr[blockIdx.x] = s;
}

// x (FP64) += f * alpha p ; r -= alpha Ap  (FP32 recursion on a scaled residual)
__device__ __forceinline__ void pcg_update_x64_kernel_body(double* __restrict__ x, float* __restrict__ r,
                                      const float* __restrict__ p, const float* __restrict__ ap,
                                      const double* __restrict__ alpha, int64_t n, int k,
                                      const PcgState* __restrict__ st, int bx_, int by_) {
  if (st->done) return;
  const int64_t idx = (int64_t)bx_ * blockDim.x + threadIdx.x;
  if (idx >= n * k) return;
  const double f = st->xscale;
  const double a = alpha[idx % k];
  x[idx] += f * a * (double)p[idx];
  r[idx] = (float)((double)r[idx] - a * (double)ap[idx]);
}

__global__ void pcg_update_x64_kernel(double* __restrict__ x, float* __restrict__ r,
                                      const float* __restrict__ p, const float* __restrict__ ap,
                                      const double* __restrict__ alpha, int64_t n, int k,
                                      const PcgState* __restrict__ st) {
  pcg_update_x64_kernel_body(x, r, p, ap, alpha, n, k, st, (int)blockIdx.x, (int)blockIdx.y);
}




// ---------------------------------------------------------------------------
// Batched mixed-precision solves (laker_batch_*): the inner PCG iterations of
// many small systems (uneven n, 2 <= k <= 32 right-hand sides each) run as
// shared launches -- one launch per kernel of the iteration for all systems
// whose refinement step is running -- and each system's refinement steps
// (FP64 residual, best iterate, next step) are taken on the host between
// polls, exactly as laker_solve_mixed takes them.  Every system keeps its
// own operator, preconditioner and Krylov buffers, its kernels run with the
// block decomposition of an unbatched launch, so a system's iterates do not
// depend on which other systems share the launches.
// ---------------------------------------------------------------------------
template <int NQ>
__global__ __launch_bounds__(256) void coldots_batched(const BColdots* __restrict__ desc, const int* __restrict__ off,
                                                       int nsys) {
  const int s = batch_system(off, nsys, (int)blockIdx.x);
  const BColdots d = desc[s];
  coldots_kernel_body<float, NQ>(d.A0, d.B0, d.A1, d.B1, d.A2, d.B2, d.n, d.k, d.rpb, d.part, d.nblk, d.st,
                                 (int)blockIdx.x - off[s], 0);
}
__global__ void alpha_batched(const BScalar* __restrict__ desc, int nsys) {
  const BScalar d = desc[blockIdx.x];
  pcg_alpha_kernel_body(d.part, d.nblk, d.k, d.rz, d.out, d.st, 0, 0);
}
__global__ void resid_batched(const BScalar* __restrict__ desc, int nsys) {
  const BScalar d = desc[blockIdx.x];
  pcg_resid_kernel_body(d.part, d.nblk, d.k, d.st, 0, 0);
}
__global__ void beta_batched(const BScalar* __restrict__ desc, int nsys) {
  const BScalar d = desc[blockIdx.x];
  pcg_beta_kernel_body(d.part, d.nblk, d.k, d.rz, d.out, d.st, 0, 0, 0);
}
__global__ void update_x64_batched(const BUpdX64* __restrict__ desc, const int* __restrict__ off, int nsys) {
  const int s = batch_system(off, nsys, (int)blockIdx.x);
  const BUpdX64 d = desc[s];
  pcg_update_x64_kernel_body(d.x, d.r, d.p, d.ap, d.alpha, d.n, d.k, d.st, (int)blockIdx.x - off[s], 0);
}
__global__ void update_p_batched(const BUpdP* __restrict__ desc, const int* __restrict__ off, int nsys) {
  const int s = batch_system(off, nsys, (int)blockIdx.x);
  const BUpdP d = desc[s];
  pcg_update_p_kernel_body<float>(d.p, d.z, d.beta, d.n, d.k, d.st, (int)blockIdx.x - off[s], 0);
}

struct BatchSys {
  Operator* op = nullptr;
  Precond* pc = nullptr;
  const double* B = nullptr;
  double* X = nullptr;
  int k = 0;
  int64_t n = 0;
  double tol = 1e-6, inner_tol = 5e-2;
  int max_outer = 20, inner_max_iter = 1000;
  PcgBuffers<float> b;
  double *R = nullptr, *Xbest = nullptr, *Rbest = nullptr;
  DevBuf tmp;
  const double* Rin = nullptr;
  double bnorm = 0.0, rnorm = 0.0, rbest = 0.0;
  bool have_best = false, final_step = false, running = false, finished = false;
  int it = 0, seg_start = 0;
  laker_mixed_result res{};
  // product geometry (product_multi_f32 for k <= 32 on the GX_SPLIT3 path)
  int gm = 0, ns = 0, jsplit = 1;
  bool clamp = false;
  const void *colpack = nullptr, *rowpack = nullptr;
  int ka = 0;
  bf16x8 *xh = nullptr, *xl = nullptr, *xm = nullptr;
  int64_t ntiles = 0;
  BatchApplyView av{};
};

struct Batch {
  hipStream_t st = nullptr;
  int max_sys = 0;
  std::vector<BatchSys*> slot;
  DevBuf states;    // PcgState per slot
  DevBuf table_d;   // descriptor tables and block offsets of one poll
  void* table_h = nullptr;  // pinned staging copy
  size_t table_bytes = 0;
  std::vector<PcgState> hs;
  ~Batch() {
    for (BatchSys* s : slot) delete s;
    if (table_h) (void)hipHostFree(table_h);
  }
};

// one FP64 refinement residual of system s and the bookkeeping of
// laker_solve_mixed after an inner step; starts the next step or finishes
static int batch_step_begin(Batch& bt, BatchSys& s);

static int batch_finish_step(Batch& bt, BatchSys& s, const PcgState& h) {
  const hipStream_t st = bt.st;
  const int64_t nk = s.n * s.k;
  s.running = false;
  s.it = h.done != PCG_RUNNING ? h.iterations : h.it;
  s.res.inner_iterations = s.it;
  if (h.done == PCG_BREAKDOWN && s.it == s.seg_start) {
    s.res.status = LAKER_EBREAKDOWN;
    s.finished = true;
    return LAKER_OK;
  }
  const bool clamp = s.op->clamp_for(LAKER_CLAMP_AUTO);
  LAKER_TRY(fused_product_f64<double>(*s.op, s.op->rows(), s.op->cols(), s.X, s.k, s.k, s.R, s.k, s.op->lambda_reg,
                                      s.X, s.k, clamp, s.B, st));
  LAKER_TRY(frob_norm(s.R, nk, s.tmp, &s.rnorm, st));
  s.res.outer_iterations += 1;
  bool stop = false;
  if (s.rnorm / s.bnorm <= s.tol) {
    s.res.converged = 1;
    s.res.residual_norm = s.rnorm;
    stop = true;
  } else if (std::isfinite(s.rnorm) && s.rnorm < s.rbest) {
    s.rbest = s.rnorm;
    s.have_best = true;
    LAKER_HIP_CHECK(hipMemcpyAsync(s.Xbest, s.X, nk * 8, hipMemcpyDeviceToDevice, st));
    LAKER_HIP_CHECK(hipMemcpyAsync(s.Rbest, s.R, nk * 8, hipMemcpyDeviceToDevice, st));
    if (s.final_step) stop = true;
  } else {
    if (s.have_best) {
      LAKER_HIP_CHECK(hipMemcpyAsync(s.X, s.Xbest, nk * 8, hipMemcpyDeviceToDevice, st));
      LAKER_HIP_CHECK(hipMemcpyAsync(s.R, s.Rbest, nk * 8, hipMemcpyDeviceToDevice, st));
    } else {
      LAKER_HIP_CHECK(hipMemsetAsync(s.X, 0, nk * 8, st));
      LAKER_HIP_CHECK(hipMemcpyAsync(s.R, s.B, nk * 8, hipMemcpyDeviceToDevice, st));
    }
    s.rnorm = s.rbest;
    if (s.final_step) stop = true;
    s.final_step = true;
  }
  if (!stop) {
    s.res.residual_norm = s.rnorm;
    if (s.res.outer_iterations >= s.max_outer || s.it >= s.inner_max_iter || s.rnorm == 0.0) stop = true;
  }
  if (stop) {
    s.res.residual_norm = std::min(s.rnorm, s.rbest);
    s.res.inner_iterations = s.it;
    s.finished = true;
    return LAKER_OK;
  }
  s.Rin = s.R;
  return batch_step_begin(bt, s);
}

static int batch_step_begin(Batch& bt, BatchSys& s) {
  const hipStream_t st = bt.st;
  const int64_t nk = s.n * s.k;
  const dim3 eg((unsigned)ceil_div(nk, 256)), eb(256);
  const double need = 0.6 * s.tol * s.bnorm / s.rnorm;
  const double seg_tol = std::min(0.5, s.final_step ? need : std::max(s.inner_tol, need));
  const double eps = std::sqrt(1.1920928955078125e-07);
  hipLaunchKernelGGL(scale_to_f32_kernel, eg, eb, 0, st, s.Rin, s.b.r, nk, 1.0 / s.rnorm);
  LAKER_KCHECK("scale_to_f32_kernel");
  LAKER_TRY(precond_apply(*s.pc, s.b.r, s.k, s.k, s.b.z, s.k, st));
  hipLaunchKernelGGL((copy_kernel<float>), eg, eb, 0, st, (const float*)s.b.z, s.b.p, nk);
  LAKER_KCHECK("copy_kernel");
  LAKER_TRY(coldots<float>(s.b, 2, s.b.r, s.b.z, s.b.r, s.b.r, nullptr, nullptr, s.n, s.k, false, st));
  hipLaunchKernelGGL(pcg_setup_kernel, dim3(1), dim3(256), 0, st, s.b.part, s.b.nblk, s.k, s.b.rz, s.b.st, seg_tol,
                     eps, s.it, s.inner_max_iter, s.rnorm);
  LAKER_KCHECK("pcg_setup_kernel");
  s.seg_start = s.it;
  s.running = true;
  return LAKER_OK;
}

// Descriptor tables of one poll, laid out in one staging buffer.
struct TableBuilder {
  std::vector<char> h;
  size_t add_bytes(const void* p, size_t bytes) {
    const size_t o = round_up(h.size(), 64);
    h.resize(o + bytes);
    memcpy(h.data() + o, p, bytes);
    return o;
  }
  template <class D>
  struct Part {
    size_t desc_off = 0, off_off = 0;
    int nsys = 0, blocks = 0;
  };
  template <class D>
  Part<D> add(const std::vector<D>& d, const std::vector<int>& blocks) {
    Part<D> p;
    p.nsys = (int)d.size();
    std::vector<int> off(d.size() + 1, 0);
    for (size_t i = 0; i < d.size(); ++i) off[i + 1] = off[i] + blocks[i];
    p.blocks = off.back();
    if (!d.empty()) {
      p.desc_off = add_bytes(d.data(), d.size() * sizeof(D));
      p.off_off = add_bytes(off.data(), off.size() * sizeof(int));
    }
    return p;
  }
  template <class D>
  BatchLaunch<D> launch(const Part<D>& p, const char* base) const {
    BatchLaunch<D> L;
    L.desc = reinterpret_cast<const D*>(base + p.desc_off);
    L.off = reinterpret_cast<const int*>(base + p.off_off);
    L.nsys = p.nsys;
    L.blocks = p.blocks;
    return L;
  }
};

// Inner iterations of every running system, `count` times, as shared launches.
static int batch_iterate(Batch& bt, const std::vector<BatchSys*>& act, int count) {
  const hipStream_t st = bt.st;
  TableBuilder tb;
  // product: groups by Gram instance (gm, ns, clamp)
  struct MGroup {
    int gm, ns;
    bool clamp;
    std::vector<BMulti2> d;
    std::vector<int> blocks;
    TableBuilder::Part<BMulti2> part;
  };
  std::vector<MGroup> mg;
  std::vector<BPackX> px;
  std::vector<int> pxb;
  std::vector<BReduce> rd;
  std::vector<int> rdb;
  std::vector<BColdots> cd3, cdr, cdz;
  std::vector<int> cdb;
  std::vector<BScalar> sa, sr, sb;
  std::vector<BUpdX64> ux;
  std::vector<int> uxb;
  std::vector<BUpdP> up;
  std::vector<BQtx> qt;
  std::vector<int> qtb;
  std::vector<BRedParts> r1, r2;
  std::vector<int> r1b, r2b;
  std::vector<BGemm> gm;
  std::vector<int> gmb;
  struct QGroup {
    int nb, P;
    std::vector<BQw> d;
    std::vector<int> blocks;
    TableBuilder::Part<BQw> part;
  };
  std::vector<QGroup> qg;
  for (BatchSys* sp : act) {
    BatchSys& s = *sp;
    const int k = s.k;
    const int64_t n = s.n, nk = n * k;
    const int kp = 32;
    px.push_back(BPackX{s.b.p, (int64_t)k, n, s.ntiles, k, 0, 1, s.xh, s.xl, s.xm});
    pxb.push_back((int)ceil_div(s.ntiles * 128, 256));
    MGroup* g = nullptr;
    for (auto& q : mg)
      if (q.gm == s.gm && q.ns == s.ns && q.clamp == s.clamp) g = &q;
    if (!g) {
      mg.push_back(MGroup{s.gm, s.ns, s.clamp, {}, {}, {}});
      g = &mg.back();
    }
    g->d.push_back(BMulti2{s.colpack, s.rowpack, s.ka, n, n, s.xh, s.xl, s.xm, k, 0, s.op->part.as<float>(),
                           (int64_t)kp, s.jsplit, (float)(s.op->clamp_limit * kLog2eD)});
    g->blocks.push_back((int)(ceil_div(n, 128) * s.jsplit));
    rd.push_back(BReduce{s.op->part.as<float>(), n, (int64_t)kp, (int64_t)k, (int64_t)k, k, s.jsplit,
                         s.op->lambda_reg, s.b.p, s.b.ap});
    rdb.push_back((int)ceil_div(nk, 256));
    cd3.push_back(BColdots{s.b.p, s.b.ap, s.b.p, s.b.p, s.b.ap, s.b.ap, n, s.b.rpb, k, s.b.nblk, s.b.part, s.b.st});
    cdr.push_back(BColdots{s.b.r, s.b.r, nullptr, nullptr, nullptr, nullptr, n, s.b.rpb, k, s.b.nblk, s.b.part,
                           s.b.st});
    cdz.push_back(BColdots{s.b.r, s.b.z, nullptr, nullptr, nullptr, nullptr, n, s.b.rpb, k, s.b.nblk, s.b.part,
                           s.b.st});
    cdb.push_back(s.b.nblk);
    sa.push_back(BScalar{s.b.part, s.b.nblk, k, s.b.rz, s.b.alpha, s.b.st});
    sr.push_back(BScalar{s.b.part, s.b.nblk, k, s.b.rz, nullptr, s.b.st});
    sb.push_back(BScalar{s.b.part, s.b.nblk, k, s.b.rz, s.b.beta, s.b.st});
    ux.push_back(BUpdX64{s.X, s.b.r, s.b.p, s.b.ap, s.b.alpha, n, k, s.b.st});
    uxb.push_back((int)ceil_div(nk, 256));
    up.push_back(BUpdP{s.b.p, s.b.z, s.b.beta, n, k, s.b.st});
    const BatchApplyView& v = s.av;
    qt.push_back(BQtx{v.Q, n, (int64_t)k, v.rps, v.nr, k, v.cblk, s.b.r, v.part});
    qtb.push_back(v.cblk * v.ns_eff);
    const int wb = (int)ceil_div(v.width, 64);
    r1.push_back(BRedParts{v.part, v.ns_eff, v.width, wb, 8, v.part2});
    r1b.push_back(wb * 8);
    r2.push_back(BRedParts{v.part2, 8, v.width, wb, 1, v.W1});
    r2b.push_back(wb);
    const int ggx = (int)ceil_div(k, 64), ggy = (int)ceil_div(v.nr, 64);
    gm.push_back(BGemm{v.nr, k, v.nr, v.nr, k, k, 1.0, 0.0, v.S, v.W1, v.W2, ggx});
    gmb.push_back(ggx * ggy);
    QGroup* h = nullptr;
    for (auto& q : qg)
      if (q.nb == v.qw_nb && q.P == v.P) h = &q;
    if (!h) {
      qg.push_back(QGroup{v.qw_nb, v.P, {}, {}, {}});
      h = &qg.back();
    }
    h->d.push_back(BQw{v.Q, v.W2, n, (int64_t)k, (int64_t)k, v.nr, v.P, k, 0, s.b.r, s.b.z, v.f});
    h->blocks.push_back(v.qw_blocks);
  }
  const std::vector<int> ones(act.size(), 1);
  auto tpx = tb.add(px, pxb);
  for (auto& g : mg) g.part = tb.add(g.d, g.blocks);
  auto trd = tb.add(rd, rdb);
  auto tc3 = tb.add(cd3, cdb);
  auto tcr = tb.add(cdr, cdb);
  auto tcz = tb.add(cdz, cdb);
  auto tsa = tb.add(sa, ones);
  auto tsr = tb.add(sr, ones);
  auto tsb = tb.add(sb, ones);
  auto tux = tb.add(ux, uxb);
  auto tup = tb.add(up, uxb);
  auto tqt = tb.add(qt, qtb);
  auto tr1 = tb.add(r1, r1b);
  auto tr2 = tb.add(r2, r2b);
  auto tgm = tb.add(gm, gmb);
  for (auto& q : qg) q.part = tb.add(q.d, q.blocks);
  const size_t bytes = round_up(tb.h.size(), 256);
  if (bytes > bt.table_bytes) {
    // the staging buffers only grow; the previous poll's copy has completed
    LAKER_HIP_CHECK(hipStreamSynchronize(st));
    if (bt.table_h) (void)hipHostFree(bt.table_h);
    bt.table_h = nullptr;
    LAKER_HIP_CHECK(hipHostMalloc(&bt.table_h, bytes * 2, hipHostMallocDefault));
    LAKER_TRY(bt.table_d.ensure(bytes * 2));
    bt.table_bytes = bytes * 2;
  }
  LAKER_HIP_CHECK(hipStreamSynchronize(st));  // the staging copy may be in use by the previous poll
  memcpy(bt.table_h, tb.h.data(), tb.h.size());
  LAKER_HIP_CHECK(hipMemcpyAsync(bt.table_d.ptr, bt.table_h, tb.h.size(), hipMemcpyHostToDevice, st));
  const char* base = bt.table_d.as<char>();
  const int nsys = (int)act.size();
  const dim3 b256(256);
  for (int q = 0; q < count; ++q) {
    // A p
    LAKER_TRY(batched_pack_x(tb.launch(tpx, base), st));
    for (auto& g : mg) LAKER_TRY(batched_multi2(tb.launch(g.part, base), g.gm, g.ns, g.clamp, st));
    LAKER_TRY(batched_reduce(tb.launch(trd, base), st));
    // p.Ap, p.p, Ap.Ap ; alpha ; x, r update ; ||r||
    {
      const auto L = tb.launch(tc3, base);
      hipLaunchKernelGGL((coldots_batched<3>), dim3((unsigned)L.blocks), b256, 0, st, L.desc, L.off, nsys);
      LAKER_KCHECK("coldots_batched");
    }
    hipLaunchKernelGGL(alpha_batched, dim3((unsigned)nsys), b256, 0, st, tb.launch(tsa, base).desc, nsys);
    LAKER_KCHECK("alpha_batched");
    {
      const auto L = tb.launch(tux, base);
      hipLaunchKernelGGL(update_x64_batched, dim3((unsigned)L.blocks), b256, 0, st, L.desc, L.off, nsys);
      LAKER_KCHECK("update_x64_batched");
    }
    {
      const auto L = tb.launch(tcr, base);
      hipLaunchKernelGGL((coldots_batched<1>), dim3((unsigned)L.blocks), b256, 0, st, L.desc, L.off, nsys);
      LAKER_KCHECK("coldots_batched");
    }
    hipLaunchKernelGGL(resid_batched, dim3((unsigned)nsys), b256, 0, st, tb.launch(tsr, base).desc, nsys);
    LAKER_KCHECK("resid_batched");
    // z = P r
    LAKER_TRY(batched_qtx(tb.launch(tqt, base), st));
    LAKER_TRY(batched_reduce_parts(tb.launch(tr1, base), st));
    LAKER_TRY(batched_reduce_parts(tb.launch(tr2, base), st));
    LAKER_TRY(batched_gemm(tb.launch(tgm, base), st));
    for (auto& q : qg) LAKER_TRY(batched_qw(tb.launch(q.part, base), q.nb, q.P, st));
    // r.z ; beta ; p
    {
      const auto L = tb.launch(tcz, base);
      hipLaunchKernelGGL((coldots_batched<1>), dim3((unsigned)L.blocks), b256, 0, st, L.desc, L.off, nsys);
      LAKER_KCHECK("coldots_batched");
    }
    hipLaunchKernelGGL(beta_batched, dim3((unsigned)nsys), b256, 0, st, tb.launch(tsb, base).desc, nsys);
    LAKER_KCHECK("beta_batched");
    {
      const auto L = tb.launch(tup, base);
      hipLaunchKernelGGL(update_p_batched, dim3((unsigned)L.blocks), b256, 0, st, L.desc, L.off, nsys);
      LAKER_KCHECK("update_p_batched");
    }
  }
  return LAKER_OK;
}

}  // namespace laker

using namespace laker;

extern "C" {

int laker_set_graphs(int32_t enable) {
  g_use_graphs.store(enable ? 1 : 0);
  return LAKER_OK;
}

int laker_pcg_solve(laker_operator_t oh, laker_precond_t ph, const void* B, int32_t k, void* X,
                    const void* X0, double tol, int32_t max_iter, double breakdown_eps,
                    int32_t restart_freq, int32_t precision, int32_t check_every, hipStream_t stream,
                    laker_pcg_result* result) {
  LAKER_CHECK_ARG(oh != nullptr, "operator handle is NULL");
  LAKER_CHECK_ARG(B && X && result && k >= 1, "bad pcg arguments");
  Operator& op = *reinterpret_cast<Operator*>(oh);
  Precond* pc = reinterpret_cast<Precond*>(ph);
  LAKER_CHECK_ARG(!pc || precond_rows(*pc) == op.n, "preconditioner and operator sizes differ");
  LAKER_CHECK_ARG(!pc || precond_dtype(*pc) == op.dtype, "preconditioner and operator dtypes differ");
  if (op.dtype == LAKER_F64) {
    const double eps = breakdown_eps >= 0 ? breakdown_eps : std::sqrt(2.220446049250313e-16);
    return pcg_run<double>(op, pc, (const double*)B, k, (double*)X, (const double*)X0, tol, max_iter, eps,
                           restart_freq, LAKER_PREC_FP64, check_every, stream, result);
  }
  const double eps = breakdown_eps >= 0 ? breakdown_eps : std::sqrt(1.1920928955078125e-07);
  LAKER_CHECK_ARG(precision == LAKER_PREC_FP32 || precision == LAKER_PREC_MIXED, "bad PCG precision");
  return pcg_run<float>(op, pc, (const float*)B, k, (float*)X, (const float*)X0, tol, max_iter, eps,
                        restart_freq, precision, check_every, stream, result);
}

int laker_solve_mixed(laker_operator_t oh, laker_precond_t ph, const double* B, int32_t k, double* X,
                      double tol, int32_t max_outer, double inner_tol, int32_t inner_max_iter,
                      int32_t inner_precision, hipStream_t st, laker_mixed_result* result) {
  LAKER_CHECK_ARG(oh != nullptr, "operator handle is NULL");
  LAKER_CHECK_ARG(B && X && result && k >= 1, "bad mixed-solve arguments");
  LAKER_CHECK_ARG(inner_tol > 0.0 && inner_tol < 1.0, "inner_tol must lie in (0, 1)");
  Operator& op = *reinterpret_cast<Operator*>(oh);
  LAKER_CHECK_ARG(op.dtype == LAKER_F32, "mixed solve needs an FP32 operator");
  LAKER_CHECK_ARG(inner_precision == LAKER_PREC_MIXED || inner_precision == LAKER_PREC_FP32,
                  "inner precision must be LAKER_PREC_MIXED or LAKER_PREC_FP32");
  Precond* pc = reinterpret_cast<Precond*>(ph);
  LAKER_CHECK_ARG(!pc || precond_rows(*pc) == op.n, "preconditioner and operator sizes differ");
  LAKER_CHECK_ARG(!pc || precond_dtype(*pc) == LAKER_F32, "mixed solve needs an FP32 preconditioner");
  const int64_t n = op.n, nk = n * k;
  const size_t vb = round_up(nk * 8, 256);
  LAKER_TRY(op.mixed_ws.ensure(3 * vb));
  double* R = op.mixed_ws.as<double>();
  double* Xbest = reinterpret_cast<double*>(op.mixed_ws.as<char>() + vb);
  double* Rbest = reinterpret_cast<double*>(op.mixed_ws.as<char>() + 2 * vb);
  PcgBuffers<float> b;
  LAKER_TRY(pcg_buffers<float>(op, k, b));
  const int nblk = b.nblk;
  DevBuf tmp;
  // k = 1 with a CCCP handle: the applies read the FP16 basis (half the bytes;
  // an SPD preconditioner within O(2^-11) of the FP32 one, used throughout
  // this solve)
  bool half_apply = false;
  if (pc && k == 1) {
    const int hr = precond_prepare_half(*pc, st);
    if (hr < 0) return hr;
    half_apply = hr == LAKER_OK;
  }
  auto papply = [&](const float* r, float* z) -> int {
    if (half_apply) return precond_apply_half(*pc, r, z, st);
    return precond_apply(*pc, r, k, k, z, k, st);
  };
  const dim3 eg((unsigned)ceil_div(nk, 256)), eb(256);
  const bool clamp = op.clamp_for(LAKER_CLAMP_AUTO);
  const double eps = std::sqrt(1.1920928955078125e-07);
  result->outer_iterations = 0;
  result->inner_iterations = 0;
  result->status = LAKER_OK;
  result->converged = 0;
  double bnorm = 0.0;
  LAKER_TRY(frob_norm(B, nk, tmp, &bnorm, st));
  result->rhs_norm = bnorm;
  result->residual_norm = bnorm;
  LAKER_HIP_CHECK(hipMemsetAsync(X, 0, nk * 8, st));
  if (bnorm == 0.0) {
    result->converged = 1;
    result->residual_norm = 0.0;
    return LAKER_OK;
  }
  // Iterative refinement.  Each outer step runs PCG from a fresh start on the
  // FP64 residual scaled to unit norm (FP32 vectors, any ||r|| in range) and
  // adds its correction straight into the FP64 X (X += ||r|| alpha p).  The
  // FP32 operator's own error limits what one step can gain on the true
  // residual, so a step stops once its recursive residual has dropped by
  // inner_tol (or by what the outer target still needs); the next FP64
  // residual then starts the next step.
  // Safeguard: when lambda is below the reduced-precision operator's error
  // (u ||K|| > lambda), refinement stops contracting.  The best FP64 iterate
  // is kept; the first step that fails to improve it is undone and replaced by
  // one last step run to the full target from the best point.
  const double* Rin = B;
  double rnorm = bnorm, rbest = bnorm;
  bool have_best = false, final_step = false;
  int it = 0;
  const int check_every = 8;
  PcgState hs;
  // one inner iteration; x accumulates in FP64 with the step's scale (device state)
  double* rr_part = b.part + (size_t)3 * nblk * k;  // k = 1: ||r||^2 partials
  const bool fuse_dots = dots_fusable(b, n, k);
  auto body = [&]() -> int {
    bool dots_done = false;
    LAKER_TRY(op_apply<float>(op, inner_precision, b.p, k, b.ap, &b.st->done, st, fuse_dots ? b.part : nullptr,
                              &dots_done));
    if (!dots_done) LAKER_TRY(coldots<float>(b, 3, b.p, b.ap, b.p, b.p, b.ap, b.ap, n, k, true, st));
    if (k == 1) {
      hipLaunchKernelGGL(pcg_update_x64_rr_alpha_kernel, eg, eb, 0, st, X, b.r, (const float*)b.p,
                         (const float*)b.ap, (const double*)b.part, nblk, b.rz, n, rr_part, b.st);
      LAKER_KCHECK("pcg_update_x64_rr_alpha_kernel");
    } else {
      hipLaunchKernelGGL(pcg_alpha_kernel, dim3(1), dim3(256), 0, st, b.part, nblk, k, b.rz, b.alpha, b.st);
      LAKER_KCHECK("pcg_alpha_kernel");
      hipLaunchKernelGGL(pcg_update_x64_kernel, eg, eb, 0, st, X, b.r, (const float*)b.p, (const float*)b.ap,
                         (const double*)b.alpha, n, k, (const PcgState*)b.st);
      LAKER_KCHECK("pcg_update_x64_kernel");
      LAKER_TRY(coldots<float>(b, 1, b.r, b.r, nullptr, nullptr, nullptr, nullptr, n, k, true, st));
      hipLaunchKernelGGL(pcg_resid_kernel, dim3(1), dim3(256), 0, st, b.part, nblk, k, b.st);
      LAKER_KCHECK("pcg_resid_kernel");
    }
    if (pc) LAKER_TRY(papply(b.r, b.z));
    else hipLaunchKernelGGL((copy_kernel<float>), eg, eb, 0, st, (const float*)b.r, b.z, nk);
    if (k == 1) {
      // the residual test of this iteration rides on the r.z dot kernel
      hipLaunchKernelGGL((pcg_resid_rz_kernel<float>), dim3(nblk), dim3(256), 0, st, (const double*)rr_part,
                         (int)eg.x, (const float*)b.r, (const float*)b.z, n, b.rpb, b.part, b.st);
      LAKER_KCHECK("pcg_resid_rz_kernel");
      hipLaunchKernelGGL((pcg_beta_update_p1_kernel<float>), eg, eb, 0, st, b.p, (const float*)b.z,
                         (const double*)b.part, nblk, b.rz, n, b.st);
      LAKER_KCHECK("pcg_beta_update_p1_kernel");
      return LAKER_OK;
    }
    LAKER_TRY(coldots<float>(b, 1, b.r, b.z, nullptr, nullptr, nullptr, nullptr, n, k, true, st));
    hipLaunchKernelGGL(pcg_beta_kernel, dim3(1), dim3(256), 0, st, b.part, nblk, k, b.rz, b.beta, b.st, 0);
    LAKER_KCHECK("pcg_beta_kernel");
    hipLaunchKernelGGL((pcg_update_p_kernel<float>), eg, eb, 0, st, b.p, (const float*)b.z,
                       (const double*)b.beta, n, k, (const PcgState*)b.st);
    LAKER_KCHECK("pcg_update_p_kernel");
    return LAKER_OK;
  };
  IterGraph ig;
  const bool graphs = g_use_graphs.load() != 0;
  bool eager_done = false;
  while (true) {
    // what the step must gain on the recursive residual to reach 0.6 tol: a
    // step has lowered the FP64 true residual at least as much as its
    // recursive one (20-25x for a 20x recursive drop), so the last step aims
    // just below the target instead of overshooting it
    const double need = 0.6 * tol * bnorm / rnorm;
    const double seg_tol = std::min(0.5, final_step ? need : std::max(inner_tol, need));
    hipLaunchKernelGGL(scale_to_f32_kernel, eg, eb, 0, st, Rin, b.r, nk, 1.0 / rnorm);
    LAKER_KCHECK("scale_to_f32_kernel");
    if (pc) LAKER_TRY(papply(b.r, b.z));
    else hipLaunchKernelGGL((copy_kernel<float>), eg, eb, 0, st, (const float*)b.r, b.z, nk);
    hipLaunchKernelGGL((copy_kernel<float>), eg, eb, 0, st, (const float*)b.z, b.p, nk);
    LAKER_KCHECK("copy_kernel");
    LAKER_TRY(coldots<float>(b, 2, b.r, b.z, b.r, b.r, nullptr, nullptr, n, k, false, st));
    hipLaunchKernelGGL(pcg_setup_kernel, dim3(1), dim3(256), 0, st, b.part, nblk, k, b.rz, b.st, seg_tol,
                       eps, it, inner_max_iter, rnorm);
    LAKER_KCHECK("pcg_setup_kernel");
    const int seg_start = it;
    while (it < inner_max_iter) {
      const int nb = std::min(check_every, inner_max_iter - it);
      // the solve's first batch runs eagerly (it sizes every work buffer);
      // later batches replay one captured graph (all state is on the device)
      LAKER_TRY(ig.run(body, nb, check_every, graphs && eager_done, st));
      eager_done = true;
      LAKER_HIP_CHECK(hipMemcpyAsync(&hs, b.st, sizeof(hs), hipMemcpyDeviceToHost, st));
      LAKER_HIP_CHECK(hipStreamSynchronize(st));
      it = hs.it;
      if (hs.done != PCG_RUNNING) break;
    }
    LAKER_HIP_CHECK(hipMemcpyAsync(&hs, b.st, sizeof(hs), hipMemcpyDeviceToHost, st));
    LAKER_HIP_CHECK(hipStreamSynchronize(st));
    if (hs.done != PCG_RUNNING) it = hs.iterations;  // launches after the stop were no-ops
    result->inner_iterations = it;
    if (hs.done == PCG_BREAKDOWN && it == seg_start) {
      // no progress at all in this step: the operator is indefinite here
      result->status = LAKER_EBREAKDOWN;
      return fail(LAKER_EBREAKDOWN,
                  "PCG breakdown: non-positive curvature detected (p^T A p <= 0). The operator may "
                  "be indefinite or the preconditioner may be unsuitable.");
    }
    // A breakdown after some progress ends the step at the last good iterate:
    // the reduced-precision operator can be slightly indefinite (rounding of
    // order u ||K|| against lambda), the FP64 residual below is exact.
    // FP64 residual R = B - A X
    LAKER_TRY(fused_product_f64<double>(op, op.rows(), op.cols(), X, k, k, R, k, op.lambda_reg, X, k,
                                        clamp, B, st));
    LAKER_TRY(frob_norm(R, nk, tmp, &rnorm, st));
    result->outer_iterations += 1;
    if (rnorm / bnorm <= tol) {
      result->converged = 1;
      result->residual_norm = rnorm;
      break;
    }
    if (std::isfinite(rnorm) && rnorm < rbest) {
      rbest = rnorm;
      have_best = true;
      LAKER_HIP_CHECK(hipMemcpyAsync(Xbest, X, nk * 8, hipMemcpyDeviceToDevice, st));
      LAKER_HIP_CHECK(hipMemcpyAsync(Rbest, R, nk * 8, hipMemcpyDeviceToDevice, st));
      if (final_step) break;
    } else {
      // no improvement: back to the best iterate (or to zero)
      if (have_best) {
        LAKER_HIP_CHECK(hipMemcpyAsync(X, Xbest, nk * 8, hipMemcpyDeviceToDevice, st));
        LAKER_HIP_CHECK(hipMemcpyAsync(R, Rbest, nk * 8, hipMemcpyDeviceToDevice, st));
      } else {
        LAKER_HIP_CHECK(hipMemsetAsync(X, 0, nk * 8, st));
        LAKER_HIP_CHECK(hipMemcpyAsync(R, B, nk * 8, hipMemcpyDeviceToDevice, st));
      }
      rnorm = rbest;
      if (final_step) break;
      final_step = true;
    }
    result->residual_norm = rnorm;
    if (result->outer_iterations >= max_outer || it >= inner_max_iter || rnorm == 0.0) break;
    Rin = R;
  }
  result->residual_norm = std::min(rnorm, rbest);
  result->inner_iterations = it;
  return LAKER_OK;
}

int laker_batch_create(int32_t max_systems, hipStream_t stream, laker_batch_t* out) {
  LAKER_CHECK_ARG(out != nullptr && max_systems >= 1 && max_systems <= 4096, "bad batch arguments");
  Batch* bt = new Batch();
  bt->st = stream;
  bt->max_sys = max_systems;
  bt->slot.assign(max_systems, nullptr);
  bt->hs.resize(max_systems);
  const int rc = bt->states.ensure((size_t)max_systems * sizeof(PcgState));
  if (rc != LAKER_OK) {
    delete bt;
    return rc;
  }
  *out = reinterpret_cast<laker_batch_t>(bt);
  return LAKER_OK;
}

int laker_batch_destroy(laker_batch_t h) {
  Batch* bt = reinterpret_cast<Batch*>(h);
  if (bt) (void)hipStreamSynchronize(bt->st);
  delete bt;
  return LAKER_OK;
}

int laker_batch_add(laker_batch_t h, laker_operator_t oh, laker_precond_t ph, const double* B, int32_t k, double* X,
                    double tol, int32_t max_outer, double inner_tol, int32_t inner_max_iter, int32_t* slot_out) {
  LAKER_CHECK_ARG(h && oh && ph && B && X && slot_out, "bad batch_add arguments");
  LAKER_CHECK_ARG(k >= 2 && k <= 32, "batched solves take 2 <= k <= 32 right-hand sides");
  LAKER_CHECK_ARG(inner_tol > 0.0 && inner_tol < 1.0, "inner_tol must lie in (0, 1)");
  Batch& bt = *reinterpret_cast<Batch*>(h);
  Operator& op = *reinterpret_cast<Operator*>(oh);
  Precond& pc = *reinterpret_cast<Precond*>(ph);
  LAKER_CHECK_ARG(op.dtype == LAKER_F32, "batched solves need an FP32 operator");
  LAKER_CHECK_ARG(precond_rows(pc) == op.n, "preconditioner and operator sizes differ");
  int slot = -1;
  for (int i = 0; i < bt.max_sys; ++i)
    if (!bt.slot[i]) {
      slot = i;
      break;
    }
  LAKER_CHECK_ARG(slot >= 0, "batch is full");
  std::unique_ptr<BatchSys> sp(new BatchSys());
  BatchSys& s = *sp;
  s.op = &op;
  s.pc = &pc;
  s.B = B;
  s.X = X;
  s.k = k;
  s.n = op.n;
  s.tol = tol;
  s.max_outer = max_outer;
  s.inner_tol = inner_tol;
  s.inner_max_iter = inner_max_iter;
  if (precond_batch_view(pc, k, &s.av) != LAKER_OK)
    return fail(LAKER_EINVAL, "batched solves need an FP32 CCCP preconditioner with N_r % 4 == 0");
  // the product of product_multi_f32 on the 3-way split (one 32-column chunk)
  const int64_t n = op.n, nk = n * k;
  const int ns16h = op.ka16h / 16, ns16 = op.ka16 / 16;
  const bool f16s = op.f16s_ok && (ns16h == 2 || ns16h == 3);
  LAKER_CHECK_ARG(f16s || ns16 == 2 || ns16 == 4 || ns16 == 6, "no batched product for this embedding width");
  s.gm = f16s ? GRAM_F16S : LAKER_PREC_MIXED;
  s.ns = f16s ? ns16h : ns16;
  s.colpack = f16s ? op.self.col16h.ptr : op.self.col16.ptr;
  s.rowpack = f16s ? op.self.row16h.ptr : op.self.row16.ptr;
  s.ka = f16s ? op.ka16h : op.ka16;
  s.clamp = op.clamp_for(LAKER_CLAMP_AUTO);
  // column splits of a batched product: the batch fills the chip with its
  // systems, so a system needs few splits (unlike a launch of its own, which
  // takes up to 64); a function of n alone, so a system's partial sums do not
  // depend on the batch
  {
    const int64_t row_blocks = ceil_div(n, 128), tiles = ceil_div(n, 32);
    int js = (int)std::max<int64_t>(1, std::min<int64_t>(8, 256 / row_blocks));
    while (js > 1 && tiles / js < 2) js /= 2;
    s.jsplit = js;
  }
  LAKER_TRY(op.ensure_part((size_t)s.jsplit * n * 32 * sizeof(float)));
  s.ntiles = ceil_div(n, 32);
  const size_t part_bytes = (size_t)s.ntiles * 128 * 2 * sizeof(bf16x8);
  LAKER_TRY(op.xpack.ensure(3 * part_bytes));
  s.xh = op.xpack.as<bf16x8>();
  s.xl = reinterpret_cast<bf16x8*>(op.xpack.as<char>() + part_bytes);
  s.xm = reinterpret_cast<bf16x8*>(op.xpack.as<char>() + 2 * part_bytes);
  LAKER_TRY(pcg_buffers<float>(op, k, s.b));
  s.b.st = bt.states.as<PcgState>() + slot;
  const size_t vb = round_up(nk * 8, 256);
  LAKER_TRY(op.mixed_ws.ensure(3 * vb));
  s.R = op.mixed_ws.as<double>();
  s.Xbest = reinterpret_cast<double*>(op.mixed_ws.as<char>() + vb);
  s.Rbest = reinterpret_cast<double*>(op.mixed_ws.as<char>() + 2 * vb);
  s.res = laker_mixed_result{};
  LAKER_TRY(frob_norm(B, nk, s.tmp, &s.bnorm, bt.st));
  s.res.rhs_norm = s.bnorm;
  s.res.residual_norm = s.bnorm;
  LAKER_HIP_CHECK(hipMemsetAsync(X, 0, nk * 8, bt.st));
  s.rnorm = s.rbest = s.bnorm;
  s.Rin = B;
  if (s.bnorm == 0.0) {
    s.res.converged = 1;
    s.res.residual_norm = 0.0;
    s.finished = true;
  } else {
    LAKER_TRY(batch_step_begin(bt, s));
  }
  bt.slot[slot] = sp.release();
  *slot_out = slot;
  return LAKER_OK;
}

int laker_batch_step(laker_batch_t h, int32_t iterations, int32_t* finished, int32_t* nfinished, int32_t* nactive) {
  LAKER_CHECK_ARG(h && finished && nfinished && iterations >= 1, "bad batch_step arguments");
  Batch& bt = *reinterpret_cast<Batch*>(h);
  std::vector<BatchSys*> act;
  for (BatchSys* s : bt.slot)
    if (s && s->running && !s->finished) act.push_back(s);
  if (!act.empty()) {
    LAKER_TRY(batch_iterate(bt, act, iterations));
    LAKER_HIP_CHECK(hipMemcpyAsync(bt.hs.data(), bt.states.ptr, bt.max_sys * sizeof(PcgState), hipMemcpyDeviceToHost,
                                   bt.st));
    LAKER_HIP_CHECK(hipStreamSynchronize(bt.st));
    for (int i = 0; i < bt.max_sys; ++i) {
      BatchSys* s = bt.slot[i];
      if (!s || !s->running || s->finished) continue;
      const PcgState& hs = bt.hs[i];
      s->it = hs.it;
      if (hs.done != PCG_RUNNING) LAKER_TRY(batch_finish_step(bt, *s, hs));
    }
  }
  int nf = 0, na = 0;
  for (int i = 0; i < bt.max_sys; ++i) {
    BatchSys* s = bt.slot[i];
    if (!s) continue;
    if (s->finished) finished[nf++] = i;
    else ++na;
  }
  if (nf) LAKER_HIP_CHECK(hipStreamSynchronize(bt.st));  // finished solutions are complete in memory
  *nfinished = nf;
  if (nactive) *nactive = na;
  return LAKER_OK;
}

int laker_batch_result(laker_batch_t h, int32_t slot, laker_mixed_result* result) {
  LAKER_CHECK_ARG(h && result, "bad batch_result arguments");
  Batch& bt = *reinterpret_cast<Batch*>(h);
  LAKER_CHECK_ARG(slot >= 0 && slot < bt.max_sys && bt.slot[slot], "no system in this batch slot");
  BatchSys* s = bt.slot[slot];
  LAKER_CHECK_ARG(s->finished, "the system in this slot has not finished");
  *result = s->res;
  delete s;
  bt.slot[slot] = nullptr;
  if (result->status == LAKER_EBREAKDOWN)
    return fail(LAKER_EBREAKDOWN,
                "PCG breakdown: non-positive curvature detected (p^T A p <= 0). The operator may "
                "be indefinite or the preconditioner may be unsuitable.");
  return LAKER_OK;
}

}  // extern "C"
