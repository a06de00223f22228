// CCCP preconditioner P = Sigma^{-1/2}, Sigma = a I + Q C Q^T, built and applied on device.
//
// Build (restates laker/preconditioner.py:144-284):
//   U = A R (probe matvec, k = N_r), column normalisation clamped at epsilon,
//   QR by Cholesky-QR2 in FP64 (shifted Cholesky-QR3 when the first Gram
//   factorisation fails), then the shrinkage-regularised CCCP fixed point in
//   FP64 with diag(R^T M^{-1} R) taken from a Cholesky factor of M = aI + C
//   (M = t F_rho is SPD, so this equals the reference's clamped eigh while the
//   clamp is inactive), and finally S = M^{-1/2} - a^{-1/2} I by a coupled
//   Newton-Schulz iteration (GEMMs only).
// Apply (restates laker/preconditioner.py:52-68):
//   Z = a^{-1/2} X + Q (S (Q^T X)), two HBM-bound passes over Q.
#include <cmath>
#include <memory>
#include <vector>

#include "dense64.h"
#include "laker_common.h"
#include "laker_internal.h"
#include "batch.h"

namespace laker {

enum PrecondKind { PRECOND_CCCP = 0, PRECOND_JACOBI = 1 };

struct Precond {
  int kind = PRECOND_CCCP;
  int64_t n = 0;
  int nr = 0;
  int dtype = LAKER_F32;
  double a = 1.0, rho = 0.0, inv_sqrt_a = 1.0;
  DevBuf Q, R, C, S;
  // FP16 copy of Q and the matching S' (made on first use by the mixed solve):
  // P' = a^{-1/2} I + Qh S' Qh^T with S' = L^{-T} S L^{-1}, L L^T = Qh^T Qh
  DevBuf Qh, Sh;
  bool has_h = false;
  DevBuf w1, w2, part, ws, scal;
  DevBuf inv_diag;  // Jacobi: 1 / max(diag, eps(dtype)), n entries of dtype
};

int64_t precond_rows(const Precond& pc) { return pc.n; }
int precond_dtype(const Precond& pc) { return pc.dtype; }

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------

// splitmix64-based standard normals (Box-Muller); used only when the caller
// passes no probe matrix.
__device__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
// bf16_exact: round each draw to the nearest BF16 value (the probe product of
// the mixed build then needs no lo part of the probes).
template <class T>
__global__ void randn_kernel(T* __restrict__ out, int64_t count, uint64_t seed, int bf16_exact = 0) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint64_t a = splitmix64(seed * 0x632BE59BD9B4E019ull + 2 * (uint64_t)i);
  const uint64_t b = splitmix64(seed * 0x632BE59BD9B4E019ull + 2 * (uint64_t)i + 1);
  const double u1 = ((a >> 11) + 1.0) * (1.0 / 9007199254740993.0);
  const double u2 = (b >> 11) * (1.0 / 9007199254740992.0);
  const double g = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  // one rounding from the FP64 draw to BF16 (not FP64 -> FP32 -> BF16)
  out[i] = bf16_exact ? (T)bf16_bits_to_f32(__builtin_bit_cast(uint16_t, static_cast<__bf16>(g))) : (T)g;
}

// Column sums of squares of an n x m row-major matrix -> out[m] (FP64).  Sum
// order per column c: 256 strided chains (chain t takes rows t, t + 256,
// t + 512, ... in order), then block_sum over the 256 chain sums.  Pass 1
// (colsumsq_chains_kernel) runs the chains with 64 consecutive columns per
// wave, so every load of a wave is one contiguous row segment; pass 2
// (colsumsq_final_kernel, one block per column) forms the block sum.
constexpr int CSQ_CHAINS = 256;

template <class T>
__global__ __launch_bounds__(256) void colsumsq_chains_kernel(const T* __restrict__ U, int64_t n, int m,
                                                              double* __restrict__ chains) {
  const int c = blockIdx.x * 64 + (threadIdx.x & 63);
  const int t = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (c >= m) return;
  double s = 0.0;
#pragma unroll 8
  for (int64_t i = t; i < n; i += CSQ_CHAINS) {
    const double v = (double)U[i * m + c];
    s += v * v;
  }
  chains[(int64_t)t * m + c] = s;
}

__global__ __launch_bounds__(CSQ_CHAINS) void colsumsq_final_kernel(const double* __restrict__ chains, int m,
                                                                    double* __restrict__ out) {
  __shared__ double sm[16];
  const int c = blockIdx.x;
  double s = block_sum(chains[(int64_t)threadIdx.x * m + c], sm);
  if (threadIdx.x == 0) out[c] = s;
}

template <class T>
static int colsumsq(const T* U, int64_t n, int m, double* out, DevBuf& chains, hipStream_t st) {
  LAKER_TRY(chains.ensure((size_t)CSQ_CHAINS * m * sizeof(double), st));
  hipLaunchKernelGGL((colsumsq_chains_kernel<T>), dim3((unsigned)ceil_div(m, 64), CSQ_CHAINS / 4), dim3(256), 0, st,
                     U, n, m, chains.as<double>());
  LAKER_KCHECK("colsumsq_chains_kernel");
  hipLaunchKernelGGL(colsumsq_final_kernel, dim3((unsigned)m), dim3(CSQ_CHAINS), 0, st,
                     (const double*)chains.as<double>(), m, out);
  LAKER_KCHECK("colsumsq_final_kernel");
  return LAKER_OK;
}

// U[:, c] /= max(sqrt(ss[c]), eps) in U's dtype (like the reference), and
// V = (double)U when V is given.  grid.y = row blocks of 64 rows
template <class T>
__global__ void colscale_kernel(T* __restrict__ U, int64_t n, int m, const double* __restrict__ ss,
                                double eps, double* __restrict__ V) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m) return;
  const T nrm = (T)sqrt(ss[c]);
  const T den = nrm > (T)eps ? nrm : (T)eps;
  const int64_t i0 = (int64_t)blockIdx.y * 64, i1 = i0 + 64 < n ? i0 + 64 : n;
  for (int64_t i = i0; i < i1; ++i) {
    const T u = U[i * m + c] / den;
    U[i * m + c] = u;
    if (V) V[i * m + c] = (double)u;
  }
}

template <class TI, class TO>
__global__ void convert_kernel(const TI* __restrict__ in, TO* __restrict__ out, int64_t count) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = (TO)in[i];
}

__global__ void add_diag_kernel(double* __restrict__ A, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[i * n + i] += v;
}

// Jacobi preconditioner (laker/solvers.py:505-546): inv = 1 / max(d, eps) in the
// solver dtype (a NaN diagonal stays NaN, as torch's clamp leaves it), and
// Z = inv * X row-wise for every column.
template <class T>
__global__ void jacobi_inv_kernel(const T* __restrict__ diag, T* __restrict__ inv, int64_t n, T eps) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T d = diag[i];
  inv[i] = T(1) / (d < eps ? eps : d);
}
template <class T>
__global__ void jacobi_apply_kernel(const T* __restrict__ inv, const T* __restrict__ X, int64_t ldx,
                                    T* __restrict__ Z, int64_t ldz, int64_t n, int k) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * k) return;
  const int64_t i = idx / k;
  const int c = (int)(idx % k);
  Z[i * ldz + c] = inv[i] * X[i * ldx + c];
}

// M = C + a I
__global__ void form_m_kernel(const double* __restrict__ C, double* __restrict__ M, int64_t n,
                              double a) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  M[idx] = C[idx] + ((idx / n == idx % n) ? a : 0.0);
}

// w[c] = (n/nr) / (sum_r Z[r][c]^2 + eps);  Rw[r][c] = R[r][c] * w[c]; grid = nr
__global__ void cccp_weights_kernel(const double* __restrict__ Z, const double* __restrict__ R,
                                    double* __restrict__ Rw, int nr, double n_over_nr, double eps) {
  __shared__ double sm[16];
  const int c = blockIdx.x;
  double s = 0.0;
  for (int r = threadIdx.x; r < nr; r += blockDim.x) {
    const double v = Z[(int64_t)r * nr + c];
    s += v * v;
  }
  s = block_sum(s, sm);
  const double w = n_over_nr / (s + eps);
  for (int r = threadIdx.x; r < nr; r += blockDim.x) Rw[(int64_t)r * nr + c] = R[(int64_t)r * nr + c] * w;
}

// Scalars of one CCCP iteration, kept on device.
struct CccpScalars {
  double trace;   // tr(F_rho)
  double t, a_new;
  double diff2, c2; // ||C_new - C||_F^2, ||C||_F^2
};

// trace of F_rho = sum_i ((1-rho)(H_ii + gamma)/s + rho); single block
__global__ void cccp_trace_kernel(const double* __restrict__ H, int nr, double gamma, double s,
                                  double rho, double a_rho, int64_t n, CccpScalars* sc) {
  __shared__ double sm[16];
  double v = 0.0;
  for (int i = threadIdx.x; i < nr; i += blockDim.x)
    v += (1.0 - rho) * (H[(int64_t)i * nr + i] + gamma) / s + rho;
  v = block_sum(v, sm);
  if (threadIdx.x == 0) {
    sc->trace = v;
    sc->t = (double)n / (a_rho * (double)(n - nr) + v);
    sc->a_new = sc->t * a_rho;
  }
}

// C_new = t ((1-rho)(H + gamma I)/s + rho I) - a_new I; partial norms per block
__global__ void cccp_update_kernel(const double* __restrict__ H, const double* __restrict__ C,
                                   double* __restrict__ Cn, int nr, double gamma, double s,
                                   double rho, const CccpScalars* __restrict__ sc,
                                   double* __restrict__ part) {
  __shared__ double sm[16];
  const double t = sc->t, an = sc->a_new;
  const int64_t total = (int64_t)nr * nr;
  double d2 = 0.0, c2 = 0.0;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const bool dg = (idx / nr) == (idx % nr);
    const double f = (1.0 - rho) * (H[idx] + (dg ? gamma : 0.0)) / s + (dg ? rho : 0.0);
    const double cn = t * f - (dg ? an : 0.0);
    const double co = C[idx];
    Cn[idx] = cn;
    d2 += (cn - co) * (cn - co);
    c2 += co * co;
  }
  d2 = block_sum(d2, sm);
  c2 = block_sum(c2, sm);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = d2;
    part[2 * blockIdx.x + 1] = c2;
  }
}

__global__ void cccp_norms_final_kernel(const double* __restrict__ part, int nb, CccpScalars* sc) {
  if (threadIdx.x == 0) {
    double d2 = 0.0, c2 = 0.0;
    for (int b = 0; b < nb; ++b) {
      d2 += part[2 * b];
      c2 += part[2 * b + 1];
    }
    sc->diff2 = d2;
    sc->c2 = c2;
  }
}

// ||A||_F^2 (single block) -> out
__global__ void frob2_kernel(const double* __restrict__ A, int64_t count, double* __restrict__ out) {
  __shared__ double sm[16];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < count; i += blockDim.x) s += A[i] * A[i];
  s = block_sum(s, sm);
  if (threadIdx.x == 0) *out = s;
}

// Newton-Schulz step helper: given P = Z Y, T = (3I - P)/2 ; err2 = ||P - I||_F^2 (block partials)
__global__ void ns_t_kernel(const double* __restrict__ P, double* __restrict__ T, int64_t n,
                            double* __restrict__ part) {
  __shared__ double sm[16];
  double e2 = 0.0;
  const int64_t total = n * n;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const bool dg = (idx / n) == (idx % n);
    const double p = P[idx];
    const double dv = p - (dg ? 1.0 : 0.0);
    e2 += dv * dv;
    T[idx] = ((dg ? 3.0 : 0.0) - p) * 0.5;
  }
  e2 = block_sum(e2, sm);
  if (threadIdx.x == 0) part[blockIdx.x] = e2;
}

__global__ void sum_parts_kernel(const double* __restrict__ part, int nb, double* __restrict__ out) {
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[b];
    *out = s;
  }
}

// Device state of the Newton-Schulz loop: status 0 running, 1 converged,
// 2 failed (NaN).
struct NsState {
  double prev_err;
  int status;
  int iterations;
};

__global__ void ns_init_kernel(NsState* s) {
  s->prev_err = 1e300;
  s->status = 0;
  s->iterations = 0;
}

// The stopping rule of one Newton-Schulz step from the partials of
// err^2 = ||Z Y - I||_F^2 (summed in block order): stop once err < tol_abs,
// or once the error stalls below 1e-10 (err > prev_err / 2).
__global__ void ns_check_kernel(const double* __restrict__ part, int nb, double tol_abs, NsState* s) {
  if (threadIdx.x != 0 || s->status != 0) return;
  double e2 = 0.0;
  for (int b = 0; b < nb; ++b) e2 += part[b];
  const double err = sqrt(e2);
  if (!(err == err)) {
    s->status = 2;
  } else if (err < tol_abs || (err > 0.5 * s->prev_err && err < 1e-10)) {
    s->status = 1;
  } else {
    s->prev_err = err;
    s->iterations += 1;
  }
}

// T = I once the loop has stopped: the steps launched after the stopping
// point then leave Y and Z exactly as they were.
__global__ void ns_hold_kernel(double* __restrict__ T, int64_t n, const NsState* __restrict__ s) {
  if (s->status == 0) return;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  T[idx] = (idx / n == idx % n) ? 1.0 : 0.0;
}

// Y = M / sqrt(fro2)
__global__ void scale_by_frob_kernel(double* __restrict__ M, int64_t n, const double* __restrict__ fro2) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const double f = 1.0 / sqrt(*fro2);
  M[idx] = M[idx] * f;
}

// S[i][j] = (Z[i][j] + Z[j][i]) / (2 sqrt(sqrt(fro2))) - g delta_ij
__global__ void sym_scale_frob_kernel(const double* __restrict__ Zm, double* __restrict__ S, int64_t n,
                                      const double* __restrict__ fro2, double g) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const int64_t i = idx / n, j = idx % n;
  const double f = 1.0 / sqrt(sqrt(*fro2));
  S[idx] = 0.5 * f * (Zm[idx] + Zm[j * n + i]) - (i == j ? g : 0.0);
}

// S = Z * f - g I
__global__ void scale_sub_diag_kernel(const double* __restrict__ Zm, double* __restrict__ S, int64_t n,
                                      double f, double g) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  S[idx] = Zm[idx] * f - ((idx / n == idx % n) ? g : 0.0);
}

// S[i][j] = f (Z[i][j] + Z[j][i]) / 2 - g delta_ij
__global__ void sym_scale_kernel(const double* __restrict__ Zm, double* __restrict__ S, int64_t n,
                                 double f, double g) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const int64_t i = idx / n, j = idx % n;
  S[idx] = 0.5 * f * (Zm[idx] + Zm[j * n + i]) - (i == j ? g : 0.0);
}

// ---------------------------------------------------------------------------
// apply kernels (k == 1)
// ---------------------------------------------------------------------------
// partial[b][c] = sum_{i in chunk b} Q[i][c] x[i]
template <class T>
__global__ __launch_bounds__(256) void qtx_vec_kernel(const T* __restrict__ Q, const T* __restrict__ x,
                                                      int64_t n, int nr, int64_t rows_per_block,
                                                      double* __restrict__ part) {
  // threads own columns c = c0 + u*256 + tid; rows are consumed in batches of
  // RB so that RB*U independent loads are in flight per thread.
  constexpr int U = 4, RB = 4;
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r1 = (r0 + rows_per_block < n) ? r0 + rows_per_block : n;
  for (int c0 = 0; c0 < nr; c0 += 256 * U) {
    double acc[U] = {0.0, 0.0, 0.0, 0.0};
    int64_t i = r0;
    for (; i + RB <= r1; i += RB) {
      T qv[RB][U];
      double xi[RB];
#pragma unroll
      for (int b = 0; b < RB; ++b) {
        xi[b] = (double)x[i + b];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int c = c0 + u * 256 + threadIdx.x;
          qv[b][u] = (c < nr) ? Q[(i + b) * nr + c] : (T)0;
        }
      }
#pragma unroll
      for (int b = 0; b < RB; ++b)
#pragma unroll
        for (int u = 0; u < U; ++u) acc[u] = fma((double)qv[b][u], xi[b], acc[u]);
    }
    for (; i < r1; ++i) {
      const double xv = (double)x[i];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int c = c0 + u * 256 + threadIdx.x;
        if (c < nr) acc[u] = fma((double)Q[i * nr + c], xv, acc[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int c = c0 + u * 256 + threadIdx.x;
      if (c < nr) part[(int64_t)blockIdx.x * nr + c] = acc[u];
    }
  }
}

// Q^T x for an FP32 Q with nr % 4 == 0: part[b][c] = sum_{i in chunk b} Q[i][c] x[i].
// Wave w of the block takes rows r0+w, r0+w+4, ... (four rows in flight);
// lane l owns the column quads l, l+64, ... (QPL quads, 16-byte loads).  The
// four waves' partial sums are combined in LDS in a fixed order.
template <int QPL>
__global__ __launch_bounds__(256) void qtx4_kernel(const float* __restrict__ Q, const float* __restrict__ x,
                                                   int64_t n, int nr, int64_t rows_per_block,
                                                   double* __restrict__ part) {
  __shared__ double red[4][QPL * 256];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nq = nr >> 2;
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r1 = (r0 + rows_per_block < n) ? r0 + rows_per_block : n;
  double acc[QPL][4];
#pragma unroll
  for (int u = 0; u < QPL; ++u) acc[u][0] = acc[u][1] = acc[u][2] = acc[u][3] = 0.0;
  int64_t i = r0 + w;
  for (; i + 12 < r1; i += 16) {
    floatx4 q[4][QPL];
    float xv[4];
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      xv[rr] = x[i + 4 * rr];
      const floatx4* row = reinterpret_cast<const floatx4*>(Q + (i + 4 * rr) * nr);
#pragma unroll
      for (int u = 0; u < QPL; ++u) {
        const int qi = lane + 64 * u;
        q[rr][u] = (qi < nq) ? row[qi] : floatx4{0.f, 0.f, 0.f, 0.f};
      }
    }
#pragma unroll
    for (int rr = 0; rr < 4; ++rr)
#pragma unroll
      for (int u = 0; u < QPL; ++u)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[u][j] = fma((double)q[rr][u][j], (double)xv[rr], acc[u][j]);
  }
  for (; i < r1; i += 4) {
    const double xv = (double)x[i];
    const floatx4* row = reinterpret_cast<const floatx4*>(Q + i * nr);
#pragma unroll
    for (int u = 0; u < QPL; ++u) {
      const int qi = lane + 64 * u;
      if (qi < nq) {
        const floatx4 q = row[qi];
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[u][j] = fma((double)q[j], xv, acc[u][j]);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < QPL; ++u)
#pragma unroll
    for (int j = 0; j < 4; ++j) red[w][(lane + 64 * u) * 4 + j] = acc[u][j];
  __syncthreads();
  for (int c = threadIdx.x; c < nr; c += 256)
    part[(int64_t)blockIdx.x * nr + c] = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
}

// out[g][c] = sum_{b in group g} part[b][c] with groups of ceil(nb / gridDim.y)
// rows; block = 64 columns x 4 row phases, fixed order.
__device__ __forceinline__ void reduce_parts_groups_kernel_body(const double* __restrict__ part, int nb,
                                                                  int nr, double* __restrict__ out, int bx_, int by_, int gy_) {
  __shared__ double sm[4][64];
  const int cx = threadIdx.x & 63, cy = threadIdx.x >> 6;
  const int c = bx_ * 64 + cx;
  const int gsz = (nb + gy_ - 1) / gy_;
  const int b0 = by_ * gsz, b1 = min(nb, b0 + gsz);
  double s = 0.0;
  if (c < nr) {
    // same additions in the same order; the unroll only keeps loads in flight
#pragma unroll 8
    for (int b = b0 + cy; b < b1; b += 4) s += part[(int64_t)b * nr + c];
  }
  sm[cy][cx] = s;
  __syncthreads();
  if (cy == 0 && c < nr) out[(int64_t)by_ * nr + c] = (sm[0][cx] + sm[1][cx]) + (sm[2][cx] + sm[3][cx]);
}

__global__ __launch_bounds__(256) void reduce_parts_groups_kernel(const double* __restrict__ part, int nb,
                                                                  int nr, double* __restrict__ out) {
  reduce_parts_groups_kernel_body(part, nb, nr, out, (int)blockIdx.x, (int)blockIdx.y, (int)gridDim.y);
}

// z[i] = f * x[i] + sum_c Q[i][c] w[c] for an FP32 Q with nr % 4 == 0: one
// wave per QW_ROWS rows, 16-byte loads, w staged in LDS.
template <int QW>
__global__ __launch_bounds__(256) void qw4_kernel(const float* __restrict__ Q, const double* __restrict__ w,
                                                  const float* __restrict__ x, float* __restrict__ z,
                                                  int64_t n, int nr, double f) {
  extern __shared__ double ws4_lds[];
  for (int c = threadIdx.x; c < nr; c += blockDim.x) ws4_lds[c] = w[c];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t i0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * QW;
  if (i0 >= n) return;
  const int nq = nr >> 2;
  double s[QW];
#pragma unroll
  for (int b = 0; b < QW; ++b) s[b] = 0.0;
  for (int qi = lane; qi < nq; qi += 64) {
    const double w0 = ws4_lds[4 * qi], w1 = ws4_lds[4 * qi + 1], w2 = ws4_lds[4 * qi + 2],
                 w3 = ws4_lds[4 * qi + 3];
    floatx4 q[QW];
#pragma unroll
    for (int b = 0; b < QW; ++b)
      q[b] = (i0 + b < n) ? reinterpret_cast<const floatx4*>(Q + (i0 + b) * nr)[qi]
                          : floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < QW; ++b) {
      s[b] = fma((double)q[b][0], w0, s[b]);
      s[b] = fma((double)q[b][1], w1, s[b]);
      s[b] = fma((double)q[b][2], w2, s[b]);
      s[b] = fma((double)q[b][3], w3, s[b]);
    }
  }
#pragma unroll
  for (int b = 0; b < QW; ++b) {
    const double t = wave_sum(s[b]);
    if (lane == 0 && i0 + b < n) z[i0 + b] = (float)(f * (double)x[i0 + b] + t);
  }
}

// Q^T X for an FP32 Q and 2 <= k <= 32 columns:
// part[s][c][j] = sum over the rows of split s of Q[i][c] X[i][j] (FP64),
// c in the block's 32 columns.  v_mfma_f32_32x32x2f32 with lane (li, h)
// supplying Q[i + h][c0 + li] (A[m = li][kk = h]) and X[i + h][li]
// (B[kk = h][n = li]); FP32 accumulation over at most 64 row pairs, then FP64.
// The four waves take interleaved row pairs and are combined in LDS in a fixed
// order.
__device__ __forceinline__ void qtx_mfma_kernel_body(const float* __restrict__ Q, int64_t n, int nr,
                                                       const float* __restrict__ X, int64_t ldx, int k,
                                                       int64_t rows_per_split, double* __restrict__ part, int bx_, int by_) {
  __shared__ double red[4][32][33];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 31, h = lane >> 5;
  const int c0 = bx_ * 32;
  const int64_t s = by_;
  const int64_t r0 = s * rows_per_split;
  const int64_t r1 = (r0 + rows_per_split < n) ? r0 + rows_per_split : n;
  floatx16 acc = {0};
  double acc64[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) acc64[r] = 0.0;
  // eight row pairs per step: all loads first, then the MFMAs
  constexpr int U = 8;
  int cnt = 0;
  for (int64_t i = r0 + 2 * w; i < r1; i += 8 * U) {
    float a[U], b[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t row = i + 8 * u + h;
      a[u] = (row < r1 && c0 + li < nr) ? Q[row * nr + c0 + li] : 0.f;
      b[u] = (row < r1 && li < k) ? X[row * ldx + li] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[u], b[u], acc, 0, 0, 0);
    cnt += U;
    if (cnt >= 64) {
#pragma unroll
      for (int r = 0; r < 16; ++r) acc64[r] += (double)acc[r];
      acc = floatx16{0};
      cnt = 0;
    }
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    acc64[r] += (double)acc[r];
    // D[m][n] at lane (li, h), register r: m = 8 (r/4) + 4 h + r%4 (Q column), n = li (X column)
    red[w][8 * (r >> 2) + 4 * h + (r & 3)][li] = acc64[r];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 32 * 32; e += 256) {
    const int m = e >> 5, j = e & 31;
    if (c0 + m < nr && j < k)
      part[(s * nr + c0 + m) * k + j] = (red[0][m][j] + red[1][m][j]) + (red[2][m][j] + red[3][m][j]);
  }
}

__global__ __launch_bounds__(256) void qtx_mfma_kernel(const float* __restrict__ Q, int64_t n, int nr,
                                                       const float* __restrict__ X, int64_t ldx, int k,
                                                       int64_t rows_per_split, double* __restrict__ part) {
  qtx_mfma_kernel_body(Q, n, nr, X, ldx, k, rows_per_split, part, (int)blockIdx.x, (int)blockIdx.y);
}

// Z[i][j0 + j] = f X[i][j0 + j] + sum_c Q[i][c] W2[c][j0 + j] for j < 16 NB, FP32 Q, on
// v_mfma_f32_16x16x4f32.  The block stages W2's columns j0 .. j0 + 16 NB - 1 in LDS
// as FP32, transposed (row j, stride P floats, zero past nr), and each wave owns
// QW_TILES tiles of 16 rows.  Lane (r = l % 16, g = l / 16) loads the quad
// Q[i][c0 + 4g .. c0 + 4g + 3] of row i = i0 + r with one 16-byte load; MFMA step t
// (t < 4) contracts over the columns c0 + 4g + t of the four lane groups:
// A[m = r][kk = g] = Q[i0 + r][c0 + 4g + t] and B[kk = g][n = r] = W2[c0 + 4g + t][j0 + r],
// the latter from one 16-byte LDS load per chunk of 16 columns.
// D[m][n] sits at lane (n = l % 16, g = l / 16), register v: m = 4g + v.
constexpr int QW_TILES = 2;
constexpr int QW_BLOCK_ROWS = 4 * 16 * QW_TILES;
static inline int qw_stride(int nr) {
  const int nr16 = (nr + 15) & ~15;
  return nr16 + (((4 - nr16) % 32) + 32) % 32;  // P = 4 (mod 32): conflict-free 16-byte reads
}
template <int NB>
__device__ __forceinline__ void qw_mfma_kernel_body(const float* __restrict__ Q, const double* __restrict__ W2,
                                                      int64_t n, int nr, int P, int k, int j0,
                                                      const float* __restrict__ X, int64_t ldx,
                                                      float* __restrict__ Z, int64_t ldz, double f, int bx_, int by_) {
  extern __shared__ float w2t[];  // 16 NB rows x P
  constexpr int JW = 16 * NB;
  for (int e = threadIdx.x; e < JW * P; e += blockDim.x) {
    const int c = e / JW, jj = e % JW;
    w2t[jj * P + c] = (c < nr && j0 + jj < k) ? (float)W2[(int64_t)c * k + j0 + jj] : 0.f;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  const int64_t i0 = (int64_t)bx_ * QW_BLOCK_ROWS + (threadIdx.x >> 6) * (16 * QW_TILES);
  if (i0 >= n) return;
  const float* qrow[QW_TILES];
#pragma unroll
  for (int tt = 0; tt < QW_TILES; ++tt) {
    const int64_t i = i0 + 16 * tt + r;
    qrow[tt] = Q + (i < n ? i : n - 1) * nr;
  }
  floatx4 acc[QW_TILES][NB];
#pragma unroll
  for (int tt = 0; tt < QW_TILES; ++tt)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[tt][nb] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
  for (int c0 = 0; c0 < nr; c0 += 16) {
    const int c = c0 + 4 * g;
    floatx4 q[QW_TILES], w[NB];
#pragma unroll
    for (int tt = 0; tt < QW_TILES; ++tt)
      q[tt] = (c < nr) ? *reinterpret_cast<const floatx4*>(qrow[tt] + c) : floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) w[nb] = *reinterpret_cast<const floatx4*>(w2t + (16 * nb + r) * P + c);
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int tt = 0; tt < QW_TILES; ++tt)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
          acc[tt][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(q[tt][t], w[nb][t], acc[tt][nb], 0, 0, 0);
  }
#pragma unroll
  for (int tt = 0; tt < QW_TILES; ++tt)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t i = i0 + 16 * tt + 4 * g + v;
        const int j = j0 + 16 * nb + r;
        if (i < n && j < k) Z[i * ldz + j] = (float)(f * (double)X[i * ldx + j] + (double)acc[tt][nb][v]);
      }
}

template <int NB>
__global__ __launch_bounds__(256) void qw_mfma_kernel(const float* __restrict__ Q, const double* __restrict__ W2,
                                                      int64_t n, int nr, int P, int k, int j0,
                                                      const float* __restrict__ X, int64_t ldx,
                                                      float* __restrict__ Z, int64_t ldz, double f) {
  qw_mfma_kernel_body<NB>(Q, W2, n, nr, P, k, j0, X, ldx, Z, ldz, f, (int)blockIdx.x, (int)blockIdx.y);
}

// out[c] = sum_b part[b][c]; block (64 columns x 4 row groups), fixed order per group
__global__ __launch_bounds__(256) void reduce_parts_cols_kernel(const double* __restrict__ part, int nb, int nr,
                                                                double* __restrict__ out) {
  __shared__ double sm[4][64];
  const int cx = threadIdx.x & 63, cy = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + cx;
  double s = 0.0;
  if (c < nr)
    for (int b = cy; b < nb; b += 4) s += part[(int64_t)b * nr + c];
  sm[cy][cx] = s;
  __syncthreads();
  if (cy == 0 && c < nr) out[c] = (sm[0][cx] + sm[1][cx]) + (sm[2][cx] + sm[3][cx]);
}

// y[r] = sum_c S[r][c] w[c]  (one wave per row of the nr x nr FP64 matrix)
__global__ __launch_bounds__(256) void gemv64_kernel(const double* __restrict__ S, const double* __restrict__ w,
                                                     double* __restrict__ y, int nr) {
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= nr) return;
  double s = 0.0;
  for (int c = lane; c < nr; c += 64) s = fma(S[(int64_t)r * nr + c], w[c], s);
  s = wave_sum(s);
  if (lane == 0) y[r] = s;
}

// y[r] = sum_c S[r][c] w1[c] with w1 = the sum of the 8 group partials of
// Q^T x, formed per block in LDS in the order of reduce_parts_groups_kernel
// ((g0 + g4) + (g1 + g5)) + ((g2 + g6) + (g3 + g7)), so the result matches
// the separate reduction bit for bit.  GS8_RPW rows per wave.
constexpr int GS8_RPW = 1;  // rows per wave (the row's sum does not depend on it)
__global__ __launch_bounds__(256) void gemv64_sum8_kernel(const double* __restrict__ S,
                                                          const double* __restrict__ part8,
                                                          double* __restrict__ y, int nr) {
  extern __shared__ double w1s[];
  for (int c = threadIdx.x; c < nr; c += 256) {
    const double* q = part8 + c;
    w1s[c] = ((q[0] + q[4 * nr]) + (q[nr] + q[5 * nr])) + ((q[2 * nr] + q[6 * nr]) + (q[3 * nr] + q[7 * nr]));
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int rr = 0; rr < GS8_RPW; ++rr) {
    const int r = (blockIdx.x * 4 + (threadIdx.x >> 6)) * GS8_RPW + rr;
    if (r < nr) {  // uniform per wave
      double v = 0.0;
#pragma unroll 8  // loads in flight; the FMAs keep their order
      for (int c = lane; c < nr; c += 64) v = fma(S[(int64_t)r * nr + c], w1s[c], v);
      v = wave_sum(v);
      if (lane == 0) y[r] = v;
    }
  }
}

// Q^T x for an FP16 Q with nr % 8 == 0: as qtx4_kernel with 8-column octets
// per 16-byte load (lane l owns octets l, l + 64, ...), FP64 accumulation.
template <int OPL>
__global__ __launch_bounds__(256) void qtx8h_kernel(const _Float16* __restrict__ Q, const float* __restrict__ x,
                                                    int64_t n, int nr, int64_t rows_per_block,
                                                    double* __restrict__ part) {
  __shared__ double red[4][OPL * 512];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int no = nr >> 3;
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r1 = (r0 + rows_per_block < n) ? r0 + rows_per_block : n;
  double acc[OPL][8];
#pragma unroll
  for (int u = 0; u < OPL; ++u)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[u][j] = 0.0;
  int64_t i = r0 + w;
  for (; i + 12 < r1; i += 16) {
    halfx8 q[4][OPL];
    float xv[4];
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      xv[rr] = x[i + 4 * rr];
      const uint4* row = reinterpret_cast<const uint4*>(Q + (i + 4 * rr) * nr);
#pragma unroll
      for (int u = 0; u < OPL; ++u) {
        const int oi = lane + 64 * u;
        q[rr][u] = __builtin_bit_cast(halfx8, (oi < no) ? row[oi] : uint4{0u, 0u, 0u, 0u});
      }
    }
#pragma unroll
    for (int rr = 0; rr < 4; ++rr)
#pragma unroll
      for (int u = 0; u < OPL; ++u)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[u][j] = fma((double)q[rr][u][j], (double)xv[rr], acc[u][j]);
  }
  for (; i < r1; i += 4) {
    const double xv = (double)x[i];
    const uint4* row = reinterpret_cast<const uint4*>(Q + i * nr);
#pragma unroll
    for (int u = 0; u < OPL; ++u) {
      const int oi = lane + 64 * u;
      if (oi < no) {
        const halfx8 q = __builtin_bit_cast(halfx8, row[oi]);
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[u][j] = fma((double)q[j], xv, acc[u][j]);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < OPL; ++u)
#pragma unroll
    for (int j = 0; j < 8; ++j) red[w][(lane + 64 * u) * 8 + j] = acc[u][j];
  __syncthreads();
  for (int c = threadIdx.x; c < nr; c += 256)
    part[(int64_t)blockIdx.x * nr + c] = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
}

// z[i] = f * x[i] + sum_c Q[i][c] w[c] for an FP16 Q with nr % 8 == 0 and
// nr <= 640: sixteen lanes per row (lane 16 r + o takes octets o + 16 t,
// t < 5, of row r), four rows per wave at a time, RPW rows per wave.  Each
// lane keeps the w entries of its five octets in registers (loaded once per
// wave), so the loop is loads and FMAs only; the row sum closes over the
// sixteen lanes with four shuffles.  Octets past the row end hold zeros.
template <int RPW>
__global__ __launch_bounds__(256) void qw8h_kernel(const _Float16* __restrict__ Q, const double* __restrict__ w,
                                                   const float* __restrict__ x, float* __restrict__ z, int64_t n,
                                                   int nr, double f) {
  constexpr int OT = 5;
  const int lane = threadIdx.x & 63, rs = lane >> 4, ol = lane & 15;
  const int no = nr >> 3;
  double wr[OT][8];
#pragma unroll
  for (int t = 0; t < OT; ++t) {
    const int oi = ol + 16 * t;
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) wr[t][jj] = (oi < no) ? w[8 * oi + jj] : 0.0;
  }
  const int64_t base = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * RPW;
#pragma unroll 2
  for (int rr = 0; rr < RPW / 4; ++rr) {
    const int64_t i = base + 4 * rr + rs;
    const uint4* row = reinterpret_cast<const uint4*>(Q + i * nr);
    halfx8 q[OT];
#pragma unroll
    for (int t = 0; t < OT; ++t) {
      const int oi = ol + 16 * t;
      q[t] = __builtin_bit_cast(halfx8, (i < n && oi < no) ? row[oi] : uint4{0u, 0u, 0u, 0u});
    }
    double s = 0.0;
#pragma unroll
    for (int t = 0; t < OT; ++t)
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) s = fma((double)q[t][jj], wr[t][jj], s);
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    s += __shfl_xor(s, 4, 64);
    s += __shfl_xor(s, 8, 64);
    if (ol == 0 && i < n) z[i] = (float)(f * (double)x[i] + s);
  }
}

// G = Qh^T Qh (nr x nr, FP64) for the FP16 basis, on MFMA.  Workgroup
// (tile pair ti <= tj of 64 columns, row split s) accumulates the 64 x 64 block
// over its rows: 32-row panels of both column groups are loaded with 16-byte
// loads, stored transposed in LDS (column-major, pitch GH_P halves) so a lane
// reads 8 consecutive rows of one column at once, and fed to
// v_mfma_f32_32x32x16_f16 (products of FP16 values are exact in FP32); the
// FP32 tile is added into FP64 every 256 rows.  part[s][i][j] (i, j < nr)
// holds the split's sums; gram_reduce_kernel adds the splits in order.
constexpr int GH_P = 40;  // LDS pitch (halves) of one transposed column: 32 rows + 8
__global__ __launch_bounds__(256) void gram_f16_kernel(const _Float16* __restrict__ Q, int64_t n, int nr,
                                                       int64_t rows_per_split, double* __restrict__ part) {
  __shared__ __align__(16) _Float16 ta[64 * GH_P];
  __shared__ __align__(16) _Float16 tb[64 * GH_P];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wi = w >> 1, wj = w & 1, h = lane >> 5;
  // blockIdx.x enumerates the tile pairs (ta, tb), ta <= tb, row by row
  const int nt = (nr + 63) / 64;
  int ta_i = 0, rem = (int)blockIdx.x;
  while (rem >= nt - ta_i) {
    rem -= nt - ta_i;
    ++ta_i;
  }
  const int c0 = 64 * ta_i, c1 = 64 * (ta_i + rem);
  const int64_t s = blockIdx.y;
  const int64_t r0 = s * rows_per_split;
  const int64_t r1 = (r0 + rows_per_split < n) ? r0 + rows_per_split : n;
  // panel load: thread t takes row t / 8 of the 32, columns 8 (t % 8) .. + 7
  const int prow = tid >> 3, poct = tid & 7;
  auto load = [&](int64_t rb, int c, uint4& v) {
    const int64_t r = rb + prow;
    const int col = c + 8 * poct;
    v = (r < r1 && col < nr) ? *reinterpret_cast<const uint4*>(Q + r * nr + col) : uint4{0u, 0u, 0u, 0u};
  };
  auto stash = [&](_Float16* t, const uint4& v) {
    const halfx8 hv = __builtin_bit_cast(halfx8, v);
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) t[(8 * poct + jj) * GH_P + prow] = hv[jj];
  };
  floatx16 acc = {0};
  double acc64[16];
#pragma unroll
  for (int v = 0; v < 16; ++v) acc64[v] = 0.0;
  uint4 pa, pb;
  if (r0 < r1) {
    load(r0, c0, pa);
    load(r0, c1, pb);
  }
  int steps = 0;
  for (int64_t rb = r0; rb < r1; rb += 32) {
    __syncthreads();  // the previous panel's MFMAs are done with the LDS
    stash(ta, pa);
    stash(tb, pb);
    __syncthreads();
    if (rb + 32 < r1) {
      load(rb + 32, c0, pa);
      load(rb + 32, c1, pb);
    }
#pragma unroll
    for (int ss = 0; ss < 2; ++ss) {
      const halfx8 av = *reinterpret_cast<const halfx8*>(ta + (32 * wi + (lane & 31)) * GH_P + 16 * ss + 8 * h);
      const halfx8 bv = *reinterpret_cast<const halfx8*>(tb + (32 * wj + (lane & 31)) * GH_P + 16 * ss + 8 * h);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv, acc, 0, 0, 0);
    }
    if (++steps == 8) {  // 256 rows in FP32, then into FP64
      steps = 0;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        acc64[v] += (double)acc[v];
        acc[v] = 0.f;
      }
    }
  }
#pragma unroll
  for (int v = 0; v < 16; ++v) acc64[v] += (double)acc[v];
  // D layout: lane l holds column n = l % 32 and rows m = 8 (v / 4) + 4 h + v % 4
  const int j = c1 + 32 * wj + (lane & 31);
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int i = c0 + 32 * wi + 8 * (v >> 2) + 4 * h + (v & 3);
    if (i < nr && j < nr) part[(s * nr + i) * nr + j] = acc64[v];
  }
}

// G[i][j] = sum over splits of part[s][min(i,j)][max(i,j)] (fixed order; the
// computed blocks are those with ti <= tj, which hold every (min, max) pair)
__global__ void gram_reduce_kernel(const double* __restrict__ part, int nsplit, int nr, double* __restrict__ G) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)nr * nr) return;
  const int i = (int)(idx / nr), j = (int)(idx % nr);
  const int a = i < j ? i : j, b = i < j ? j : i;
  double v = 0.0;
  for (int s = 0; s < nsplit; ++s) v += part[((int64_t)s * nr + a) * nr + b];
  G[idx] = v;
}

__global__ void to_half_kernel(const float* __restrict__ a, _Float16* __restrict__ out, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = (_Float16)a[i];
}

// z[i] = f * x[i] + sum_c Q[i][c] w[c]   (one wave per QW_ROWS rows)
constexpr int QW_ROWS = 2;
template <class T>
__global__ __launch_bounds__(256) void qw_vec_kernel(const T* __restrict__ Q, const double* __restrict__ w,
                                                     const T* __restrict__ x, T* __restrict__ z,
                                                     int64_t n, int nr, double f) {
  extern __shared__ double ws_lds[];
  for (int c = threadIdx.x; c < nr; c += blockDim.x) ws_lds[c] = w[c];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t i0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * QW_ROWS;
  if (i0 >= n) return;
  double s[QW_ROWS];
#pragma unroll
  for (int b = 0; b < QW_ROWS; ++b) s[b] = 0.0;
  for (int c = lane; c < nr; c += 64) {
    const double wc = ws_lds[c];
#pragma unroll
    for (int b = 0; b < QW_ROWS; ++b)
      if (i0 + b < n) s[b] = fma((double)Q[(i0 + b) * nr + c], wc, s[b]);
  }
#pragma unroll
  for (int b = 0; b < QW_ROWS; ++b) {
    const double t = wave_sum(s[b]);
    if (lane == 0 && i0 + b < n) z[i0 + b] = (T)(f * (double)x[i0 + b] + t);
  }
}

template <class T>
__global__ void scale_copy_kernel(const T* __restrict__ X, int64_t ldx, T* __restrict__ Z, int64_t ldz,
                                  int64_t n, int k, double f) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * k) return;
  const int64_t i = idx / k;
  const int c = (int)(idx % k);
  Z[i * ldz + c] = (T)(f * (double)X[i * ldx + c]);
}

// ---------------------------------------------------------------------------
template <class T>
static int apply_t(Precond& pc, const T* X, int k, int64_t ldx, T* Z, int64_t ldz, hipStream_t st) {
  const int64_t n = pc.n;
  const int nr = pc.nr;
  LAKER_TRY(pc.w1.ensure((size_t)nr * k * 8));
  LAKER_TRY(pc.w2.ensure((size_t)nr * k * 8));
  double* W1 = pc.w1.as<double>();
  double* W2 = pc.w2.as<double>();
  const T* Q = pc.Q.as<T>();
  if constexpr (sizeof(T) == 4) {
    if (k == 1 && ldx == 1 && ldz == 1 && nr % 4 == 0 && nr <= 1024) {
      const int64_t nb = std::min<int64_t>(1024, ceil_div(n, 64));
      const int64_t rpb = ceil_div(n, nb);
      const int nb_eff = (int)ceil_div(n, rpb);
      constexpr int RG = 8;  // first-level reduction groups
      LAKER_TRY(pc.part.ensure(((size_t)nb_eff + RG) * nr * 8));
      double* part = pc.part.as<double>();
      double* part2 = part + (size_t)nb_eff * nr;
      const int qpl = (nr / 4 + 63) / 64;
      const dim3 qg((unsigned)nb_eff), qb(256);
      if (qpl == 1) hipLaunchKernelGGL((qtx4_kernel<1>), qg, qb, 0, st, Q, X, n, nr, rpb, part);
      else if (qpl == 2) hipLaunchKernelGGL((qtx4_kernel<2>), qg, qb, 0, st, Q, X, n, nr, rpb, part);
      else if (qpl == 3) hipLaunchKernelGGL((qtx4_kernel<3>), qg, qb, 0, st, Q, X, n, nr, rpb, part);
      else hipLaunchKernelGGL((qtx4_kernel<4>), qg, qb, 0, st, Q, X, n, nr, rpb, part);
      LAKER_KCHECK("qtx4_kernel");
      const unsigned cb = (unsigned)ceil_div(nr, 64);
      hipLaunchKernelGGL(reduce_parts_groups_kernel, dim3(cb, RG), dim3(256), 0, st, (const double*)part,
                         nb_eff, nr, part2);
      LAKER_KCHECK("reduce_parts_groups_kernel");
      static_assert(RG == 8, "gemv64_sum8_kernel sums eight group partials");
      hipLaunchKernelGGL(gemv64_sum8_kernel, dim3((unsigned)ceil_div(nr, 4 * GS8_RPW)), dim3(256), nr * sizeof(double), st,
                         pc.S.as<double>(), (const double*)part2, W2, nr);
      LAKER_KCHECK("gemv64_sum8_kernel");
      hipLaunchKernelGGL((qw4_kernel<2>), dim3((unsigned)ceil_div(n, 8)), dim3(256), nr * sizeof(double), st,
                         Q, (const double*)W2, X, Z, n, nr, pc.inv_sqrt_a);
      LAKER_KCHECK("qw4_kernel");
      return LAKER_OK;
    }
  }
  if (k == 1 && ldx == 1 && ldz == 1) {
    const int64_t nb = std::min<int64_t>(512, ceil_div(n, 16));
    const int64_t rpb = ceil_div(n, nb);
    const int64_t nb_eff = ceil_div(n, rpb);
    LAKER_TRY(pc.part.ensure((size_t)nb_eff * nr * 8));
    hipLaunchKernelGGL((qtx_vec_kernel<T>), dim3((unsigned)nb_eff), dim3(256), 0, st, Q, X, n, nr, rpb,
                       pc.part.as<double>());
    LAKER_KCHECK("qtx_vec_kernel");
    hipLaunchKernelGGL(reduce_parts_cols_kernel, dim3((unsigned)ceil_div(nr, 64)), dim3(256), 0, st,
                       pc.part.as<double>(), (int)nb_eff, nr, W1);
    LAKER_KCHECK("reduce_parts_cols_kernel");
    hipLaunchKernelGGL(gemv64_kernel, dim3((unsigned)ceil_div(nr, 4)), dim3(256), 0, st, pc.S.as<double>(), W1,
                       W2, nr);
    LAKER_KCHECK("gemv64_kernel");
    hipLaunchKernelGGL((qw_vec_kernel<T>), dim3((unsigned)ceil_div(n, 4 * QW_ROWS)), dim3(256),
                       nr * sizeof(double), st, Q, W2, X, Z, n, nr, pc.inv_sqrt_a);
    LAKER_KCHECK("qw_vec_kernel");
    LAKER_HIP_CHECK(hipGetLastError());
    return LAKER_OK;
  }
  if constexpr (sizeof(T) == 4) {
    if (k >= 2 && k <= 32 && nr % 4 == 0 && (size_t)16 * qw_stride(nr) * sizeof(float) <= 64 * 1024) {
      // W1 = Q^T X on MFMA with split partials, two-level fixed-order reduction
      const int cblk = (int)ceil_div(nr, 32);
      const int64_t nsplit = std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), ceil_div(1024, cblk)));
      const int64_t rps = ceil_div(n, nsplit);
      const int64_t ns_eff = ceil_div(n, rps);
      constexpr int RG = 8;
      const int64_t width = (int64_t)nr * k;
      LAKER_TRY(pc.part.ensure(((size_t)ns_eff + RG) * width * 8));
      double* part = pc.part.as<double>();
      double* part2 = part + (size_t)ns_eff * width;
      hipLaunchKernelGGL(qtx_mfma_kernel, dim3((unsigned)cblk, (unsigned)ns_eff), dim3(256), 0, st, Q, n, nr, X,
                         ldx, k, rps, part);
      LAKER_KCHECK("qtx_mfma_kernel");
      const unsigned wb = (unsigned)ceil_div(width, 64);
      hipLaunchKernelGGL(reduce_parts_groups_kernel, dim3(wb, RG), dim3(256), 0, st, (const double*)part,
                         (int)ns_eff, (int)width, part2);
      LAKER_KCHECK("reduce_parts_groups_kernel");
      hipLaunchKernelGGL(reduce_parts_groups_kernel, dim3(wb, 1), dim3(256), 0, st, (const double*)part2, RG,
                         (int)width, W1);
      LAKER_KCHECK("reduce_parts_groups_kernel");
      LAKER_TRY((gemm<double, double, double>(false, false, nr, k, nr, 1.0, pc.S.as<double>(), nr, W1, k,
                                               0.0, W2, k, nullptr, st)));
      const int P = qw_stride(nr);
      const unsigned qb = (unsigned)ceil_div(n, QW_BLOCK_ROWS);
      if (k > 16 && (size_t)32 * P * sizeof(float) <= 64 * 1024) {
        hipLaunchKernelGGL((qw_mfma_kernel<2>), dim3(qb), dim3(256), (size_t)32 * P * sizeof(float), st, Q,
                           (const double*)W2, n, nr, P, k, 0, X, ldx, Z, ldz, pc.inv_sqrt_a);
        LAKER_KCHECK("qw_mfma_kernel");
      } else {
        for (int j0 = 0; j0 < k; j0 += 16) {
          hipLaunchKernelGGL((qw_mfma_kernel<1>), dim3(qb), dim3(256), (size_t)16 * P * sizeof(float), st, Q,
                             (const double*)W2, n, nr, P, k, j0, X, ldx, Z, ldz, pc.inv_sqrt_a);
          LAKER_KCHECK("qw_mfma_kernel");
        }
      }
      return LAKER_OK;
    }
  }
  // general k: W1 = Q^T X ; W2 = S W1 ; Z = a^{-1/2} X + Q W2
  LAKER_TRY((gemm<T, T, double>(true, false, nr, k, n, 1.0, Q, nr, X, ldx, 0.0, W1, k, &pc.ws, st)));
  LAKER_TRY((gemm<double, double, double>(false, false, nr, k, nr, 1.0, pc.S.as<double>(), nr, W1, k,
                                           0.0, W2, k, nullptr, st)));
  hipLaunchKernelGGL((scale_copy_kernel<T>), dim3((unsigned)ceil_div(n * k, 256)), dim3(256), 0, st, X,
                     ldx, Z, ldz, n, k, pc.inv_sqrt_a);
  LAKER_KCHECK("scale_copy_kernel");
  LAKER_HIP_CHECK(hipGetLastError());
  LAKER_TRY((gemm<T, double, T>(false, false, n, k, nr, 1.0, Q, nr, W2, k, 1.0, Z, ldz, nullptr, st)));
  return LAKER_OK;
}

int precond_apply(Precond& pc, const void* X, int k, int64_t ldx, void* Z, int64_t ldz, hipStream_t st) {
  if (pc.kind == PRECOND_JACOBI) {
    const dim3 g((unsigned)ceil_div(pc.n * k, 256)), b(256);
    if (pc.dtype == LAKER_F32)
      hipLaunchKernelGGL((jacobi_apply_kernel<float>), g, b, 0, st, pc.inv_diag.as<float>(), (const float*)X, ldx,
                         (float*)Z, ldz, pc.n, k);
    else
      hipLaunchKernelGGL((jacobi_apply_kernel<double>), g, b, 0, st, pc.inv_diag.as<double>(), (const double*)X,
                         ldx, (double*)Z, ldz, pc.n, k);
    LAKER_KCHECK("jacobi_apply_kernel");
    return LAKER_OK;
  }
  if (pc.dtype == LAKER_F32) return apply_t<float>(pc, (const float*)X, k, ldx, (float*)Z, ldz, st);
  return apply_t<double>(pc, (const double*)X, k, ldx, (double*)Z, ldz, st);
}

static int read_scalar(const void* dev, void* host, size_t bytes, hipStream_t st) {
  LAKER_HIP_CHECK(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, st));
  LAKER_HIP_CHECK(hipStreamSynchronize(st));
  return LAKER_OK;
}

// FP16 copy of the CCCP basis for the k = 1 applies of the mixed solve (half
// the bytes of the two passes over Q).  Rounding Q alone could leave
// a^{-1/2}(I - Qh Qh^T) slightly indefinite, so the preconditioner is formed
// on the orthonormal basis Q0 = Qh L^{-T} of span(Qh) instead (L L^T =
// Qh^T Qh): P' = a^{-1/2}(I - Q0 Q0^T) + Q0 M^{-1/2} Q0^T = a^{-1/2} I +
// Qh S' Qh^T with S' = L^{-T} S L^{-1}, symmetric positive definite like P and
// within O(2^-11) of it.  (The Gram sums FP16 products exactly in FP32 over
// 256-row blocks and in FP64 beyond: Q0 is orthonormal to ~1e-6, far inside
// the margin a^{-1/2}(I - Q0 Q0^T) needs against M^{-1/2}.)  Returns 1 when the handle does not qualify (not an
// FP32 CCCP handle, N_r not a multiple of 8 or above 640, Qh^T Qh not
// positive definite), LAKER_OK once Qh and S' exist.
int precond_prepare_half(Precond& pc, hipStream_t st) {
  if (pc.has_h) return LAKER_OK;
  if (pc.kind != PRECOND_CCCP || pc.dtype != LAKER_F32 || pc.nr <= 0 || pc.nr % 8 != 0 || pc.nr > 640) return 1;
  const int64_t n = pc.n;
  const int nr = pc.nr;
  const size_t nn = (size_t)nr * nr;
  LAKER_TRY(pc.Qh.ensure((size_t)n * nr * sizeof(_Float16)));
  hipLaunchKernelGGL(to_half_kernel, dim3((unsigned)ceil_div(n * nr, 256)), dim3(256), 0, st, pc.Q.as<float>(),
                     pc.Qh.as<_Float16>(), n * nr);
  LAKER_KCHECK("to_half_kernel");
  DevBuf G, Li, T, info;
  TriWork tw;
  tw.bind(st);
  LAKER_TRY(G.ensure(nn * 8, st));
  LAKER_TRY(Li.ensure(nn * 8, st));
  LAKER_TRY(T.ensure(nn * 8, st));
  LAKER_TRY(info.ensure(sizeof(int), st));
  LAKER_TRY(pc.Sh.ensure(nn * 8));
  // Qh^T Qh on FP16 MFMA (exact products, FP32 over 256 rows, FP64 beyond)
  {
    const int nt = (int)ceil_div(nr, 64);
    const int npairs = nt * (nt + 1) / 2;
    const int64_t nsplit = std::max<int64_t>(1, std::min<int64_t>(64, ceil_div(n, 2048)));
    const int64_t rps = round_up(ceil_div(n, nsplit), 32);
    const int64_t ns_eff = ceil_div(n, rps);
    DevBuf gpart;
    LAKER_TRY(gpart.ensure((size_t)ns_eff * nn * 8, st));
    hipLaunchKernelGGL(gram_f16_kernel, dim3((unsigned)npairs, (unsigned)ns_eff), dim3(256), 0, st,
                       pc.Qh.as<_Float16>(), n, nr, rps, gpart.as<double>());
    LAKER_KCHECK("gram_f16_kernel");
    hipLaunchKernelGGL(gram_reduce_kernel, dim3((unsigned)ceil_div((int64_t)nn, 256)), dim3(256), 0, st,
                       (const double*)gpart.as<double>(), (int)ns_eff, nr, G.as<double>());
    LAKER_KCHECK("gram_reduce_kernel");
  }
  LAKER_HIP_CHECK(hipMemsetAsync(info.ptr, 0, sizeof(int), st));
  LAKER_TRY(potrf_lower(G.as<double>(), nr, nr, info.as<int>(), &pc.ws, tw, st));
  LAKER_TRY(set_identity(Li.as<double>(), nr, 1.0, st));
  LAKER_TRY(trsm_lower_left(G.as<double>(), nr, nr, Li.as<double>(), nr, nr, &pc.ws, tw, st));  // L^{-1}
  LAKER_TRY((gemm<double, double, double>(false, false, nr, nr, nr, 1.0, pc.S.as<double>(), nr, Li.as<double>(),
                                           nr, 0.0, T.as<double>(), nr, &pc.ws, st)));  // S L^{-1}
  LAKER_TRY((gemm<double, double, double>(true, false, nr, nr, nr, 1.0, Li.as<double>(), nr, T.as<double>(), nr,
                                           0.0, pc.Sh.as<double>(), nr, &pc.ws, st)));  // L^{-T} S L^{-1}
  int hinfo = 0;
  LAKER_HIP_CHECK(hipMemcpyAsync(&hinfo, info.ptr, sizeof(int), hipMemcpyDeviceToHost, st));
  LAKER_HIP_CHECK(hipStreamSynchronize(st));  // before the temporaries are freed
  if (hinfo != 0) return 1;
  pc.has_h = true;
  return LAKER_OK;
}

// z = P' x for one FP32 vector with the FP16 basis (see precond_prepare_half):
// the k = 1 path of apply_t with Qh and S'.
int precond_apply_half(Precond& pc, const float* X, float* Z, hipStream_t st) {
  if (!pc.has_h) return fail(LAKER_ESTATE, "precond_apply_half: no FP16 basis");
  const int64_t n = pc.n;
  const int nr = pc.nr;
  LAKER_TRY(pc.w2.ensure((size_t)nr * 8));
  double* W2 = pc.w2.as<double>();
  const int64_t nb = std::min<int64_t>(1024, ceil_div(n, 64));
  const int64_t rpb = ceil_div(n, nb);
  const int nb_eff = (int)ceil_div(n, rpb);
  constexpr int RG = 8;  // first-level reduction groups
  LAKER_TRY(pc.part.ensure(((size_t)nb_eff + RG) * nr * 8));
  double* part = pc.part.as<double>();
  double* part2 = part + (size_t)nb_eff * nr;
  const _Float16* qh = pc.Qh.as<_Float16>();
  const dim3 qg((unsigned)nb_eff), qb(256);
  if ((nr / 8 + 63) / 64 == 1) hipLaunchKernelGGL((qtx8h_kernel<1>), qg, qb, 0, st, qh, X, n, nr, rpb, part);
  else hipLaunchKernelGGL((qtx8h_kernel<2>), qg, qb, 0, st, qh, X, n, nr, rpb, part);
  LAKER_KCHECK("qtx8h_kernel");
  hipLaunchKernelGGL(reduce_parts_groups_kernel, dim3((unsigned)ceil_div(nr, 64), RG), dim3(256), 0, st,
                     (const double*)part, nb_eff, nr, part2);
  LAKER_KCHECK("reduce_parts_groups_kernel");
  static_assert(RG == 8, "gemv64_sum8_kernel sums eight group partials");
  hipLaunchKernelGGL(gemv64_sum8_kernel, dim3((unsigned)ceil_div(nr, 4 * GS8_RPW)), dim3(256), nr * sizeof(double), st,
                     pc.Sh.as<double>(), (const double*)part2, W2, nr);
  LAKER_KCHECK("gemv64_sum8_kernel");
  // 128 rows per block (4 waves x 32 rows)
  hipLaunchKernelGGL((qw8h_kernel<32>), dim3((unsigned)ceil_div(n, 128)), dim3(256), 0, st, qh, (const double*)W2,
                     X, Z, n, nr, pc.inv_sqrt_a);
  LAKER_KCHECK("qw8h_kernel");
  return LAKER_OK;
}

// One Cholesky-QR pass in FP64: G = Vin^T Vin (+ shift I), L = chol(G),
// Vout = Vin L^{-T}, Racc <- L^T Racc.  A failed factorisation leaves a
// non-zero *info (device); the caller reads it.
static int cholqr_pass(const double* Vin, double* Vout, int64_t n, int nr, double shift, double* G, double* Linv,
                       double* Racc, double* tmp, int* info, DevBuf* ws, TriWork& tw, hipStream_t st) {
  // lower triangle only: the Cholesky reads no other part of G
  LAKER_TRY((gemm<double, double, double>(true, false, nr, nr, n, 1.0, Vin, nr, Vin, nr, 0.0, G, nr, ws, st,
                                           GEMM_LOWER)));
  if (shift > 0.0) {
    hipLaunchKernelGGL(add_diag_kernel, dim3((unsigned)ceil_div(nr, 256)), dim3(256), 0, st, G, (int64_t)nr,
                       shift);
    LAKER_KCHECK("add_diag_kernel");
  }
  LAKER_HIP_CHECK(hipMemsetAsync(info, 0, sizeof(int), st));
  LAKER_TRY(potrf_lower(G, nr, nr, info, ws, tw, st));
  // Linv = L^{-1}
  LAKER_TRY(set_identity(Linv, nr, 1.0, st));
  LAKER_TRY(trsm_lower_left(G, nr, nr, Linv, nr, nr, ws, tw, st));
  // Vout = Vin Linv^T  (Linv^T upper triangular: column j needs k <= j)
  LAKER_TRY((gemm<double, double, double>(false, true, n, nr, nr, 1.0, Vin, nr, Linv, nr, 0.0, Vout, nr,
                                           nullptr, st, GEMM_BTRI)));
  // Racc <- L^T Racc   (L^T upper = R factor of this pass)
  LAKER_TRY((gemm<double, double, double>(true, false, nr, nr, nr, 1.0, G, nr, Racc, nr, 0.0, tmp, nr,
                                           nullptr, st)));
  LAKER_HIP_CHECK(hipMemcpyAsync(Racc, tmp, (size_t)nr * nr * 8, hipMemcpyDeviceToDevice, st));
  return LAKER_OK;
}

// Orthonormal basis of the n x nr block src (dtype LAKER_F32/F64) in FP64:
// Cholesky-QR2 (two passes), or a shifted Cholesky-QR3 when the first Gram
// matrix is not numerically positive definite.  V receives the basis, Racc the
// accumulated upper-triangular factor (src = V Racc).  LAKER_ENUMERIC if every
// route fails (rank-deficient block).  The two passes of Cholesky-QR2 run
// without a host round trip; their factorisation flags are read once at the
// end (a failed first pass makes the second one's output meaningless, and the
// shifted route then starts over from src).
static int cholqr_block(const void* src, int dtype, int64_t n, int nr, double* V, double* Racc, DevBuf* ws,
                        hipStream_t st, bool v_loaded = false) {
  const int64_t nn = (int64_t)nr * nr;
  DevBuf Vtmp, G, Linv, tmp, infob;
  TriWork tri;
  tri.bind(st);
  LAKER_TRY(Vtmp.ensure((size_t)n * nr * 8, st));
  LAKER_TRY(G.ensure(nn * 8, st));
  LAKER_TRY(Linv.ensure(nn * 8, st));
  LAKER_TRY(tmp.ensure(nn * 8, st));
  LAKER_TRY(infob.ensure(3 * sizeof(int), st));
  int* info = infob.as<int>();
  double* Vt = Vtmp.as<double>();
  auto load_v = [&]() -> int {
    if (v_loaded) {  // V already holds the FP64 copy of src (first attempt only)
      v_loaded = false;
      return set_identity(Racc, nr, 1.0, st);
    }
    if (dtype == LAKER_F32)
      hipLaunchKernelGGL((convert_kernel<float, double>), dim3((unsigned)ceil_div(n * nr, 256)), dim3(256), 0,
                         st, (const float*)src, V, n * nr);
    else
      LAKER_HIP_CHECK(hipMemcpyAsync(V, src, (size_t)n * nr * 8, hipMemcpyDeviceToDevice, st));
    LAKER_HIP_CHECK(hipGetLastError());
    return set_identity(Racc, nr, 1.0, st);
  };
  auto pass = [&](const double* vin, double* vout, double shift, int q) {
    return cholqr_pass(vin, vout, n, nr, shift, G.as<double>(), Linv.as<double>(), Racc, tmp.as<double>(),
                       info + q, ws, tri, st);
  };
  auto read_info = [&](int* h, int count) -> int {
    LAKER_HIP_CHECK(hipMemcpyAsync(h, info, count * sizeof(int), hipMemcpyDeviceToHost, st));
    LAKER_HIP_CHECK(hipStreamSynchronize(st));
    return LAKER_OK;
  };
  LAKER_TRY(load_v());
  LAKER_TRY(pass(V, Vt, 0.0, 0));
  LAKER_TRY(pass(Vt, V, 0.0, 1));
  int hinfo[3] = {0, 0, 0};
  LAKER_TRY(read_info(hinfo, 2));
  if (hinfo[0] == 0) return hinfo[1] == 0 ? LAKER_OK : LAKER_ENUMERIC;
  // shifted Cholesky-QR: shift = 11 (n nr + nr (nr + 1)) u ||U||_F^2, ||U||_F^2 <= nr
  // (columns of unit norm; callers normalise first), then two plain passes
  LAKER_TRY(load_v());
  const double u = 1.1102230246251565e-16;
  const double shift = 11.0 * ((double)n * nr + (double)nr * (nr + 1)) * u * (double)nr;
  LAKER_TRY(pass(V, Vt, shift, 0));
  LAKER_TRY(pass(Vt, V, 0.0, 1));
  LAKER_TRY(pass(V, Vt, 0.0, 2));
  LAKER_HIP_CHECK(hipMemcpyAsync(V, Vt, (size_t)n * nr * 8, hipMemcpyDeviceToDevice, st));
  LAKER_TRY(read_info(hinfo, 3));
  return (hinfo[0] == 0 && hinfo[1] == 0 && hinfo[2] == 0) ? LAKER_OK : LAKER_ENUMERIC;
}

// Steps 2-5 of the build on probe responses U (n x nr, op dtype; consumed).
static int build_from_u(Operator& op, std::unique_ptr<Precond>& pc, DevBuf& U, int nr, double gamma,
                        double epsilon, double base_rho, int max_iter, double tol, hipStream_t st,
                        int* iterations) {
  const int64_t n = pc->n;
  const size_t esz = pc->dtype == LAKER_F32 ? 4 : 8;
  const int64_t nn = (int64_t)nr * nr;
  (void)op;
  // 2. column normalisation in the solver dtype; the scaling pass also writes
  // the FP64 copy the Cholesky-QR starts from
  DevBuf ss, chains, V, Racc;
  LAKER_TRY(ss.ensure((size_t)nr * 8, st));
  LAKER_TRY(V.ensure((size_t)n * nr * 8, st));
  LAKER_TRY(Racc.ensure(nn * 8, st));
  const dim3 sg((unsigned)ceil_div(nr, 256), (unsigned)ceil_div(n, 64)), sb(256);
  if (pc->dtype == LAKER_F32) {
    LAKER_TRY(colsumsq(U.as<float>(), n, nr, ss.as<double>(), chains, st));
    hipLaunchKernelGGL((colscale_kernel<float>), sg, sb, 0, st, U.as<float>(), n, nr, (const double*)ss.as<double>(),
                       epsilon, V.as<double>());
    LAKER_KCHECK("colscale_kernel");
  } else {
    LAKER_TRY(colsumsq(U.as<double>(), n, nr, ss.as<double>(), chains, st));
    hipLaunchKernelGGL((colscale_kernel<double>), sg, sb, 0, st, U.as<double>(), n, nr,
                       (const double*)ss.as<double>(), epsilon, V.as<double>());
    LAKER_KCHECK("colscale_kernel");
  }
  LAKER_HIP_CHECK(hipGetLastError());

  // 3. Cholesky-QR2 in FP64 (shifted Cholesky-QR3 on failure)
  {
    const int rc = cholqr_block(U.ptr, pc->dtype, n, nr, V.as<double>(), Racc.as<double>(), &pc->ws, st, true);
    if (rc == LAKER_ENUMERIC) return fail(LAKER_ENUMERIC, "CCCP build: Cholesky-QR of the probe responses failed");
    LAKER_TRY(rc);
  }
  U.release();
  LAKER_TRY(pc->Q.ensure((size_t)n * nr * esz));
  if (pc->dtype == LAKER_F32)
    hipLaunchKernelGGL((convert_kernel<double, float>), dim3((unsigned)ceil_div(n * nr, 256)), dim3(256), 0,
                       st, V.as<double>(), pc->Q.as<float>(), n * nr);
  else
    LAKER_HIP_CHECK(hipMemcpyAsync(pc->Q.ptr, V.ptr, (size_t)n * nr * 8, hipMemcpyDeviceToDevice, st));
  LAKER_HIP_CHECK(hipGetLastError());
  V.release();
  LAKER_TRY(pc->R.ensure(nn * 8));
  LAKER_HIP_CHECK(hipMemcpyAsync(pc->R.ptr, Racc.ptr, nn * 8, hipMemcpyDeviceToDevice, st));
  double* Rq = pc->R.as<double>();

  // 4. CCCP fixed point (FP64)
  double rho;
  if (nr >= n) {
    rho = base_rho;
  } else {
    const double ratio = (double)nr / (double)n;
    rho = base_rho + (1.0 - base_rho) * (1.0 - ratio) * std::min(1.0, gamma * 10.0);
    rho = std::min(rho, 0.5);
  }
  pc->rho = rho;
  const double s = 1.0 + gamma / (double)n;
  const double a_rho = (1.0 - rho) * gamma / s + rho;
  DevBuf Cb, Cn, Mb, Zb, Rw, Hb, scb, partb, infob, Linv;
  TriWork tri;
  tri.bind(st);
  LAKER_TRY(infob.ensure(sizeof(int), st));
  LAKER_TRY(Linv.ensure(nn * 8, st));
  LAKER_TRY(Cb.ensure(nn * 8, st));
  LAKER_TRY(Cn.ensure(nn * 8, st));
  LAKER_TRY(Mb.ensure(nn * 8, st));
  LAKER_TRY(Zb.ensure(nn * 8, st));
  LAKER_TRY(Rw.ensure(nn * 8, st));
  LAKER_TRY(Hb.ensure(nn * 8, st));
  LAKER_TRY(scb.ensure(sizeof(CccpScalars), st));
  const int upd_blocks = (int)std::min<int64_t>(256, ceil_div(nn, 256));
  LAKER_TRY(partb.ensure((size_t)upd_blocks * 2 * 8, st));
  LAKER_HIP_CHECK(hipMemsetAsync(Cb.ptr, 0, nn * 8, st));
  double a = 1.0;
  int it_done = 0;
  const double n_over_nr = (double)n / (double)nr;
  for (int it = 0; it < max_iter; ++it) {
    LAKER_HIP_CHECK(hipMemsetAsync(infob.ptr, 0, sizeof(int), st));
    LAKER_HIP_CHECK(hipMemcpyAsync(Zb.ptr, Rq, nn * 8, hipMemcpyDeviceToDevice, st));
    if (it > 0) {
      // Z = L^{-1} R with L L^T = M = aI + C.  The first iteration starts from
      // C = 0, a = 1: M = I, whose factor and solve return R unchanged, bit for bit
      hipLaunchKernelGGL(form_m_kernel, dim3((unsigned)ceil_div(nn, 256)), dim3(256), 0, st, Cb.as<double>(),
                         Mb.as<double>(), (int64_t)nr, a);
      LAKER_KCHECK("form_m_kernel");
      LAKER_TRY(potrf_lower(Mb.as<double>(), nr, nr, infob.as<int>(), &pc->ws, tri, st));
      LAKER_TRY(trsm_lower_left(Mb.as<double>(), nr, nr, Zb.as<double>(), nr, nr, &pc->ws, tri, st));
    }
    hipLaunchKernelGGL(cccp_weights_kernel, dim3(nr), dim3(256), 0, st, Zb.as<double>(), Rq, Rw.as<double>(),
                       nr, n_over_nr, epsilon);
    LAKER_KCHECK("cccp_weights_kernel");
    LAKER_TRY((gemm<double, double, double>(false, true, nr, nr, nr, 1.0, Rw.as<double>(), nr, Rq, nr, 0.0,
                                             Hb.as<double>(), nr, &pc->ws, st)));
    hipLaunchKernelGGL(cccp_trace_kernel, dim3(1), dim3(256), 0, st, Hb.as<double>(), nr, gamma, s, rho,
                       a_rho, n, scb.as<CccpScalars>());
    LAKER_KCHECK("cccp_trace_kernel");
    hipLaunchKernelGGL(cccp_update_kernel, dim3(upd_blocks), dim3(256), 0, st, Hb.as<double>(),
                       Cb.as<double>(), Cn.as<double>(), nr, gamma, s, rho, scb.as<CccpScalars>(),
                       partb.as<double>());
    LAKER_KCHECK("cccp_update_kernel");
    hipLaunchKernelGGL(cccp_norms_final_kernel, dim3(1), dim3(64), 0, st, partb.as<double>(), upd_blocks,
                       scb.as<CccpScalars>());
    LAKER_KCHECK("cccp_norms_final_kernel");
    LAKER_HIP_CHECK(hipGetLastError());
    CccpScalars hs;
    int hinfo = 0;
    LAKER_HIP_CHECK(hipMemcpyAsync(&hinfo, infob.ptr, sizeof(int), hipMemcpyDeviceToHost, st));
    LAKER_TRY(read_scalar(scb.ptr, &hs, sizeof(hs), st));
    if (hinfo != 0) return fail(LAKER_ENUMERIC, "CCCP build: M = aI + C is not positive definite");
    const double a_prev = a;
    a = hs.a_new;
    Cb.swap(Cn);
    const double iso_rel = std::fabs(a - a_prev) / (std::fabs(a_prev) + 1e-12);
    const double q_rel = std::sqrt(hs.diff2) / (std::sqrt(hs.c2) + 1e-12);
    it_done = it + 1;
    if (std::max(iso_rel, q_rel) < tol) break;
  }
  pc->a = a;
  pc->inv_sqrt_a = 1.0 / std::sqrt(a);
  LAKER_TRY(pc->C.ensure(nn * 8));
  LAKER_HIP_CHECK(hipMemcpyAsync(pc->C.ptr, Cb.ptr, nn * 8, hipMemcpyDeviceToDevice, st));

  // 5. S = M^{-1/2} - a^{-1/2} I by coupled Newton-Schulz on M / c, c = ||M||_F.
  // The stopping test runs on the device (ns_check_kernel) and the host reads
  // its verdict every NS_POLL steps; steps after the stop multiply by I.
  hipLaunchKernelGGL(form_m_kernel, dim3((unsigned)ceil_div(nn, 256)), dim3(256), 0, st, Cb.as<double>(),
                     Mb.as<double>(), (int64_t)nr, a);
  LAKER_KCHECK("form_m_kernel");
  double* fro2 = scb.as<double>();
  hipLaunchKernelGGL(frob2_kernel, dim3(1), dim3(1024), 0, st, Mb.as<double>(), nn, fro2);
  LAKER_KCHECK("frob2_kernel");
  // Y = M / c  (in Mb), Z = I (in Zb)
  hipLaunchKernelGGL(scale_by_frob_kernel, dim3((unsigned)ceil_div(nn, 256)), dim3(256), 0, st, Mb.as<double>(),
                     (int64_t)nr, (const double*)fro2);
  LAKER_KCHECK("scale_by_frob_kernel");
  LAKER_TRY(set_identity(Zb.as<double>(), nr, 1.0, st));
  const int ns_blocks = (int)std::min<int64_t>(256, ceil_div(nn, 256));
  DevBuf nsb;
  LAKER_TRY(nsb.ensure(sizeof(NsState), st));
  NsState* nss = nsb.as<NsState>();
  LAKER_TRY(partb.ensure((size_t)std::max(ns_blocks, upd_blocks) * 2 * 8, st));
  hipLaunchKernelGGL(ns_init_kernel, dim3(1), dim3(1), 0, st, nss);
  LAKER_KCHECK("ns_init_kernel");
  constexpr int NS_POLL = 8;
  const double ns_tol = 1e-13 * std::sqrt((double)nr);
  NsState hns{};
  for (int it = 0; it < 200; ++it) {
    // P = Z Y -> Hb ; T -> Rw ; Y <- Y T (Cn) ; Z <- T Z (Linv)
    LAKER_TRY((gemm<double, double, double>(false, false, nr, nr, nr, 1.0, Zb.as<double>(), nr,
                                             Mb.as<double>(), nr, 0.0, Hb.as<double>(), nr, &pc->ws, st)));
    hipLaunchKernelGGL(ns_t_kernel, dim3(ns_blocks), dim3(256), 0, st, Hb.as<double>(), Rw.as<double>(),
                       (int64_t)nr, partb.as<double>());
    LAKER_KCHECK("ns_t_kernel");
    hipLaunchKernelGGL(ns_check_kernel, dim3(1), dim3(64), 0, st, (const double*)partb.as<double>(), ns_blocks,
                       ns_tol, nss);
    LAKER_KCHECK("ns_check_kernel");
    hipLaunchKernelGGL(ns_hold_kernel, dim3((unsigned)ceil_div(nn, 256)), dim3(256), 0, st, Rw.as<double>(),
                       (int64_t)nr, (const NsState*)nss);
    LAKER_KCHECK("ns_hold_kernel");
    LAKER_TRY((gemm<double, double, double>(false, false, nr, nr, nr, 1.0, Mb.as<double>(), nr,
                                             Rw.as<double>(), nr, 0.0, Cn.as<double>(), nr, &pc->ws, st)));
    LAKER_TRY((gemm<double, double, double>(false, false, nr, nr, nr, 1.0, Rw.as<double>(), nr,
                                             Zb.as<double>(), nr, 0.0, Linv.as<double>(), nr, &pc->ws, st)));
    Mb.swap(Cn);
    Zb.swap(Linv);
    if ((it + 1) % NS_POLL == 0 || it == 199) {
      LAKER_TRY(read_scalar(nss, &hns, sizeof(hns), st));
      if (hns.status != 0) break;
    }
  }
  if (hns.status != 1) return fail(LAKER_ENUMERIC, "CCCP build: Newton-Schulz inverse square root did not converge");
  LAKER_TRY(pc->S.ensure(nn * 8));
  // S = (Z + Z^T)/(2 sqrt(c)) - a^{-1/2} I
  hipLaunchKernelGGL(sym_scale_frob_kernel, dim3((unsigned)ceil_div(nn, 256)), dim3(256), 0, st, Zb.as<double>(),
                     pc->S.as<double>(), (int64_t)nr, (const double*)fro2, pc->inv_sqrt_a);
  LAKER_KCHECK("sym_scale_frob_kernel");
  if (iterations) *iterations = it_done;
  return LAKER_OK;
}


// ---------------------------------------------------------------------------
// Batched launches of the multi-RHS apply's kernels (see batch.h).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void qtx_mfma_batched(const BQtx* __restrict__ desc, const int* __restrict__ off,
                                                        int nsys) {
  const int s = batch_system(off, nsys, (int)blockIdx.x);
  const BQtx d = desc[s];
  const int lb = (int)blockIdx.x - off[s];
  qtx_mfma_kernel_body(d.Q, d.n, d.nr, d.X, d.ldx, d.k, d.rps, d.part, lb % d.gx, lb / d.gx);
}

__global__ __launch_bounds__(256) void reduce_parts_batched(const BRedParts* __restrict__ desc,
                                                            const int* __restrict__ off, int nsys) {
  const int s = batch_system(off, nsys, (int)blockIdx.x);
  const BRedParts d = desc[s];
  const int lb = (int)blockIdx.x - off[s];
  reduce_parts_groups_kernel_body(d.part, d.nb, d.width, d.out, lb % d.gx, lb / d.gx, d.gy);
}

template <int NB>
__global__ __launch_bounds__(256) void qw_mfma_batched(const BQw* __restrict__ desc, const int* __restrict__ off,
                                                       int nsys) {
  const int s = batch_system(off, nsys, (int)blockIdx.x);
  const BQw d = desc[s];
  qw_mfma_kernel_body<NB>(d.Q, d.W2, d.n, d.nr, d.P, d.k, d.j0, d.X, d.ldx, d.Z, d.ldz, d.f,
                          (int)blockIdx.x - off[s], 0);
}

int batched_qtx(const BatchLaunch<BQtx>& L, hipStream_t st) {
  if (L.nsys <= 0) return LAKER_OK;
  hipLaunchKernelGGL(qtx_mfma_batched, dim3((unsigned)L.blocks), dim3(256), 0, st, L.desc, L.off, L.nsys);
  LAKER_KCHECK("qtx_mfma_batched");
  return LAKER_OK;
}

int batched_reduce_parts(const BatchLaunch<BRedParts>& L, hipStream_t st) {
  if (L.nsys <= 0) return LAKER_OK;
  hipLaunchKernelGGL(reduce_parts_batched, dim3((unsigned)L.blocks), dim3(256), 0, st, L.desc, L.off, L.nsys);
  LAKER_KCHECK("reduce_parts_batched");
  return LAKER_OK;
}

int batched_qw(const BatchLaunch<BQw>& L, int nb, int P, hipStream_t st) {
  if (L.nsys <= 0) return LAKER_OK;
  const size_t lds = (size_t)16 * nb * P * sizeof(float);
  if (nb == 2)
    hipLaunchKernelGGL((qw_mfma_batched<2>), dim3((unsigned)L.blocks), dim3(256), lds, st, L.desc, L.off, L.nsys);
  else
    hipLaunchKernelGGL((qw_mfma_batched<1>), dim3((unsigned)L.blocks), dim3(256), lds, st, L.desc, L.off, L.nsys);
  LAKER_KCHECK("qw_mfma_batched");
  return LAKER_OK;
}

int qw_block_rows() { return QW_BLOCK_ROWS; }

int precond_batch_view(Precond& pc, int k, BatchApplyView* v) {
  const int64_t n = pc.n;
  const int nr = pc.nr;
  if (pc.kind != PRECOND_CCCP || pc.dtype != LAKER_F32 || k < 2 || k > 32 || nr % 4 != 0 ||
      (size_t)16 * qw_stride(nr) * sizeof(float) > 64 * 1024)
    return 1;
  // the geometry of apply_t's multi-RHS MFMA branch
  const int cblk = (int)ceil_div(nr, 32);
  const int64_t nsplit = std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), ceil_div(1024, cblk)));
  const int64_t rps = ceil_div(n, nsplit);
  const int64_t ns_eff = ceil_div(n, rps);
  constexpr int RG = 8;
  const int64_t width = (int64_t)nr * k;
  LAKER_TRY(pc.w1.ensure((size_t)nr * k * 8));
  LAKER_TRY(pc.w2.ensure((size_t)nr * k * 8));
  LAKER_TRY(pc.part.ensure(((size_t)ns_eff + RG) * width * 8));
  const int P = qw_stride(nr);
  v->Q = pc.Q.as<float>();
  v->S = pc.S.as<double>();
  v->W1 = pc.w1.as<double>();
  v->W2 = pc.w2.as<double>();
  v->part = pc.part.as<double>();
  v->part2 = v->part + (size_t)ns_eff * width;
  v->n = n;
  v->rps = rps;
  v->nr = nr;
  v->cblk = cblk;
  v->ns_eff = (int)ns_eff;
  v->width = (int)width;
  v->P = P;
  v->qw_blocks = (int)ceil_div(n, QW_BLOCK_ROWS);
  // one pass: qw_mfma_kernel<2> for k > 16 (when its LDS fits), else <1> (k <= 16)
  v->qw_nb = (k > 16 && (size_t)32 * P * sizeof(float) <= 64 * 1024) ? 2 : 1;
  if (v->qw_nb == 1 && k > 16) return 1;
  v->f = pc.inv_sqrt_a;
  return LAKER_OK;
}

}  // namespace laker

using namespace laker;

extern "C" {

int laker_precond_build(laker_operator_t oh, int32_t nr, double gamma, double epsilon,
                        double base_rho, int32_t max_iter, double tol, const void* probes,
                        uint64_t seed, int32_t probe_precision, hipStream_t st,
                        laker_precond_t* out, int32_t* iterations) {
  LAKER_CHECK_ARG(oh != nullptr && out != nullptr, "operator/out handle is NULL");
  Operator& op = *reinterpret_cast<Operator*>(oh);
  const int64_t n = op.n;
  LAKER_CHECK_ARG(nr >= 1 && nr <= n, "num_probes must be in [1, n]");
  LAKER_CHECK_ARG(nr <= 4096, "num_probes > 4096 not supported");
  LAKER_CHECK_ARG(max_iter >= 1, "max_iter must be >= 1");
  std::unique_ptr<Precond> pc(new Precond());
  pc->n = n;
  pc->nr = nr;
  pc->dtype = op.dtype;
  const size_t esz = op.dtype == LAKER_F32 ? 4 : 8;
  const int64_t nn = (int64_t)nr * nr;

  // 1. probes and responses U = A R
  DevBuf probe_buf, U;
  const void* Rp = probes;
  // probes drawn here for the hi/lo probe product are BF16-exact Gaussians,
  // so that product needs two BF16 products instead of three
  const bool xbf16 = !Rp && op.dtype == LAKER_F32 && probe_precision == LAKER_PREC_HILO;
  if (!Rp) {
    LAKER_TRY(probe_buf.ensure((size_t)n * nr * esz, st));
    if (op.dtype == LAKER_F32)
      hipLaunchKernelGGL((randn_kernel<float>), dim3((unsigned)ceil_div(n * nr, 256)), dim3(256), 0, st,
                         probe_buf.as<float>(), n * nr, seed, xbf16 ? 1 : 0);
    else
      hipLaunchKernelGGL((randn_kernel<double>), dim3((unsigned)ceil_div(n * nr, 256)), dim3(256), 0, st,
                         probe_buf.as<double>(), n * nr, seed);
    LAKER_KCHECK("randn_kernel");
    Rp = probe_buf.ptr;
  }
  LAKER_TRY(U.ensure((size_t)n * nr * esz, st));
  const bool clamp = op.clamp_for(LAKER_CLAMP_AUTO);
  if (op.dtype == LAKER_F32) {
    int prec = probe_precision;
    if (prec == LAKER_PREC_FP64) prec = LAKER_PREC_FP32;
    if (xbf16) prec = PREC_HILO_XBF16;
    LAKER_TRY(fused_product_f32(op, prec, op.rows(), op.cols(), (const float*)Rp, nr, nr, U.as<float>(),
                                nr, op.lambda_reg, (const float*)Rp, nr, clamp, false, 0, st));
  } else {
    LAKER_TRY(fused_product_f64<double>(op, op.rows(), op.cols(), (const double*)Rp, nr, nr,
                                        U.as<double>(), nr, op.lambda_reg, (const double*)Rp, nr, clamp,
                                        nullptr, st));
  }
  probe_buf.release();

  int its = 0;
  LAKER_TRY(build_from_u(op, pc, U, nr, gamma, epsilon, base_rho, max_iter, tol, st, &its));
  if (iterations) *iterations = its;
  *out = reinterpret_cast<laker_precond_t>(pc.release());
  return LAKER_OK;
}

int laker_precond_build_from_responses(const void* U_in, int64_t n, int32_t nr, int32_t dtype,
                                       double gamma, double epsilon, double base_rho, int32_t max_iter,
                                       double tol, hipStream_t st, laker_precond_t* out,
                                       int32_t* iterations) {
  LAKER_CHECK_ARG(U_in != nullptr && out != nullptr, "responses/out pointer is NULL");
  LAKER_CHECK_ARG(n >= 1 && nr >= 1 && nr <= n && nr <= 4096, "bad n / num_probes");
  LAKER_CHECK_ARG(dtype == LAKER_F32 || dtype == LAKER_F64, "dtype must be LAKER_F32 or LAKER_F64");
  std::unique_ptr<Precond> pc(new Precond());
  pc->n = n;
  pc->nr = nr;
  pc->dtype = dtype;
  const size_t esz = dtype == LAKER_F32 ? 4 : 8;
  DevBuf U;
  LAKER_TRY(U.ensure((size_t)n * nr * esz, st));
  LAKER_HIP_CHECK(hipMemcpyAsync(U.ptr, U_in, (size_t)n * nr * esz, hipMemcpyDeviceToDevice, st));
  Operator dummy;
  int its = 0;
  LAKER_TRY(build_from_u(dummy, pc, U, nr, gamma, epsilon, base_rho, max_iter, tol, st, &its));
  if (iterations) *iterations = its;
  *out = reinterpret_cast<laker_precond_t>(pc.release());
  return LAKER_OK;
}

int laker_precond_create_jacobi(const void* diagonal, int64_t n, int32_t dtype, hipStream_t st,
                                laker_precond_t* out) {
  LAKER_CHECK_ARG(diagonal != nullptr && out != nullptr && n >= 1, "bad Jacobi preconditioner arguments");
  LAKER_CHECK_ARG(dtype == LAKER_F32 || dtype == LAKER_F64, "dtype must be LAKER_F32 or LAKER_F64");
  std::unique_ptr<Precond> pc(new Precond());
  pc->kind = PRECOND_JACOBI;
  pc->n = n;
  pc->nr = 0;
  pc->dtype = dtype;
  const dim3 g((unsigned)ceil_div(n, 256)), b(256);
  if (dtype == LAKER_F32) {
    LAKER_TRY(pc->inv_diag.ensure((size_t)n * 4));
    hipLaunchKernelGGL((jacobi_inv_kernel<float>), g, b, 0, st, (const float*)diagonal, pc->inv_diag.as<float>(), n,
                       1.1920928955078125e-07f);
  } else {
    LAKER_TRY(pc->inv_diag.ensure((size_t)n * 8));
    hipLaunchKernelGGL((jacobi_inv_kernel<double>), g, b, 0, st, (const double*)diagonal,
                       pc->inv_diag.as<double>(), n, 2.220446049250313e-16);
  }
  LAKER_KCHECK("jacobi_inv_kernel");
  *out = reinterpret_cast<laker_precond_t>(pc.release());
  return LAKER_OK;
}

int laker_precond_from_factors(const void* Q, int64_t n, int32_t nr, int32_t dtype, double isotropic_coef,
                               double rho, const double* R, const double* C, const double* S, hipStream_t st,
                               laker_precond_t* out) {
  LAKER_CHECK_ARG(Q && R && C && S && out, "from_factors needs Q, R, C, S and an out handle");
  LAKER_CHECK_ARG(nr >= 1 && nr <= n && nr <= 4096, "num_probes must be in [1, min(n, 4096)]");
  LAKER_CHECK_ARG(dtype == LAKER_F32 || dtype == LAKER_F64, "dtype must be LAKER_F32 or LAKER_F64");
  LAKER_CHECK_ARG(isotropic_coef > 0.0 && std::isfinite(isotropic_coef), "isotropic_coef must be positive");
  std::unique_ptr<Precond> pc(new Precond());
  pc->n = n;
  pc->nr = nr;
  pc->dtype = dtype;
  pc->a = isotropic_coef;
  pc->rho = rho;
  pc->inv_sqrt_a = 1.0 / std::sqrt(isotropic_coef);
  const size_t qb = (size_t)n * nr * (dtype == LAKER_F32 ? 4 : 8), nn8 = (size_t)nr * nr * 8;
  LAKER_TRY(pc->Q.ensure(qb));
  LAKER_TRY(pc->R.ensure(nn8));
  LAKER_TRY(pc->C.ensure(nn8));
  LAKER_TRY(pc->S.ensure(nn8));
  LAKER_HIP_CHECK(hipMemcpyAsync(pc->Q.ptr, Q, qb, hipMemcpyDeviceToDevice, st));
  LAKER_HIP_CHECK(hipMemcpyAsync(pc->R.ptr, R, nn8, hipMemcpyDeviceToDevice, st));
  LAKER_HIP_CHECK(hipMemcpyAsync(pc->C.ptr, C, nn8, hipMemcpyDeviceToDevice, st));
  LAKER_HIP_CHECK(hipMemcpyAsync(pc->S.ptr, S, nn8, hipMemcpyDeviceToDevice, st));
  *out = reinterpret_cast<laker_precond_t>(pc.release());
  return LAKER_OK;
}

int laker_orthonormalize(void* X, int64_t n, int32_t k, int32_t dtype, hipStream_t st) {
  LAKER_CHECK_ARG(X != nullptr && k >= 1 && n >= k, "orthonormalize needs an n x k block with n >= k >= 1");
  LAKER_CHECK_ARG(k <= 4096, "orthonormalize: k > 4096 not supported");
  LAKER_CHECK_ARG(dtype == LAKER_F32 || dtype == LAKER_F64, "dtype must be LAKER_F32 or LAKER_F64");
  const int64_t nk = n * k;
  const dim3 eg((unsigned)ceil_div(nk, 256)), eb(256);
  DevBuf W, ss, V, Racc, ws;
  LAKER_TRY(W.ensure((size_t)nk * 8, st));
  LAKER_TRY(ss.ensure((size_t)k * 8, st));
  LAKER_TRY(V.ensure((size_t)nk * 8, st));
  LAKER_TRY(Racc.ensure((size_t)k * k * 8, st));
  ws.bind(st);
  // FP64 copy with unit columns (Q is unchanged by a positive column scaling)
  if (dtype == LAKER_F32)
    hipLaunchKernelGGL((convert_kernel<float, double>), eg, eb, 0, st, (const float*)X, W.as<double>(), nk);
  else
    LAKER_HIP_CHECK(hipMemcpyAsync(W.ptr, X, (size_t)nk * 8, hipMemcpyDeviceToDevice, st));
  LAKER_KCHECK("convert_kernel");
  DevBuf chains;
  LAKER_TRY(colsumsq(W.as<double>(), n, k, ss.as<double>(), chains, st));
  hipLaunchKernelGGL((colscale_kernel<double>), dim3((unsigned)ceil_div(k, 256), (unsigned)ceil_div(n, 64)), eb, 0, st,
                     W.as<double>(), n, k, (const double*)ss.as<double>(), 1e-300, V.as<double>());
  LAKER_KCHECK("colscale_kernel");
  const int rc = cholqr_block(W.ptr, LAKER_F64, n, k, V.as<double>(), Racc.as<double>(), &ws, st, true);
  if (rc == LAKER_ENUMERIC) return fail(LAKER_ENUMERIC, "orthonormalize: the block is numerically rank-deficient");
  LAKER_TRY(rc);
  if (dtype == LAKER_F32)
    hipLaunchKernelGGL((convert_kernel<double, float>), eg, eb, 0, st, V.as<double>(), (float*)X, nk);
  else
    LAKER_HIP_CHECK(hipMemcpyAsync(X, V.ptr, (size_t)nk * 8, hipMemcpyDeviceToDevice, st));
  LAKER_KCHECK("convert_kernel");
  return LAKER_OK;
}

int laker_randn(void* out, int64_t count, int32_t dtype, uint64_t seed, hipStream_t st) {
  LAKER_CHECK_ARG(out != nullptr && count >= 0, "bad randn arguments");
  LAKER_CHECK_ARG(dtype == LAKER_F32 || dtype == LAKER_F64, "dtype must be LAKER_F32 or LAKER_F64");
  if (count == 0) return LAKER_OK;
  const dim3 g((unsigned)ceil_div(count, 256)), b(256);
  if (dtype == LAKER_F32)
    hipLaunchKernelGGL((randn_kernel<float>), g, b, 0, st, (float*)out, count, seed);
  else
    hipLaunchKernelGGL((randn_kernel<double>), g, b, 0, st, (double*)out, count, seed);
  LAKER_KCHECK("randn_kernel");
  return LAKER_OK;
}

int laker_precond_copy_out(laker_precond_t h, int32_t which, void* dst, hipStream_t st) {
  LAKER_CHECK_ARG(h != nullptr && dst != nullptr, "bad copy_out arguments");
  Precond* pc = reinterpret_cast<Precond*>(h);
  if (pc->kind == PRECOND_JACOBI) {
    LAKER_CHECK_ARG(which == 4, "a Jacobi preconditioner holds only its inverse diagonal (which = 4)");
    LAKER_HIP_CHECK(hipMemcpyAsync(dst, pc->inv_diag.ptr, (size_t)pc->n * (pc->dtype == LAKER_F32 ? 4 : 8),
                                   hipMemcpyDeviceToDevice, st));
    return LAKER_OK;
  }
  const int64_t nn = (int64_t)pc->nr * pc->nr;
  const void* src = nullptr;
  size_t bytes = 0;
  switch (which) {
    case 0: src = pc->Q.ptr; bytes = (size_t)pc->n * pc->nr * (pc->dtype == LAKER_F32 ? 4 : 8); break;
    case 1: src = pc->R.ptr; bytes = nn * 8; break;
    case 2: src = pc->C.ptr; bytes = nn * 8; break;
    case 3: src = pc->S.ptr; bytes = nn * 8; break;
    default: return fail(LAKER_EINVAL, "which must be 0 (Q), 1 (R), 2 (C) or 3 (S)");
  }
  LAKER_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st));
  return LAKER_OK;
}

int laker_precond_destroy(laker_precond_t h) {
  delete reinterpret_cast<Precond*>(h);
  return LAKER_OK;
}

int laker_precond_info(laker_precond_t h, double* isotropic_coef, int32_t* nr, double* rho,
                       const void** q_basis, const double** qr_r, const double** c_matrix,
                       const double** s_matrix) {
  LAKER_CHECK_ARG(h != nullptr, "preconditioner handle is NULL");
  Precond* pc = reinterpret_cast<Precond*>(h);
  LAKER_CHECK_ARG(pc->kind == PRECOND_CCCP, "not a CCCP preconditioner");
  if (isotropic_coef) *isotropic_coef = pc->a;
  if (nr) *nr = pc->nr;
  if (rho) *rho = pc->rho;
  if (q_basis) *q_basis = pc->Q.ptr;
  if (qr_r) *qr_r = pc->R.as<double>();
  if (c_matrix) *c_matrix = pc->C.as<double>();
  if (s_matrix) *s_matrix = pc->S.as<double>();
  return LAKER_OK;
}

int laker_precond_apply(laker_precond_t h, const void* X, int32_t k, int64_t ldx, void* Z, int64_t ldz,
                        hipStream_t stream) {
  LAKER_CHECK_ARG(h != nullptr, "Preconditioner has not been built. Call build() first.");
  LAKER_CHECK_ARG(X && Z && k >= 1 && ldx >= k && ldz >= k, "bad apply arguments");
  return precond_apply(*reinterpret_cast<Precond*>(h), X, k, ldx, Z, ldz, stream);
}

}  // extern "C"
