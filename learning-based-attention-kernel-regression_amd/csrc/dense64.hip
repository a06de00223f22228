// Small dense FP64 linear algebra for the CCCP preconditioner (N_r <= 2048)
// and the tall-skinny products of its QR and apply:
//   gemm      C = alpha op(A) op(B) + beta C    (row-major, split-K, FP64 accumulation)
//   potrf     blocked right-looking Cholesky, lower, in place
//   trsm_llnn X = L^{-1} B, blocked forward substitution, in place on B
// These replace the torch.linalg.qr / eigh calls of laker/preconditioner.py:186-193,
// :207-280 and laker/utils.py:109-134 in the FP64 device CCCP build.
#include "laker_common.h"
#include "laker_internal.h"
#include "dense64.h"
#include "batch.h"

namespace laker {

constexpr int GBM = 64, GBN = 64, GBK = 32;
constexpr int GPF = GBK * GBM / 256;  // elements of each operand block per thread

// One 64x64 tile of op(A) op(B) over the K range [k0, k1) on the FP64 MFMA
// (v_mfma_f64_16x16x4f64).  Operands are converted to double while staged in
// LDS (k-major), 32 k per block; the next block is loaded into registers while
// the MFMAs consume the current one, so one global-load latency per block is
// hidden.  Wave w owns the 32x32 quadrant (w / 2, w % 2) as 2x2 MFMA 16x16
// blocks.  Lane l supplies A[m = l % 16][k = l / 16], B[k = l / 16][n = l % 16]
// and receives D[(l / 16) + 4 r][l % 16].
template <class TA, class TB, bool TRA, bool TRB>
__device__ inline void gemm_tile(const TA* __restrict__ A, int64_t lda, const TB* __restrict__ B,
                                 int64_t ldb, int64_t M, int64_t N, int64_t k0, int64_t k1,
                                 int64_t m0, int64_t n0, doublex4 acc[2][2],
                                 double (*As)[GBM + 1], double (*Bs)[GBN + 1]) {
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6, wm = w >> 1, wn = w & 1;
  const int lm = lane & 15, lk = lane >> 4;
  // element u of this thread: (kk, mm) of the A block, (kk, nn) of the B block;
  // consecutive threads take consecutive addresses of the stored layout
  auto a_idx = [&](int u, int& kk, int& mm) {
    const int e = tid + 256 * u;
    if (TRA) { mm = e % GBM; kk = e / GBM; } else { kk = e % GBK; mm = e / GBK; }
  };
  auto b_idx = [&](int u, int& kk, int& nn) {
    const int e = tid + 256 * u;
    if (TRB) { kk = e % GBK; nn = e / GBK; } else { nn = e % GBN; kk = e / GBN; }
  };
  double pa[GPF], pb[GPF];
  auto load = [&](int64_t kb) {
#pragma unroll
    for (int u = 0; u < GPF; ++u) {
      int kk, mm, nn;
      a_idx(u, kk, mm);
      const int64_t gm = m0 + mm, gk = kb + kk;
      pa[u] = (gm < M && gk < k1) ? (TRA ? (double)A[gk * lda + gm] : (double)A[gm * lda + gk]) : 0.0;
      b_idx(u, kk, nn);
      const int64_t gn = n0 + nn, gk2 = kb + kk;
      pb[u] = (gn < N && gk2 < k1) ? (TRB ? (double)B[gn * ldb + gk2] : (double)B[gk2 * ldb + gn]) : 0.0;
    }
  };
  if (k0 < k1) load(k0);
  for (int64_t kb = k0; kb < k1; kb += GBK) {
    __syncthreads();  // the previous block's MFMAs are done with the LDS
#pragma unroll
    for (int u = 0; u < GPF; ++u) {
      int kk, mm, nn;
      a_idx(u, kk, mm);
      As[kk][mm] = pa[u];
      b_idx(u, kk, nn);
      Bs[kk][nn] = pb[u];
    }
    __syncthreads();
    if (kb + GBK < k1) load(kb + GBK);
#pragma unroll
    for (int s4 = 0; s4 < GBK / 4; ++s4) {
      double a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[4 * s4 + lk][32 * wm + 16 * i + lm];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Bs[4 * s4 + lk][32 * wn + 16 * j + lm];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }
}

// grid: (ceil(N/64), ceil(M/64), ksplit).  ksplit == 1 writes C directly.
template <class TA, class TB, class TC, bool TRA, bool TRB>
__device__ __forceinline__ void gemm_kernel_body(int64_t M, int64_t N, int64_t K, double alpha,
                                                   const TA* __restrict__ A, int64_t lda,
                                                   const TB* __restrict__ B, int64_t ldb,
                                                   double beta, TC* __restrict__ C, int64_t ldc,
                                                   double* __restrict__ part, int64_t kchunk, int bx_, int by_, int bz_,
                                                   int kmode = GEMM_FULL) {
  __shared__ double As[GBK][GBM + 1];
  __shared__ double Bs[GBK][GBN + 1];
  const int64_t m0 = (int64_t)by_ * GBM, n0 = (int64_t)bx_ * GBN;
  const int64_t k0 = (int64_t)bz_ * kchunk;
  int64_t k1 = (k0 + kchunk < K) ? k0 + kchunk : K;
  if (kmode == GEMM_LOWER && n0 > m0) {
    // a tile strictly above the diagonal: not formed; zeros when C is overwritten
    // directly (with split-K the reduction writes them)
    if (!part && beta == 0.0) {
      for (int e = threadIdx.x; e < GBM * GBN; e += blockDim.x) {
        const int64_t gm = m0 + e / GBN, gn = n0 + e % GBN;
        if (gm < M && gn < N) C[gm * ldc + gn] = (TC)0.0;
      }
    }
    return;
  }
  // op(B)[k][n] = 0 for k > n: the tile's columns need k < n0 + GBN only
  if (kmode == GEMM_BTRI && n0 + GBN < k1) k1 = n0 + GBN;
  doublex4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = doublex4{0.0, 0.0, 0.0, 0.0};
  gemm_tile<TA, TB, TRA, TRB>(A, lda, B, ldb, M, N, k0, k1, m0, n0, acc, As, Bs);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wm = w >> 1, wn = w & 1;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gm = m0 + 32 * wm + 16 * i + (lane >> 4) + 4 * r;
        const int64_t gn = n0 + 32 * wn + 16 * j + (lane & 15);
        if (gm < M && gn < N) {
          if (part) {
            part[((int64_t)bz_ * M + gm) * N + gn] = acc[i][j][r];
          } else {
            double v = alpha * acc[i][j][r];
            if (beta != 0.0) v += beta * (double)C[gm * ldc + gn];
            C[gm * ldc + gn] = (TC)v;
          }
        }
      }
}

template <class TA, class TB, class TC, bool TRA, bool TRB>
__global__ __launch_bounds__(256) void gemm_kernel(int64_t M, int64_t N, int64_t K, double alpha,
                                                   const TA* __restrict__ A, int64_t lda,
                                                   const TB* __restrict__ B, int64_t ldb,
                                                   double beta, TC* __restrict__ C, int64_t ldc,
                                                   double* __restrict__ part, int64_t kchunk, int kmode) {
  gemm_kernel_body<TA, TB, TC, TRA, TRB>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, part, kchunk, (int)blockIdx.x,
                                         (int)blockIdx.y, (int)blockIdx.z, kmode);
}

template <class TC>
__global__ void gemm_splitk_reduce(const double* __restrict__ part, int64_t M, int64_t N, int ks,
                                   double alpha, double beta, TC* __restrict__ C, int64_t ldc, int kmode) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= M * N) return;
  const int64_t m = idx / N, n = idx % N;
  if (kmode == GEMM_LOWER && n / GBN > m / GBM) {  // tile not formed
    if (beta == 0.0) C[m * ldc + n] = (TC)0.0;
    return;
  }
  double s = 0.0;
  for (int z = 0; z < ks; ++z) s += part[(int64_t)z * M * N + idx];
  double v = alpha * s;
  if (beta != 0.0) v += beta * (double)C[m * ldc + n];
  C[m * ldc + n] = (TC)v;
}

template <class TA, class TB, class TC>
int gemm(bool tra, bool trb, int64_t M, int64_t N, int64_t K, double alpha, const TA* A,
         int64_t lda, const TB* B, int64_t ldb, double beta, TC* C, int64_t ldc, DevBuf* ws,
         hipStream_t st, int kmode) {
  if (M <= 0 || N <= 0) return LAKER_OK;
  const int64_t tiles = ceil_div(M, GBM) * ceil_div(N, GBN);
  int ks = 1;
  if (ws) {
    while (tiles * ks < 512 && K / (ks * 2) >= 256 && ks < 256) ks *= 2;
  }
  const int64_t kchunk = round_up(ceil_div(std::max<int64_t>(K, 1), ks), GBK);
  ks = (int)ceil_div(std::max<int64_t>(K, 1), kchunk);
  double* part = nullptr;
  if (ks > 1) {
    LAKER_TRY(ws->ensure((size_t)ks * M * N * sizeof(double)));
    part = ws->as<double>();
  }
  dim3 grid((unsigned)ceil_div(N, GBN), (unsigned)ceil_div(M, GBM), (unsigned)ks), block(256);
#define LAKER_GEMM(TRA_, TRB_)                                                                  \
  hipLaunchKernelGGL((gemm_kernel<TA, TB, TC, TRA_, TRB_>), grid, block, 0, st, M, N, K, alpha, A, \
                     lda, B, ldb, beta, C, ldc, part, kchunk, kmode)
  if (tra) {
    if (trb) LAKER_GEMM(true, true); else LAKER_GEMM(true, false);
  } else {
    if (trb) LAKER_GEMM(false, true); else LAKER_GEMM(false, false);
  }
#undef LAKER_GEMM
  LAKER_HIP_CHECK(hipGetLastError());
  if (ks > 1) {
    hipLaunchKernelGGL((gemm_splitk_reduce<TC>), dim3((unsigned)ceil_div(M * N, 256)), dim3(256), 0,
                       st, part, M, N, ks, alpha, beta, C, ldc, kmode);
    LAKER_KCHECK("gemm_splitk_reduce");
    LAKER_HIP_CHECK(hipGetLastError());
  }
  return LAKER_OK;
}

#define LAKER_INST_GEMM(TA, TB, TC)                                                             \
  template int gemm<TA, TB, TC>(bool, bool, int64_t, int64_t, int64_t, double, const TA*,     \
                                int64_t, const TB*, int64_t, double, TC*, int64_t, DevBuf*,  \
                                hipStream_t, int);
LAKER_INST_GEMM(double, double, double)
LAKER_INST_GEMM(float, float, double)
LAKER_INST_GEMM(float, double, double)
LAKER_INST_GEMM(float, double, float)
LAKER_INST_GEMM(double, double, float)
LAKER_INST_GEMM(float, float, float)
LAKER_INST_GEMM(double, float, double)
LAKER_INST_GEMM(double, float, float)
#undef LAKER_INST_GEMM

// ---------------------------------------------------------------------------
// Cholesky (lower), blocked with NB = 64.  The diagonal block is factored in
// LDS together with the inverse of its factor; the panel solve and the
// triangular solves then become GEMMs with those 64x64 inverses.
// ---------------------------------------------------------------------------
constexpr int PNB = 64;

// Unblocked Cholesky of the nb x nb diagonal block at (k0,k0) and the inverse
// of its factor into dinv (64 x 64 row-major, identity-padded beyond nb), by
// four waves: lane i of wave w holds row i, columns w, w + 4, ..., w + 60 of
// the block in registers (the steps are unrolled at compile time, so every
// register index is fixed) and nothing branches on data.  Right-looking: at
// step k the wave owning column k takes the pivot from lane k, scales its
// entries of the column by 1 / sqrt(pivot) and publishes the column in LDS;
// after one barrier every wave applies l_ik l_jk to its entries j > k, so
// every entry receives its updates in the order k = 0, 1, ...  With a
// quarter of a row per lane, the column values a step needs are read from
// LDS together and the barrier is the only wait.  Sets *info = global index
// + 1 of the first non-positive pivot (replaced by 1e300 so the factorisation
// runs on without NaNs).
//
// The inverse runs one step behind the factorisation inside the same barrier
// intervals: column t = i of L^{-1} by forward substitution, column-oriented
// (once x_p is known, every later row takes v_q -= L[q][p] x_p, so each v_q
// receives its terms in the order p = 0, 1, ... of the row-oriented dot
// product).  Step p needs column p of L, published one interval earlier, and
// its pivot division runs on another wave than the factorisation's square
// root, so the two chains overlap and the block takes 65 barriers, not 128.
// v_p stays exactly zero for p < t (zero right-hand side, zero x so far), so
// x_p = v_p / L[p][p] is zero there without a test.  Wave w holds the rows
// 4 c + w of the column.
__device__ __forceinline__ double lane_bcast(double v, int src) {
  const int2 w = __builtin_bit_cast(int2, v);
  return __builtin_bit_cast(double, int2{__builtin_amdgcn_readlane(w.x, src), __builtin_amdgcn_readlane(w.y, src)});
}

constexpr int PW = PNB / 4;  // columns (factorisation) or rows (inverse) per wave

struct PotrfShared {
  double col[2][PNB];          // column k of L, double-buffered across steps
  double xb[2][PNB];           // row p of the inverse, double-buffered across steps
  double Lc[PNB * (PNB + 1)];  // L column-major: Lc[c (PNB + 1) + r] = L[r][c]
  int bad_at;                  // first non-positive pivot (local index) or -1
};

// Barrier interval K for wave w: column K of the factor (K < PNB) and entry
// P = K - 1 of the inverse columns (K >= 1), then the updates both publish.
template <int K>
__device__ __forceinline__ void potrf_step(double (&a)[PW], double (&v)[PW], int i, int w, int nb,
                                           double* __restrict__ dinv, PotrfShared& sh) {
  constexpr int P = K - 1;
  if constexpr (K < PNB) {
    if ((K & 3) == w) {  // uniform over the wave
      constexpr int kl = K >> 2;
      double piv = lane_bcast(a[kl], K);
      const bool bad = !(piv > 0.0);
      if (bad && K < nb && i == 0 && sh.bad_at < 0) sh.bad_at = K;
      piv = bad ? 1e300 : piv;
      const double d = sqrt(piv);
      const double r = 1.0 / d;
      a[kl] = i == K ? d : (i > K ? a[kl] * r : 0.0);
      sh.col[K & 1][i] = a[kl];
      sh.Lc[K * (PNB + 1) + i] = a[kl];
    }
  }
  if constexpr (P >= 0) {
    if ((P & 3) == w) {
      const double xp = v[P >> 2] / sh.Lc[P * (PNB + 1) + P];
      dinv[P * PNB + i] = xp;
      sh.xb[P & 1][i] = xp;
    }
  }
  __syncthreads();
  if constexpr (K < PNB) {
    constexpr int c1 = K >> 2;  // columns 4 c + w > K: all c > c1, and c = c1 when w > K & 3
    const double lik = sh.col[K & 1][i];
    double lj[PW];
#pragma unroll
    for (int c = c1; c < PW; ++c) lj[c] = sh.col[K & 1][4 * c + w];
    if (w > (K & 3)) a[c1] -= lik * lj[c1];
#pragma unroll
    for (int c = c1 + 1; c < PW; ++c) a[c] -= lik * lj[c];
  }
  if constexpr (P >= 0 && (P >> 2) < PW) {
    constexpr int c1 = P >> 2;
    const double* lp = sh.Lc + P * (PNB + 1);
    const double xp = sh.xb[P & 1][i];
    double lq[PW];
#pragma unroll
    for (int c = c1; c < PW; ++c) lq[c] = lp[4 * c + w];
    if (w > (P & 3)) v[c1] -= lq[c1] * xp;
#pragma unroll
    for (int c = c1 + 1; c < PW; ++c) v[c] -= lq[c] * xp;
  }
}

template <int K>
struct PotrfSteps {
  static __device__ __forceinline__ void run(double (&a)[PW], double (&v)[PW], int i, int w, int nb,
                                             double* __restrict__ dinv, PotrfShared& sh) {
    potrf_step<K>(a, v, i, w, nb, dinv, sh);
    PotrfSteps<K + 1>::run(a, v, i, w, nb, dinv, sh);
  }
};
template <>
struct PotrfSteps<PNB + 1> {
  static __device__ __forceinline__ void run(double (&)[PW], double (&)[PW], int, int, int, double* __restrict__,
                                             PotrfShared&) {}
};

__global__ __launch_bounds__(256) void potrf_diag_kernel(double* __restrict__ A, int64_t lda, int64_t k0, int nb,
                                                         int* __restrict__ info, double* __restrict__ dinv) {
  __shared__ PotrfShared sh;
  const int i = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (threadIdx.x == 0) sh.bad_at = -1;
  double a[PW], v[PW];
  {
    const bool rin = i < nb;
    const double* row = A + (k0 + (rin ? i : 0)) * lda + k0;
#pragma unroll
    for (int c = 0; c < PW; ++c) {
      const int cc = 4 * c + w;
      const double g = row[cc < nb ? cc : 0];
      a[c] = (rin && cc < nb) ? g : (i == cc ? 1.0 : 0.0);
      v[c] = cc == i ? 1.0 : 0.0;
    }
  }
  __syncthreads();  // bad_at initialised
  PotrfSteps<0>::run(a, v, i, w, nb, dinv, sh);
  if (i < nb) {
    double* row = A + (k0 + i) * lda + k0;
#pragma unroll
    for (int c = 0; c < PW; ++c) {
      const int cc = 4 * c + w;
      if (cc < nb) row[cc] = cc <= i ? a[c] : 0.0;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && sh.bad_at >= 0 && *info == 0) *info = (int)(k0 + sh.bad_at + 1);
}

__global__ void zero_upper_kernel(double* __restrict__ A, int64_t lda, int64_t n) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const int64_t i = idx / n, j = idx % n;
  if (j > i) A[i * lda + j] = 0.0;
}

int potrf_lower(double* A, int64_t lda, int64_t n, int* info_dev, DevBuf* ws, TriWork& tw,
                hipStream_t st) {
  const int64_t nblk = ceil_div(n, PNB);
  LAKER_TRY(tw.dinv.ensure((size_t)nblk * PNB * PNB * sizeof(double)));
  double* dinv = tw.dinv.as<double>();
  for (int64_t k0 = 0; k0 < n; k0 += PNB) {
    const int nb = (int)std::min<int64_t>(PNB, n - k0);
    double* dk = dinv + (k0 / PNB) * PNB * PNB;
    hipLaunchKernelGGL(potrf_diag_kernel, dim3(1), dim3(256), 0, st, A, lda, k0, nb, info_dev, dk);
    LAKER_KCHECK("potrf_diag_kernel");
    const int64_t rem = n - k0 - nb;
    if (rem > 0) {
      // L21 = A21 L11^{-T} in place: the product has one column tile (nb <= 64)
      // and K = nb is too short for split-K, so each workgroup reads its rows
      // of A21 in full before it writes them
      double* A21 = A + (k0 + nb) * lda + k0;
      LAKER_TRY((gemm<double, double, double>(false, true, rem, nb, nb, 1.0, A21, lda, dk, PNB, 0.0,
                                               A21, lda, ws, st)));
      // A22 -= L21 L21^T on the tiles at or below the diagonal (the factor
      // never reads the upper triangle; it is zeroed at the end)
      double* A22 = A + (k0 + nb) * lda + k0 + nb;
      LAKER_TRY((gemm<double, double, double>(false, true, rem, rem, nb, -1.0, A21, lda, A21, lda,
                                               1.0, A22, lda, ws, st, GEMM_LOWER)));
    }
  }
  hipLaunchKernelGGL(zero_upper_kernel, dim3((unsigned)ceil_div(n * n, 256)), dim3(256), 0, st, A,
                     lda, n);
  LAKER_KCHECK("zero_upper_kernel");
  return LAKER_OK;
}

// ---------------------------------------------------------------------------
// X = L^{-1} B in place (B: n x m, row-major ldb), L lower n x n, with the
// diagonal-block inverses that potrf_lower left in tw.
// ---------------------------------------------------------------------------
int trsm_lower_left(const double* L, int64_t ldl, int64_t n, double* B, int64_t ldb, int64_t m,
                    DevBuf* ws, TriWork& tw, hipStream_t st) {
  const int64_t nblk = ceil_div(n, PNB);
  if (tw.dinv.bytes < (size_t)nblk * PNB * PNB * sizeof(double))
    return fail(LAKER_ESTATE, "trsm_lower_left: no diagonal-block inverses (call potrf_lower first)");
  const double* dinv = tw.dinv.as<double>();
  for (int64_t k0 = 0; k0 < n; k0 += PNB) {
    const int nb = (int)std::min<int64_t>(PNB, n - k0);
    const double* dk = dinv + (k0 / PNB) * PNB * PNB;
    // B_k <- L_kk^{-1} B_k in place: one row tile (nb <= 64) and K = nb (no
    // split-K), so each workgroup reads its columns of B_k in full before it
    // writes them
    LAKER_TRY((gemm<double, double, double>(false, false, nb, m, nb, 1.0, dk, PNB, B + k0 * ldb, ldb,
                                             0.0, B + k0 * ldb, ldb, ws, st)));
    const int64_t rem = n - k0 - nb;
    if (rem > 0) {
      // B[k0+nb:, :] -= L[k0+nb:, k0:k0+nb] * B_k
      LAKER_TRY((gemm<double, double, double>(false, false, rem, m, nb, -1.0,
                                               L + (k0 + nb) * ldl + k0, ldl, B + k0 * ldb, ldb, 1.0,
                                               B + (k0 + nb) * ldb, ldb, ws, st)));
    }
  }
  return LAKER_OK;
}

// ---------------------------------------------------------------------------
// Elementwise helpers
// ---------------------------------------------------------------------------
__global__ void set_identity_kernel(double* __restrict__ A, int64_t n, double diag) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  A[idx] = (idx / n == idx % n) ? diag : 0.0;
}

int set_identity(double* A, int64_t n, double diag, hipStream_t st) {
  hipLaunchKernelGGL(set_identity_kernel, dim3((unsigned)ceil_div(n * n, 256)), dim3(256), 0, st, A,
                     n, diag);
  LAKER_KCHECK("set_identity_kernel");
  LAKER_HIP_CHECK(hipGetLastError());
  return LAKER_OK;
}

// Batched FP64 C = alpha A B + beta C (no transposes, no split-K), see batch.h.
__global__ __launch_bounds__(256) void gemm_batched(const BGemm* __restrict__ desc, const int* __restrict__ off,
                                                    int nsys) {
  const int s = batch_system(off, nsys, (int)blockIdx.x);
  const BGemm d = desc[s];
  const int lb = (int)blockIdx.x - off[s];
  const int64_t kk = d.K > 1 ? d.K : 1;
  const int64_t kchunk = (kk + GBK - 1) / GBK * GBK;  // one K chunk (no split-K)
  gemm_kernel_body<double, double, double, false, false>(d.M, d.N, d.K, d.alpha, d.A, d.lda, d.B, d.ldb, d.beta,
                                                         d.C, d.ldc, (double*)nullptr, kchunk, lb % d.gx,
                                                         lb / d.gx, 0);
}

int batched_gemm(const BatchLaunch<BGemm>& L, hipStream_t st) {
  if (L.nsys <= 0) return LAKER_OK;
  hipLaunchKernelGGL(gemm_batched, dim3((unsigned)L.blocks), dim3(256), 0, st, L.desc, L.off, L.nsys);
  LAKER_KCHECK("gemm_batched");
  return LAKER_OK;
}

}  // namespace laker

