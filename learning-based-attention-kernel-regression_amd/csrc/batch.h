// Batched launches: one kernel launch runs the same kernel body for several
// independent systems (ragged sizes).  System s owns the flat blocks
// [off[s], off[s + 1]) of the launch; a block finds its system by binary
// search and runs the body with its system's descriptor and its block index
// within the system, so each system's arithmetic (block decomposition,
// summation order) is exactly that of a launch for it alone.
#pragma once
#include "laker_common.h"

namespace laker {

struct PcgState;

// System of flat block b (off[0] = 0 <= b < off[nsys]).
__device__ inline int batch_system(const int* __restrict__ off, int nsys, int b) {
  int lo = 0, hi = nsys - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Flat launch of one batched kernel: per-system descriptors and block offsets
// (device memory), their count, the total block count.
template <class D>
struct BatchLaunch {
  const D* desc = nullptr;
  const int* off = nullptr;
  int nsys = 0;
  int blocks = 0;
};

// ---- descriptors (one per system and kernel)
struct BPackX {  // pack_x_kernel
  const float* X;
  int64_t ldx, n_cols, ntiles;
  int kcols, c_base, nq;
  bf16x8 *xh, *xl, *xm;
};
struct BMulti2 {  // mv_multi2_kernel
  const void *colpack, *rowpack;
  int ka;
  int64_t n_rows, n_cols;
  const bf16x8 *xh, *xl, *xm;
  int kcols, c_base;
  float* part;
  int64_t ldp;
  int jsplit;
  float clampv;
};
struct BReduce {  // mv_reduce_kernel<float, float, float, false>
  const float* part;
  int64_t n_rows, ldp, ldx, ldy;
  int k, jsplit;
  double lam;
  const float* Xr;
  float* Y;
};
struct BColdots {  // coldots_kernel<float, NQ>
  const float *A0, *B0, *A1, *B1, *A2, *B2;
  int64_t n, rpb;
  int k, nblk;
  double* part;
  const PcgState* st;
};
struct BScalar {  // pcg_alpha / pcg_resid / pcg_beta kernels (one block per system)
  const double* part;
  int nblk, k;
  double *rz, *out;
  PcgState* st;
};
struct BUpdX64 {  // pcg_update_x64_kernel
  double* x;
  float* r;
  const float *p, *ap;
  const double* alpha;
  int64_t n;
  int k;
  const PcgState* st;
};
struct BUpdP {  // pcg_update_p_kernel<float>
  float* p;
  const float* z;
  const double* beta;
  int64_t n;
  int k;
  const PcgState* st;
};
struct BQtx {  // qtx_mfma_kernel, grid (gx, gy) flattened x-fastest
  const float* Q;
  int64_t n, ldx, rps;
  int nr, k, gx;
  const float* X;
  double* part;
};
struct BRedParts {  // reduce_parts_groups_kernel, grid (gx, gy)
  const double* part;
  int nb, width, gx, gy;
  double* out;
};
struct BGemm {  // gemm_kernel<double, double, double, false, false>, grid (gx, gy), no split-K
  int64_t M, N, K, lda, ldb, ldc;
  double alpha, beta;
  const double *A, *B;
  double* C;
  int gx;
};
struct BQw {  // qw_mfma_kernel<NB>
  const float* Q;
  const double* W2;
  int64_t n, ldx, ldz;
  int nr, P, k, j0;
  const float* X;
  float* Z;
  double f;
};

// Launchers (each in the translation unit of its kernel body).
int batched_pack_x(const BatchLaunch<BPackX>& L, hipStream_t st);
// multi-RHS product on the 3-way BF16 split (GX_SPLIT3), one 32-column chunk:
// gm = GRAM_F16S or LAKER_PREC_MIXED, ns its Gram steps; all systems share them.
int batched_multi2(const BatchLaunch<BMulti2>& L, int gm, int ns, bool clamp, hipStream_t st);
int batched_reduce(const BatchLaunch<BReduce>& L, hipStream_t st);
int batched_qtx(const BatchLaunch<BQtx>& L, hipStream_t st);
int batched_reduce_parts(const BatchLaunch<BRedParts>& L, hipStream_t st);
int batched_gemm(const BatchLaunch<BGemm>& L, hipStream_t st);
int batched_qw(const BatchLaunch<BQw>& L, int nb, int P, hipStream_t st);
// Buffers and launch geometry of one system's multi-RHS preconditioner apply
// (2 <= k <= 32), exactly those of its unbatched apply.
struct BatchApplyView {
  const float* Q;
  const double* S;
  double *W1, *W2, *part, *part2;
  int64_t n, rps;
  int nr, cblk, ns_eff, width, P, qw_blocks, qw_nb;
  double f;
};
struct Precond;
// Fills v for pc and k; 1 when the handle has no multi-RHS MFMA apply for k.
int precond_batch_view(Precond& pc, int k, BatchApplyView* v);
int qw_block_rows();

// GX_SPLIT3 multi2 launch geometry shared with product_multi_f32
int multi2_split3_jt(int gm, int ns);
size_t multi2_split3_lds(int ns, int jt);

}  // namespace laker
