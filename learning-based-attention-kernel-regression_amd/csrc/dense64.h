// FP64 dense helpers (device, stream-ordered).  See dense64.hip.
#pragma once
#include "laker_common.h"

namespace laker {

// Shapes of gemm's operands and output:
//   GEMM_FULL   every output tile
//   GEMM_LOWER  only the 64 x 64 output tiles on or below the diagonal are
//               formed (the lower triangle of a symmetric product); with
//               beta == 0 the tiles above are written as zeros, else untouched
//   GEMM_BTRI   op(B)[k][n] == 0 for k > n (op(B) upper triangular): each
//               output tile stops its K loop after its last column
constexpr int GEMM_FULL = 0, GEMM_LOWER = 1, GEMM_BTRI = 2;

// C = alpha * op(A) * op(B) + beta * C, row-major, FP64 accumulation.
// op(A) is M x K, op(B) is K x N.  ws (optional) enables split-K.
template <class TA, class TB, class TC>
int gemm(bool tra, bool trb, int64_t M, int64_t N, int64_t K, double alpha, const TA* A,
         int64_t lda, const TB* B, int64_t ldb, double beta, TC* C, int64_t ldc, DevBuf* ws,
         hipStream_t st, int kmode = GEMM_FULL);

// Work space of potrf_lower / trsm_lower_left: the inverses of the 64 x 64
// diagonal blocks of the last factor (dinv).
struct TriWork {
  DevBuf dinv;
  void bind(hipStream_t st) { dinv.bind(st); }
};

// In-place lower Cholesky of the n x n matrix A (upper triangle zeroed).
// info_dev (device int, caller-zeroed) receives index+1 of the first
// non-positive pivot.  Leaves the diagonal-block inverses of L in tw.
int potrf_lower(double* A, int64_t lda, int64_t n, int* info_dev, DevBuf* ws, TriWork& tw,
                hipStream_t st);

// B <- L^{-1} B  (L lower n x n, B n x m), with the diagonal-block inverses
// that potrf_lower left in tw for this L.
int trsm_lower_left(const double* L, int64_t ldl, int64_t n, double* B, int64_t ldb, int64_t m,
                    DevBuf* ws, TriWork& tw, hipStream_t st);

int set_identity(double* A, int64_t n, double diag, hipStream_t st);

}  // namespace laker
