// Hand-written backward of the fused matvec, recomputing Gram tiles instead of
// storing them.  For Y = lambda X + G X with G_ij = exp(min(s e_i.e_j, c)):
//   dX   = lambda dY + G dY                       (G symmetric)
//   dE_i = s sum_c [ dY_ic (G'(X_c o E))_i + X_ic (G'(dY_c o E))_i ]
// where G' is G with clamped entries zeroed (their derivative is 0) and
// (X_c o E)_j = X_jc e_j.  Both products are fused multi-RHS matvecs with
// 2kd right-hand sides.  This replaces autograd through exp_safe's
// out-of-place branch (laker/kernels.py:100-101,111-112) on the path of
// laker/training.py:186-211 and laker/core.py:782-849.
#include "laker_common.h"
#include "laker_internal.h"

namespace laker {

// U[j][c*d + q] = X[j][c] * E[j][q]   (n x k*d)
__global__ void outer_rhs_kernel(const float* __restrict__ X, const double* __restrict__ E64, int64_t n,
                                 int k, int d, float* __restrict__ U, int64_t ldu, int col0) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int kd = k * d;
  if (idx >= n * kd) return;
  const int64_t j = idx / kd;
  const int r = (int)(idx % kd);
  const int c = r / d, q = r % d;
  U[j * ldu + col0 + r] = X[j * k + c] * (float)E64[j * d + q];
}

// dE[i][q] (+)= s * sum_c ( A[i][c] W[i][colA + c*d + q] + B[i][c] W[i][colB + c*d + q] )
__global__ void contract_grad_kernel(const float* __restrict__ A, const float* __restrict__ B,
                                     const float* __restrict__ W, int64_t ldw, int colA, int colB,
                                     int64_t n, int k, int d, double s, float* __restrict__ dE,
                                     int accumulate) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * d) return;
  const int64_t i = idx / d;
  const int q = (int)(idx % d);
  double acc = 0.0;
  for (int c = 0; c < k; ++c) {
    if (A) acc += (double)A[i * k + c] * (double)W[i * ldw + colA + c * d + q];
    if (B) acc += (double)B[i * k + c] * (double)W[i * ldw + colB + c * d + q];
  }
  acc *= s;
  if (accumulate) acc += (double)dE[idx];
  dE[idx] = (float)acc;
}

}  // namespace laker

using namespace laker;

extern "C" {

int laker_matvec_backward(laker_operator_t h, const float* X, const float* dY, int32_t k, float* dX,
                          float* dE, hipStream_t st) {
  LAKER_CHECK_ARG(h != nullptr, "operator handle is NULL");
  LAKER_CHECK_ARG(dY && k >= 1, "bad backward arguments");
  Operator& op = *reinterpret_cast<Operator*>(h);
  LAKER_CHECK_ARG(op.dtype == LAKER_F32, "hand-written backward supports FP32 operators");
  const bool clamp = op.clamp_for(LAKER_CLAMP_AUTO);
  const int64_t n = op.n;
  const int d = op.d;
  if (dX) {
    LAKER_TRY(fused_product_f32(op, LAKER_PREC_FP32, op.rows(), op.cols(), dY, k, k, dX, k, op.lambda_reg,
                                dY, k, clamp, false, 0, st));
  }
  if (dE) {
    LAKER_CHECK_ARG(X != nullptr, "X is required for the embedding gradient");
    const int kd = k * d;
    const int64_t w = 2 * kd;
    LAKER_TRY(op.scratch.ensure((size_t)n * w * 2 * sizeof(float)));
    float* U = op.scratch.as<float>();
    float* W = U + n * w;
    const double* E64 = op.self.e64.as<double>();
    const dim3 g((unsigned)ceil_div(n * kd, 256)), b(256);
    hipLaunchKernelGGL(outer_rhs_kernel, g, b, 0, st, X, E64, n, k, d, U, w, 0);
    LAKER_KCHECK("outer_rhs_kernel");
    hipLaunchKernelGGL(outer_rhs_kernel, g, b, 0, st, dY, E64, n, k, d, U, w, kd);
    LAKER_KCHECK("outer_rhs_kernel");
    LAKER_HIP_CHECK(hipGetLastError());
    LAKER_TRY(fused_product_f32(op, LAKER_PREC_FP32, op.rows(), op.cols(), U, (int)w, w, W, w, 0.0,
                                nullptr, 0, clamp, true, 0, st));
    // W[:, 0:kd] = G'(X o E), W[:, kd:2kd] = G'(dY o E)
    hipLaunchKernelGGL(contract_grad_kernel, dim3((unsigned)ceil_div(n * d, 256)), b, 0, st, dY, X, W, w, 0,
                       kd, n, k, d, op.scale, dE, 0);
    LAKER_KCHECK("contract_grad_kernel");
    LAKER_HIP_CHECK(hipGetLastError());
  }
  return LAKER_OK;
}

int laker_cross_matvec_backward(laker_operator_t h, const float* Eq, int64_t m, const float* X, int32_t k,
                                const float* dYq, float* dEq, float* dE, float* dX, int32_t clamp_policy,
                                hipStream_t st) {
  LAKER_CHECK_ARG(h != nullptr, "operator handle is NULL");
  LAKER_CHECK_ARG(Eq && dYq && m >= 1 && k >= 1, "bad cross backward arguments");
  Operator& op = *reinterpret_cast<Operator*>(h);
  LAKER_CHECK_ARG(op.dtype == LAKER_F32, "hand-written backward supports FP32 operators");
  const bool clamp = op.clamp_for(clamp_policy);
  const int64_t n = op.n;
  const int d = op.d;
  const int kd = k * d;
  std::lock_guard<std::mutex> g(op.query_mu);
  // pack the query set on both sides
  LAKER_TRY(pack_store(op.query, Eq, LAKER_F32, m, d, op.ka32, op.ka16, op.scale, true, true, st));
  Side qrows = Operator::view(op.query, true), qcols = Operator::view(op.query, false);
  const dim3 b(256);
  if (dX) {
    // dX = G_q^T dYq  (rows: training set, cols: queries)
    LAKER_TRY(fused_product_f32(op, LAKER_PREC_FP32, op.rows(), qcols, dYq, k, k, dX, k, 0.0, nullptr, 0,
                                clamp, false, 0, st));
  }
  const size_t need = ((size_t)n * kd * 2 + (size_t)m * kd * 2) * sizeof(float);
  LAKER_TRY(op.scratch.ensure(need));
  float* U = op.scratch.as<float>();       // n x kd : X o E
  float* Wq = U + n * kd;                  // m x kd : G'_q U
  float* V = Wq + m * kd;                  // m x kd : dYq o Eq
  float* Wn = V + m * kd;                  // n x kd : G'_q^T V
  if (dEq) {
    LAKER_CHECK_ARG(X != nullptr, "X is required for the query-embedding gradient");
    hipLaunchKernelGGL(outer_rhs_kernel, dim3((unsigned)ceil_div(n * kd, 256)), b, 0, st, X,
                       op.self.e64.as<double>(), n, k, d, U, (int64_t)kd, 0);
    LAKER_KCHECK("outer_rhs_kernel");
    LAKER_HIP_CHECK(hipGetLastError());
    LAKER_TRY(fused_product_f32(op, LAKER_PREC_FP32, qrows, op.cols(), U, kd, kd, Wq, kd, 0.0, nullptr, 0,
                                clamp, true, 0, st));
    hipLaunchKernelGGL(contract_grad_kernel, dim3((unsigned)ceil_div(m * d, 256)), b, 0, st, dYq,
                       (const float*)nullptr, Wq, (int64_t)kd, 0, 0, m, k, d, op.scale, dEq, 1);
    LAKER_KCHECK("contract_grad_kernel");
    LAKER_HIP_CHECK(hipGetLastError());
  }
  if (dE) {
    LAKER_CHECK_ARG(X != nullptr, "X is required for the embedding gradient");
    hipLaunchKernelGGL(outer_rhs_kernel, dim3((unsigned)ceil_div(m * kd, 256)), b, 0, st, dYq,
                       op.query.e64.as<double>(), m, k, d, V, (int64_t)kd, 0);
    LAKER_KCHECK("outer_rhs_kernel");
    LAKER_HIP_CHECK(hipGetLastError());
    LAKER_TRY(fused_product_f32(op, LAKER_PREC_FP32, op.rows(), qcols, V, kd, kd, Wn, kd, 0.0, nullptr, 0,
                                clamp, true, 0, st));
    hipLaunchKernelGGL(contract_grad_kernel, dim3((unsigned)ceil_div(n * d, 256)), b, 0, st, X,
                       (const float*)nullptr, Wn, (int64_t)kd, 0, 0, n, k, d, op.scale, dE, 1);
    LAKER_KCHECK("contract_grad_kernel");
    LAKER_HIP_CHECK(hipGetLastError());
  }
  return LAKER_OK;
}

}  // extern "C"
