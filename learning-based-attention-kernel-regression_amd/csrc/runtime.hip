// Error reporting, device buffers and small column kernels of the LAKER HIP library.
#include "laker_common.h"
#include "laker_internal.h"

namespace laker {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int DevBuf::ensure(size_t want) {
  if (want <= bytes && ptr) return LAKER_OK;
  if (want == 0) return LAKER_OK;
  release();
  hipError_t e = hipMalloc(&ptr, want);
  if (e != hipSuccess) {
    ptr = nullptr;
    bytes = 0;
    (void)hipGetLastError();
    return fail(LAKER_ENOMEM, "device allocation of " + std::to_string(want) +
                                  " bytes failed: " + hipGetErrorString(e));
  }
  bytes = want;
  return LAKER_OK;
}

void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

// out[c] = sum_i X[i,c] * Y[i,c], one block per column chunk, FP64 accumulation.
template <class T>
__global__ void coldot_kernel(const T* __restrict__ X, const T* __restrict__ Y, int64_t n, int k,
                              int64_t ld, double* __restrict__ out) {
  __shared__ double sm[16];
  const int c = blockIdx.x;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += (double)X[i * ld + c] * (double)Y[i * ld + c];
  s = block_sum(s, sm);
  if (threadIdx.x == 0) out[c] = s;
}

}  // namespace laker

using namespace laker;

extern "C" {

const char* laker_last_error(void) { return g_last_error.c_str(); }

const char* laker_version(void) { return "laker-hip 0.4.0-mi355x"; }

int laker_coldot(const void* X, const void* Y, int64_t n, int32_t k, int64_t ld, int32_t dtype,
                 double* out, hipStream_t stream) {
  LAKER_CHECK_ARG(X && Y && out && n >= 1 && k >= 1 && ld >= k, "bad coldot arguments");
  if (dtype == LAKER_F32)
    hipLaunchKernelGGL((coldot_kernel<float>), dim3(k), dim3(1024), 0, stream, (const float*)X,
                       (const float*)Y, n, k, ld, out);
  else
    hipLaunchKernelGGL((coldot_kernel<double>), dim3(k), dim3(1024), 0, stream, (const double*)X,
                       (const double*)Y, n, k, ld, out);
  LAKER_HIP_CHECK(hipGetLastError());
  return LAKER_OK;
}

}  // extern "C"
