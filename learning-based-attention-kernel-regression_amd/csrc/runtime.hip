// Error reporting, device buffers and small column kernels of the LAKER HIP library.
#include "laker_common.h"
#include "laker_internal.h"

#include <mutex>
#include <vector>

namespace laker {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

// ---------------------------------------------------------------------------
// Stream-bound block cache behind DevBuf.  hipFree waits for the whole device
// and large hipMalloc calls take hundreds of microseconds, so a call that
// frees its temporaries would drain the GPU between kernels.  A released
// stream-bound block is kept instead and handed out again to the same stream
// (and device) only: stream order guarantees that the work of its previous
// user has finished by the time the next user's kernels run, so no host wait
// and no event is needed.  (Event queries and device synchronisation are
// avoided here on purpose: issued while another thread captures a graph on a
// stream, they invalidate that capture.)  At most kCacheCap bytes stay
// cached; beyond that the oldest blocks go back to the driver (hipFree waits
// for them), as do all cached blocks of a device when an allocation fails.
// Unbound buffers are plain hipMalloc / hipFree.
// ---------------------------------------------------------------------------
namespace {

struct CachedBlock {
  void* ptr;
  size_t bytes;
  int dev;
  hipStream_t st;
  uint64_t seq;  // release order
};

// Never destroyed: handles may be released during interpreter shutdown,
// after the library's static objects are gone.
struct CacheState {
  std::mutex mu;
  std::vector<CachedBlock> blocks;
  size_t bytes = 0;
  uint64_t seq = 0;
};
CacheState& cache_state() {
  static CacheState* s = new CacheState();
  return *s;
}
constexpr size_t kCacheCap = size_t(16) << 30;
constexpr size_t kBigGran = size_t(2) << 20;

size_t block_bytes(size_t want) {
  return want < (size_t(1) << 20) ? (size_t)round_up((int64_t)want, 512) : (size_t)round_up((int64_t)want, kBigGran);
}

int current_device() {
  int dev = 0;
  (void)hipGetDevice(&dev);
  return dev;
}

void erase_cached(size_t i) {
  CacheState& c = cache_state();
  c.bytes -= c.blocks[i].bytes;
  c.blocks[i] = c.blocks.back();
  c.blocks.pop_back();
}

// caller holds the lock; returns the cached blocks of dev (all when dev < 0) to the driver
void free_cached(int dev) {
  CacheState& c = cache_state();
  for (size_t i = 0; i < c.blocks.size();) {
    if (dev >= 0 && c.blocks[i].dev != dev) {
      ++i;
      continue;
    }
    (void)hipFree(c.blocks[i].ptr);
    erase_cached(i);
  }
}

void* cache_alloc(size_t want, int dev, hipStream_t st, size_t* got, hipError_t* err) {
  const size_t size = block_bytes(want);
  CacheState& c = cache_state();
  {
    std::lock_guard<std::mutex> lk(c.mu);
    int best = -1;
    for (size_t i = 0; i < c.blocks.size(); ++i) {
      const CachedBlock& b = c.blocks[i];
      if (b.dev != dev || b.st != st || b.bytes < size || b.bytes > size + size / 4 + kBigGran) continue;
      if (best < 0 || b.bytes < c.blocks[best].bytes) best = (int)i;
    }
    if (best >= 0) {
      void* p = c.blocks[best].ptr;
      *got = c.blocks[best].bytes;
      erase_cached((size_t)best);
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, size);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    {
      std::lock_guard<std::mutex> lk(c.mu);
      free_cached(dev);
    }
    e = hipMalloc(&p, size);
  }
  *err = e;
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  *got = size;
  return p;
}

void cache_give(void* p, size_t bytes, int dev, hipStream_t st) {
  CacheState& c = cache_state();
  std::lock_guard<std::mutex> lk(c.mu);
  c.blocks.push_back(CachedBlock{p, bytes, dev, st, ++c.seq});
  c.bytes += bytes;
  while (c.bytes > kCacheCap && !c.blocks.empty()) {
    size_t old = 0;
    for (size_t i = 1; i < c.blocks.size(); ++i)
      if (c.blocks[i].seq < c.blocks[old].seq) old = i;
    (void)hipFree(c.blocks[old].ptr);
    erase_cached(old);
  }
}

}  // namespace

int DevBuf::ensure(size_t want) {
  if (tagged) return ensure(want, tag);  // a bound buffer stays on its stream
  if (want <= bytes && ptr) return LAKER_OK;
  if (want == 0) return LAKER_OK;
  release();
  hipError_t e = hipMalloc(&ptr, want);
  if (e != hipSuccess) {
    ptr = nullptr;
    (void)hipGetLastError();
    {
      std::lock_guard<std::mutex> lk(cache_state().mu);
      free_cached(current_device());
    }
    e = hipMalloc(&ptr, want);
  }
  if (e != hipSuccess) {
    ptr = nullptr;
    bytes = 0;
    (void)hipGetLastError();
    return fail(LAKER_ENOMEM, "device allocation of " + std::to_string(want) +
                                  " bytes failed: " + hipGetErrorString(e));
  }
  bytes = want;
  dev = current_device();
  return LAKER_OK;
}

int DevBuf::ensure(size_t want, hipStream_t st) {
  if (ptr && tagged && tag != st) tagged = false;  // used on two streams: unbound from now on
  if (want <= bytes && ptr) return LAKER_OK;
  if (want == 0) return LAKER_OK;
  if (ptr && !tagged) return ensure(want);  // an allocated unbound buffer stays unbound
  release();
  const int d = current_device();
  size_t got = 0;
  hipError_t e = hipSuccess;
  void* p = cache_alloc(want, d, st, &got, &e);
  if (!p)
    return fail(LAKER_ENOMEM, "device allocation of " + std::to_string(want) +
                                  " bytes failed: " + hipGetErrorString(e));
  ptr = p;
  bytes = got;
  dev = d;
  tag = st;
  tagged = true;
  return LAKER_OK;
}

void DevBuf::bind(hipStream_t st) {
  if (!ptr) {
    tag = st;
    tagged = true;
  }
}

void DevBuf::release() {
  if (!ptr) return;
  if (tagged) cache_give(ptr, bytes, dev, tag);
  else (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
  tagged = false;
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
