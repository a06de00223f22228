// Shared definitions for the LAKER HIP library (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <utility>

#include "../../include/laker_hip.h"

namespace laker {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef short shortx8 __attribute__((ext_vector_type(8)));

constexpr float kLog2e = 1.4426950408889634f;
constexpr double kLog2eD = 1.4426950408889634;

// Thread-local error text returned by laker_last_error().
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define LAKER_HIP_CHECK(expr)                                                          \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return ::laker::fail(LAKER_EHIP, std::string("HIP error: ") + hipGetErrorString(_e) + \
                                           " at " __FILE__ ":" + std::to_string(__LINE__)); \
  } while (0)

// Status of the kernel launch just issued, named for the error text.
#define LAKER_KCHECK(name)                                                               \
  do {                                                                                   \
    hipError_t _e = hipGetLastError();                                                   \
    if (_e != hipSuccess)                                                                \
      return ::laker::fail(LAKER_EHIP, std::string("HIP error launching ") + (name) + ": " + \
                                           hipGetErrorString(_e) + " at " __FILE__ ":" +  \
                                           std::to_string(__LINE__));                    \
  } while (0)

#define LAKER_CHECK_ARG(cond, msg)                                \
  do {                                                            \
    if (!(cond)) return ::laker::fail(LAKER_EINVAL, (msg));       \
  } while (0)

#define LAKER_TRY(expr)            \
  do {                             \
    int _rc = (expr);              \
    if (_rc != LAKER_OK) return _rc; \
  } while (0)

inline int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }
inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Round-to-nearest-even float -> bf16 bit pattern (finite inputs): the
// hardware conversion (v_cvt_pk_bf16_f32) in device code.
__device__ inline uint16_t f32_to_bf16_bits(float f) {
  return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f));
}
__device__ __host__ inline float bf16_bits_to_f32(uint16_t b) {
  return __builtin_bit_cast(float, (uint32_t)b << 16);
}

// Device memory of the library.  An unbound buffer (ensure(bytes) on an
// empty buffer) is plain hipMalloc / hipFree memory.  ensure(bytes, stream)
// on an empty buffer binds it to that stream: its block comes from, and on
// release goes back to, a cache of blocks that only the same stream reuses
// (stream order makes that safe without any host wait).  A bound buffer must
// be used on its stream only -- temporaries inside one library call; a buffer
// used on a second stream becomes unbound (plain hipFree on release).
// Handle-owned buffers stay unbound.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipStream_t tag = nullptr;
  bool tagged = false;
  int dev = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), bytes(o.bytes), tag(o.tag), tagged(o.tagged), dev(o.dev) {
    o.ptr = nullptr;
    o.bytes = 0;
    o.tagged = false;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      ptr = o.ptr;
      bytes = o.bytes;
      tag = o.tag;
      tagged = o.tagged;
      dev = o.dev;
      o.ptr = nullptr;
      o.bytes = 0;
      o.tagged = false;
    }
    return *this;
  }
  void swap(DevBuf& o) noexcept {
    std::swap(ptr, o.ptr);
    std::swap(bytes, o.bytes);
    std::swap(tag, o.tag);
    std::swap(tagged, o.tagged);
    std::swap(dev, o.dev);
  }
  int ensure(size_t want);
  int ensure(size_t want, hipStream_t st);
  // Marks an empty buffer as a temporary of stream st: later ensure() calls
  // without a stream (helpers handed the buffer) allocate it bound to st.
  void bind(hipStream_t st);
  void release();
  ~DevBuf() { release(); }
  template <class T> T* as() const { return reinterpret_cast<T*>(ptr); }
};

// Block-wide sum of doubles (blockDim.x a multiple of 64, <= 1024).
__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline float wave_sumf(float v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline double block_sum(double v, double* smem /* >= 16 doubles */) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) smem[wid] = v;
  __syncthreads();
  const int nw = blockDim.x >> 6;
  double t = (threadIdx.x < (unsigned)nw) ? smem[threadIdx.x] : 0.0;
  if (wid == 0) t = wave_sum(t);
  __syncthreads();
  if (threadIdx.x == 0) smem[0] = t;
  __syncthreads();
  double r = smem[0];
  __syncthreads();
  return r;
}

}  // namespace laker
