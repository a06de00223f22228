// Internal types shared by the translation units of the LAKER HIP library.
#pragma once
#include <mutex>

#include "laker_common.h"

namespace laker {

// Non-owning view of one side (rows or columns) of a fused product.
struct Side {
  int64_t n = 0;
  int64_t npad = 0;                // rows padded to a multiple of 64, zero-filled
  const void* pack32 = nullptr;    // FP32 MFMA operand rows (ka32 floats)
  const void* pack16 = nullptr;    // BF16 hi/lo MFMA operand rows (ka16 bf16)
  const double* e64 = nullptr;     // FP64 copy, n x d
};

// Owning storage of a packed embedding set.
struct SideStore {
  int64_t n = 0, npad = 0;
  DevBuf col32, row32, col16, row16, e64, user;
  DevBuf col16h, row16h;  // FP16 2-way split operands (square operator only, ka16h halves)
};

struct Operator {
  int64_t n = 0;
  int d = 0;
  int dtype = LAKER_F32;
  double lambda_reg = 0.0;
  double scale = 1.0;
  int clamp_policy = LAKER_CLAMP_AUTO;
  double clamp_limit = 80.0;
  double max_sq_norm = 0.0;
  bool skip_clamp = true;
  int ka32 = 0, ka16 = 0;
  int ka16h = 0;         // round_up(3d, 16): FP16 2-way split operand width
  bool f16s_ok = false;  // the scaled embedding fits FP16: k = 1 square products use the FP16 split
  SideStore self;
  SideStore query;  // packed query set of the last cross product
  std::mutex query_mu;
  DevBuf part;      // split-K partial sums of the FP32 / split-precision products
  // Partial sums of the FP64 products (refinement residual, FP64 operators).
  // Kept apart from `part` so that an FP64 residual between replays of a
  // captured graph of FP32 products never reallocates a buffer the graph
  // still points to.
  DevBuf part64;
  DevBuf scratch;   // backward right-hand sides
  DevBuf pcg_ws;    // PCG vectors, partial sums and device scalars
  DevBuf mixed_ws;  // refinement residual and FP32 correction vectors
  DevBuf xpack;     // multi-RHS right-hand sides packed as BF16 MFMA operands
  DevBuf sym_map;   // (row block, segment) of each workgroup of the symmetric product
  int sym_nwg = 0;
  bool sym_enabled = true;  // symmetric k = 1 product for the square operator
  DevBuf part_c;    // column partial sums of the symmetric product
  DevBuf sym64_map; // (row block, segment) of each workgroup of the symmetric FP64 product
  int sym64_nwg = 0;
  DevBuf part_c64;  // column partial sums of the symmetric FP64 product

  int ensure_part(size_t bytes) { return part.ensure(bytes); }
  bool clamp_for(int policy) const {
    if (policy == LAKER_CLAMP_ALWAYS) return true;
    if (policy == LAKER_CLAMP_NEVER) return false;
    if (clamp_policy == LAKER_CLAMP_ALWAYS) return true;
    if (clamp_policy == LAKER_CLAMP_NEVER) return false;
    return !skip_clamp;
  }
  Side rows() const { return view(self, true); }
  Side cols() const { return view(self, false); }
  static Side view(const SideStore& s, bool as_rows) {
    Side v;
    v.n = s.n;
    v.npad = s.npad;
    v.pack32 = as_rows ? s.row32.ptr : s.col32.ptr;
    v.pack16 = as_rows ? s.row16.ptr : s.col16.ptr;
    v.e64 = s.e64.as<double>();
    return v;
  }
  const float* e_user32() const { return self.user.as<float>(); }
};

// Pack an embedding set (row side pre-scaled by scale*log2(e)).
int pack_f16_split(SideStore& s, const void* E, int dtype, int64_t n, int d, int ka, double scale, hipStream_t st);
int pack_store(SideStore& s, const void* E, int dtype, int64_t n, int d, int ka32, int ka16,
               double scale, bool need_col, bool need_row, hipStream_t st);

// Y (rows.n x k) = [lam * Xr] + exp(min(s*Erow.Ecol^T, c)) X   (FP32 vectors)
// deriv: use the backward weights (0 where the clamp is active).
// accumulate: add into Y instead of overwriting.
// skip: optional device flag; when it reads non-zero at launch time the k <= 4
// kernels return at once and Y is left untouched (a finished solve's trailing
// launches cost nothing).
// dots3 (optional, k = 1): when the square operator's symmetric kernel runs,
// its reduction also writes the per-256-row-block partials of Xr.Y, Xr.Xr and
// Y.Y to dots3[q * ceil(n / 256) + block] and sets *dots3_done.
// Internal product precision (not part of the C ABI): the exp tile split hi/lo
// times an X whose entries are exactly representable in BF16 (two BF16
// products, hi.X + lo.X; the lo part of X would be zero).
constexpr int PREC_HILO_XBF16 = 7;

int fused_product_f32(Operator& op, int precision, const Side& rows, const Side& cols,
                      const float* X, int k, int64_t ldx, float* Y, int64_t ldy, double lam,
                      const float* Xr, int64_t ldxr, bool clamp, bool deriv, int accumulate,
                      hipStream_t st, const int* skip = nullptr, double* dots3 = nullptr,
                      bool* dots3_done = nullptr);

// FP64 product; when B != nullptr writes the residual Y = B - (lam Xr + G X).
template <class XT>
int fused_product_f64(Operator& op, const Side& rows, const Side& cols, const XT* X, int k,
                      int64_t ldx, double* Y, int64_t ldy, double lam, const XT* Xr, int64_t ldxr,
                      bool clamp, const double* B, hipStream_t st);

}  // namespace laker
