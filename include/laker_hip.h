/*
 * laker_hip.h -- C ABI of the MI355X-native LAKER solver core.
 *
 * Plain pointers, sizes and hipStream_t only.  All device pointers are
 * caller-owned except the opaque handles, whose internal device buffers
 * (packed embeddings, preconditioner factors, Krylov workspaces) are
 * allocated by the library at create time and released by *_destroy.
 * Every call is stream-ordered on the given stream; calls on one handle must
 * not overlap.  Return value: LAKER_OK or a negative LAKER_E* code, with the
 * message available from laker_last_error() (thread-local).
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference repository root).
 */
#ifndef LAKER_HIP_H
#define LAKER_HIP_H

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (Python maps EINVAL->ValueError, the rest->RuntimeError) */
#define LAKER_OK 0
#define LAKER_EINVAL -1     /* bad shape / argument */
#define LAKER_EHIP -2       /* HIP runtime error */
#define LAKER_ENOMEM -3     /* device allocation failed */
#define LAKER_EBREAKDOWN -4 /* PCG non-positive curvature (laker/solvers.py:232-246) */
#define LAKER_ESTATE -5     /* object not built (laker/preconditioner.py:303-304) */
#define LAKER_ENUMERIC -6   /* factorisation failure (non-SPD small matrix) */

/* ---- element types */
#define LAKER_F32 0
#define LAKER_F64 1

/* ---- matvec arithmetic */
#define LAKER_PREC_FP32 0  /* Gram on FP32 MFMA, FP32 exp and products */
#define LAKER_PREC_MIXED 1 /* Gram on BF16 MFMA from a 3-way bf16 split of E (6 products, ~2^-24); the
                              square operator's k = 1 product instead uses a 2-way FP16 split (3 products,
                              ~2^-22) while max |e| * max(1, scale log2 e) < 2^14; for k > 4 the exp tile
                              times X also on BF16 MFMA from 3-way splits (6 products, ~2^-24) */
#define LAKER_PREC_FP64 2  /* Gram, exp and accumulation in FP64 (refinement residual) */
#define LAKER_PREC_BF16 3  /* split Gram; for k > 4 one BF16 product of the exp tile and X (~2^-8) */
#define LAKER_PREC_HILO 4  /* split Gram; for k > 4 hi/lo BF16 products (3 products, ~2^-16):
                              the probe matvec of the mixed CCCP build */

/* ---- clamp policy for exp(min(s*g, 80|700)) (laker/kernels.py:80-113,181-185) */
#define LAKER_CLAMP_AUTO 0   /* clamp unless max_i ||e_i||^2 * scale < 80 (fp32) / 700 (fp64) */
#define LAKER_CLAMP_ALWAYS 1 /* always clamp (tiled predict, laker/core.py:653) */
#define LAKER_CLAMP_NEVER 2  /* never clamp (to_dense, laker/kernels.py:296-298) */

typedef struct laker_operator_s* laker_operator_t;
typedef struct laker_precond_s* laker_precond_t;

/* Text of the last error on this thread ("" if none). */
const char* laker_last_error(void);
/* Library version string. */
const char* laker_version(void);

/* ---------------------------------------------------------------------------
 * Exact attention-kernel operator  A = lambda I + exp(scale * E E^T)
 * Replaces AttentionKernelOperator.__init__ (laker/kernels.py:154-185).
 * E: n x d row-major (ld = d), dtype LAKER_F32 or LAKER_F64, device memory.
 * The handle copies and packs E (FP32 MFMA operands, BF16 3-way split and
 * FP16 2-way split operands, FP64 copy) and evaluates the skip-clamp
 * predicate on device.
 */
int laker_operator_create(const void* E, int64_t n, int32_t d, int32_t dtype, double lambda_reg,
                          double scale, int32_t clamp_policy, hipStream_t stream,
                          laker_operator_t* out);
int laker_operator_destroy(laker_operator_t op);
/* skip_clamp flag (laker/kernels.py:183-185) and max squared row norm. */
int laker_operator_info(laker_operator_t op, int32_t* skip_clamp, double* max_sq_norm);
/* Change lambda without repacking (regularisation path / bilevel use). */
int laker_operator_set_lambda(laker_operator_t op, double lambda_reg);

/* Y = (lambda I + G) X.  X, Y: n x k, column stride... row-major with leading
 * dimension ldx/ldy (elements), dtype = operator dtype.
 * Replaces AttentionKernelOperator.matvec / matvec_impl (laker/kernels.py:187-270)
 * and dense_attention_matvec (laker/kernels.py:121-130). */
int laker_matvec(laker_operator_t op, const void* X, int32_t k, int64_t ldx, void* Y, int64_t ldy,
                 int32_t precision, hipStream_t stream);

/* Allocation-free Y = (lambda I + G) X for an LAKER_F32 operator, 1 <= k <= 4
 * and precision LAKER_PREC_FP32 or LAKER_PREC_MIXED: the same result as
 * laker_matvec to FP32 rounding (split-K row products instead of the
 * symmetric k = 1 kernel).  No reference counterpart: it is for callers that
 * cannot allocate device memory during the call (HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It calls no hipMalloc,
 * hipFree or allocating routine; its partial sums live in a 16 MiB device
 * array that the library holds on every GPU it runs on until the process
 * ends, so n <= 262144 for k <= 2 and n <= 131072 for k = 3, 4 (LAKER_EINVAL
 * beyond).  All calls on one GPU share that array: calls that overlap on one
 * GPU (concurrent streams or threads) race and must be serialised by the
 * caller. */
int laker_matvec_noalloc(laker_operator_t op, const float* X, int32_t k, int64_t ldx, float* Y, int64_t ldy,
                         int32_t precision, hipStream_t stream);

/* R = B - (lambda I + G) X with FP64 X/B/R and FP64 Gram/exp/accumulation:
 * the iterative-refinement residual (no reference counterpart; it checks the
 * reference's stopping quantity ||b - A x|| / ||b||, laker/solvers.py:266-267). */
int laker_residual_fp64(laker_operator_t op, const double* B, const double* X, int32_t k,
                        double* R, hipStream_t stream);

/* Y = exp(min(scale * Eq E^T, clamp)) X for m query rows (no lambda term).
 * Replaces the predict path kernel_eval(q, E) @ alpha (laker/core.py:580-661,
 * laker/kernels.py:301-354) without materialising m x n.
 * clamp_policy LAKER_CLAMP_AUTO uses the operator's skip-clamp flag.
 * Eq has the operator's dtype.  X and Y have the operator's dtype, except
 * with precision LAKER_PREC_FP64 on an LAKER_F32 operator: then X and Y are
 * FP64 and the product runs on the FP64 copy of E in FP64 (predictions from
 * the FP64 solution of laker_solve_mixed), with the FP32 clamp limit. */
int laker_cross_matvec(laker_operator_t op, const void* Eq, int64_t m, const void* X, int32_t k,
                       int64_t ldx, void* Y, int64_t ldy, int32_t clamp_policy, int32_t precision,
                       hipStream_t stream);

/* Transposed cross product Y = exp(min(scale * E Eq^T, clamp)) Xq  (n x k),
 * Xq: m x k.  Used by the hand-written backward of predict and the batched
 * variance right-hand sides. */
int laker_cross_matvec_t(laker_operator_t op, const void* Eq, int64_t m, const void* Xq, int32_t k,
                         int64_t ldx, void* Y, int64_t ldy, int32_t clamp_policy,
                         hipStream_t stream);

/* K = exp(min(scale * X Y^T, clamp)) materialised (m x p, row-major, ldk).
 * Replaces AttentionKernelOperator.kernel_eval (laker/kernels.py:301-354). */
int laker_kernel_eval(const void* X, int64_t m, const void* Y, int64_t p, int32_t d, int32_t dtype,
                      double scale, int32_t clamp, void* K, int64_t ldk, hipStream_t stream);

/* out[i] = lambda + exp(scale * ||e_i||^2)  (no clamp).
 * Replaces AttentionKernelOperator.diagonal (laker/kernels.py:272-283). */
int laker_diagonal(laker_operator_t op, void* out, hipStream_t stream);

/* Dense lambda I + exp(scale E E^T) (no clamp), n x n row-major.
 * Replaces AttentionKernelOperator.to_dense (laker/kernels.py:285-299). */
int laker_to_dense(laker_operator_t op, void* out, hipStream_t stream);

/* Hand-written backward of Y = lambda X + G X with G = exp(min(s E E^T, c)):
 * given dY (n x k), writes dX = lambda dY + G dY and
 * dE_i = s * sum_j G'_ij (dY_i . X_j + X_i . dY_j) E_j  (G' = 0 where clamped).
 * Tiles are recomputed, nothing n x n is stored.  Replaces autograd through
 * exp_safe's out-of-place branch (laker/kernels.py:100-101,111-112) used by
 * laker/training.py:186-211.  dX or dE may be NULL to skip. */
int laker_matvec_backward(laker_operator_t op, const float* X, const float* dY, int32_t k,
                          float* dX, float* dE, hipStream_t stream);

/* Backward of Yq = exp(min(s Eq E^T, c)) X  (predict_train, laker/core.py:782-849):
 * dEq (m x d), dE (n x d) accumulated (+=), dX (n x k) written.  Any may be NULL. */
int laker_cross_matvec_backward(laker_operator_t op, const float* Eq, int64_t m, const float* X,
                                int32_t k, const float* dYq, float* dEq, float* dE, float* dX,
                                int32_t clamp_policy, hipStream_t stream);

/* ---------------------------------------------------------------------------
 * CCCP preconditioner  P = Sigma^{-1/2},  Sigma = a I + Q C Q^T
 * Replaces CCCPPreconditioner.build (laker/preconditioner.py:144-284).
 * probes: n x nr row-major in the operator dtype (device), or NULL to draw
 * N(0,1) probes on device from `seed`.  probe_precision selects the matvec
 * used for the probe responses (LAKER_PREC_FP32, _MIXED, _HILO or _BF16).
 * The QR is a shifted Cholesky-QR3 in FP64; the N_r x N_r CCCP loop, its
 * Cholesky factorisations and the final inverse square root run in FP64 on
 * device.  *iterations receives the CCCP iteration count. */
int laker_precond_build(laker_operator_t op, int32_t nr, double gamma, double epsilon,
                        double base_rho, int32_t max_iter, double tol, const void* probes,
                        uint64_t seed, int32_t probe_precision, hipStream_t stream,
                        laker_precond_t* out, int32_t* iterations);
/* Same build from probe responses U = A R computed by the caller (n x nr,
 * dtype LAKER_F32/LAKER_F64, device): the path for operators that are Python
 * callables (CCCPPreconditioner.build(operator, n) with a generic operator,
 * laker/preconditioner.py:186-284). */
int laker_precond_build_from_responses(const void* U, int64_t n, int32_t nr, int32_t dtype,
                                       double gamma, double epsilon, double base_rho, int32_t max_iter,
                                       double tol, hipStream_t stream, laker_precond_t* out,
                                       int32_t* iterations);
int laker_precond_destroy(laker_precond_t pc);
/* Copy a factor into caller memory: which = 0 Q (n x nr, op dtype), 1 R, 2 C,
 * 3 S (nr x nr FP64).  Backs the q_basis / qr_r / q_basis_matrix attributes
 * (laker/preconditioner.py:137-142). */
int laker_precond_copy_out(laker_precond_t pc, int32_t which, void* dst, hipStream_t stream);
/* isotropic_coef a, N_r, shrinkage rho; device pointers to Q (n x nr, op dtype),
 * R of the QR (nr x nr FP64), C (nr x nr FP64) and S = M^{-1/2} - a^{-1/2} I (FP64). */
int laker_precond_info(laker_precond_t pc, double* isotropic_coef, int32_t* nr, double* rho,
                       const void** q_basis, const double** qr_r, const double** c_matrix,
                       const double** s_matrix);
/* Z = P X.  Replaces CCCPPreconditioner.apply / preconditioner_apply_core
 * (laker/preconditioner.py:52-68,286-352). */
int laker_precond_apply(laker_precond_t pc, const void* X, int32_t k, int64_t ldx, void* Z,
                        int64_t ldz, hipStream_t stream);
/* Jacobi preconditioner P = diag(lambda I + G)^{-1}: the handle stores
 * inv = 1 / max(diag, eps(dtype)) computed in dtype (LAKER_F32/F64) from the
 * device diagonal (n entries), and laker_precond_apply then returns
 * Z[i][c] = inv[i] * X[i][c] for every column, so it also drives
 * laker_pcg_solve and laker_solve_mixed.  copy_out(which = 4) copies inv out;
 * laker_precond_info rejects it.  Replaces JacobiPreconditioner.__init__ /
 * apply (laker/solvers.py:505-546). */
int laker_precond_create_jacobi(const void* diagonal, int64_t n, int32_t dtype, hipStream_t stream,
                                laker_precond_t* out);

/* A CCCP handle from stored factors, without the probe matvec or the CCCP
 * loop: Q (n x nr, dtype), R, C and S = M^{-1/2} - a^{-1/2} I (nr x nr FP64,
 * as laker_precond_copy_out returns them), isotropic_coef a > 0 and rho; all
 * device pointers, copied into the handle.  The reference's save format holds
 * no preconditioner (laker/persistence.py:74-328); this backs the compatible
 * extension that stores the factors, so a loaded model solves at once. */
int laker_precond_from_factors(const void* Q, int64_t n, int32_t nr, int32_t dtype, double isotropic_coef,
                               double rho, const double* R, const double* C, const double* S,
                               hipStream_t stream, laker_precond_t* out);

/* In-place orthonormal basis of the n x k block X (row-major, ld = k, device,
 * LAKER_F32/F64, n >= k): columns scaled to unit norm, then Cholesky-QR2 in
 * FP64 (shifted Cholesky-QR3 if the first Gram factorisation fails).  The
 * result equals the Q of a Householder QR up to column signs.  LAKER_ENUMERIC
 * for a numerically rank-deficient block.  Replaces the torch.linalg.qr of
 * the "power_iter" probe strategy (laker/preconditioner.py:174-184). */
int laker_orthonormalize(void* X, int64_t n, int32_t k, int32_t dtype, hipStream_t stream);

/* count N(0,1) draws (LAKER_F32/F64) into device memory: the same stream
 * laker_precond_build draws when it is given no probe matrix and its
 * probe_precision is LAKER_PREC_FP32, _MIXED or _BF16, so a caller can
 * transform those probes first (the "power_iter" strategy) and pass them
 * back.  With LAKER_PREC_HILO on an FP32 operator the build rounds each of
 * these draws to the nearest BF16 value before use (its probe product then
 * takes two BF16 products instead of three); probes passed in by the caller
 * are used as given.  Replaces the device torch.randn of
 * laker/preconditioner.py:168-171 for probe_source="device" (a different
 * generator than torch's). */
int laker_randn(void* out, int64_t count, int32_t dtype, uint64_t seed, hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Device-resident PCG.  Replaces PreconditionedConjugateGradient.solve /
 * solve_1d / solve_2d (laker/solvers.py:130-394) when both the operator and
 * the preconditioner are native.  B, X: n x k (op dtype).  X0 may be NULL.
 * k == 1 follows solve_1d, k > 1 follows solve_2d (per-column alpha/beta,
 * global Frobenius stopping ratio, breakdown if any column).
 * check_every: host polls the device status every that many iterations. */
typedef struct laker_pcg_result {
  int32_t iterations;    /* reference `iterations` */
  int32_t status;        /* LAKER_OK, LAKER_EBREAKDOWN */
  int32_t converged;     /* 1 if ||r||/||b|| <= tol was reached */
  int32_t zero_rhs;      /* ||b|| == 0 early return (laker/solvers.py:181-182) */
  double residual_norm;  /* reference `residual_norm` (recurrence residual) */
  double rhs_norm;
} laker_pcg_result;

int laker_pcg_solve(laker_operator_t op, laker_precond_t pc, const void* B, int32_t k, void* X,
                    const void* X0, double tol, int32_t max_iter, double breakdown_eps,
                    int32_t restart_freq, int32_t precision, int32_t check_every,
                    hipStream_t stream, laker_pcg_result* result);

/* Mixed-precision solve (iterative refinement): each outer step runs PCG
 * from a fresh start on the FP64 residual B - A X scaled to unit norm, with
 * operator products on split MFMA Gram tiles (LAKER_PREC_MIXED, or FP32
 * MFMA with LAKER_PREC_FP32), FP32 vectors and FP64 scalars, and adds its
 * correction straight into the FP64 X.  A step stops when its recursive
 * residual has dropped by inner_tol (or by what the outer target still
 * needs).  B, X: n x k FP64.  Stops when ||B - A X||_F / ||B||_F <= tol in
 * FP64, after at most max_outer FP64 residuals or inner_max_iter iterations
 * in total.  Non-positive curvature inside a step ends that step at its last
 * good iterate (the reduced-precision operator can be slightly indefinite when
 * lambda is below its rounding error); only a step that makes no progress at
 * all returns LAKER_EBREAKDOWN.  The best FP64 iterate is kept: the first
 * step that does not lower the FP64 residual is undone and followed by one
 * last step run to the full target (refinement cannot contract when lambda is
 * below the reduced-precision operator's rounding error); converged = 0 then
 * reports that the target was not reached.  No reference counterpart (new
 * precision mode behind the same solve call). */
typedef struct laker_mixed_result {
  int32_t outer_iterations; /* refinement steps (FP64 residuals) */
  int32_t inner_iterations; /* total PCG iterations */
  int32_t status;
  int32_t converged;
  double residual_norm;     /* FP64 true residual ||B - A X||_F */
  double rhs_norm;
} laker_mixed_result;

int laker_solve_mixed(laker_operator_t op, laker_precond_t pc, const double* B, int32_t k,
                      double* X, double tol, int32_t max_outer, double inner_tol,
                      int32_t inner_max_iter, int32_t inner_precision, hipStream_t stream,
                      laker_mixed_result* result);

/* ---------------------------------------------------------------------------
 * Batched mixed-precision solves: a batch runs the inner PCG iterations of
 * many small systems -- uneven n, 2 <= k <= 32 right-hand sides each, an FP32
 * operator and an FP32 CCCP handle each -- as shared kernel launches (one
 * launch per kernel of the iteration for all systems whose refinement step is
 * running), and takes each system's refinement steps (FP64 residual, best
 * iterate, next step) on the host between polls, with the arithmetic and
 * stopping rules of laker_solve_mixed.  A system's iterates do not depend on
 * which other systems share the batch.  Systems join with laker_batch_add
 * (their set-up runs at once) whenever a slot is free, and leave when
 * laker_batch_result collects them; the operator, preconditioner, B (n x k
 * FP64) and X (n x k FP64) must stay valid until then.  All work is ordered
 * on the batch's stream; one host thread drives a batch.  No reference
 * counterpart: the serving loop's continuous batching
 * (the reference solves one system per call, laker/models.py:392-455). */
typedef struct laker_batch_s* laker_batch_t;
int laker_batch_create(int32_t max_systems, hipStream_t stream, laker_batch_t* out);
int laker_batch_destroy(laker_batch_t batch);
int laker_batch_add(laker_batch_t batch, laker_operator_t op, laker_precond_t pc, const double* B, int32_t k,
                    double* X, double tol, int32_t max_outer, double inner_tol, int32_t inner_max_iter,
                    int32_t* slot);
/* Runs `iterations` inner iterations of every running system, then the
 * refinement bookkeeping of the systems whose step ended; writes the slots
 * whose solve has finished (results ready) to finished[0 .. *nfinished) and
 * the number of systems still solving to *nactive. */
int laker_batch_step(laker_batch_t batch, int32_t iterations, int32_t* finished, int32_t* nfinished,
                     int32_t* nactive);
/* Result of a finished slot; frees the slot.  LAKER_EBREAKDOWN as
 * laker_solve_mixed returns it. */
int laker_batch_result(laker_batch_t batch, int32_t slot, laker_mixed_result* result);

/* HIP-graph replay of the device PCG iterations (laker_pcg_solve,
 * laker_solve_mixed): after a first eager batch, each batch of polled
 * iterations is one captured graph launch (iteration state lives on the
 * device, so results are identical with and without graphs).  enable = 1 (the
 * default) or 0; process-wide.  No reference counterpart: a launch-overhead
 * option of the serving path. */
int laker_set_graphs(int32_t enable);

/* ---------------------------------------------------------------------------
 * Column-wise vector kernels for the generic PCG path, where operator or
 * preconditioner are Python callables (laker/solvers.py:225-286, 334-385).
 * out[c] (FP64, device) = sum_i X[i,c] * Y[i,c] for c < k; X, Y n x k
 * row-major with leading dimension ld, dtype LAKER_F32/LAKER_F64. */
int laker_coldot(const void* X, const void* Y, int64_t n, int32_t k, int64_t ld, int32_t dtype,
                 double* out, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* LAKER_HIP_H */
