/*
 * mfla_hip.h -- C ABI of the MI355X (gfx950) Float64xN dense linear-algebra
 * library, a drop-in device backend for MultiFloatLinearAlgebra.jl (MFLA).
 *
 * Each entry point replaces one MFLA generic function for device-resident
 * Float64xN (N = 1..4) matrices; the reference function is cited beside it.
 * A Julia wrapper binds these with ccall (one `ccall((:mfla_gemm, lib), ...)`
 * per entry point, matrices passed as mfla_matrix by value); the Python
 * package multifloatlinearalgebra.jl_amd/mfla binds them with ctypes.
 *
 * Data model. A Float64xN matrix is LIMB-PLANAR: N planes of binary64, each a
 * column-major rows x cols matrix with leading dimension `ld`; plane l starts
 * `limb_stride` elements after plane 0. Element (i, j) (zero-based) limb l is
 * data[l*limb_stride + j*ld + i]. All pointers in mfla_matrix are device
 * pointers. Scalars (alpha, beta) are host arrays of N binary64 limbs.
 *
 * Conventions (MFLA's):
 *  - Status codes follow MFLA: 0 success, -1 non-finite input, k > 0 the
 *    one-based column/pivot where the factorization stopped. They are returned
 *    through `int64_t *info` (host pointer). Pivot vectors are one-based int64.
 *  - Return value: MFLA_OK or a negative error mirroring the Julia exception
 *    the reference throws before touching memory: MFLA_ERR_DIMENSION
 *    (DimensionMismatch), MFLA_ERR_ARGUMENT (ArgumentError: bad symbol,
 *    unsupported N), MFLA_ERR_ALIAS (ArgumentError: output aliases an input),
 *    MFLA_ERR_FAILED_FACTOR (PosDefException / SingularException on solve).
 *  - beta == 0 means the output is not read.
 *  - Calls are stream-ordered on the handle's stream. Factorizations and
 *    backward_error synchronise the stream before returning (they report
 *    host-visible status); the BLAS-like calls are asynchronous.
 *  - Every reduction runs in ascending index order with a fixed
 *    multi-double network: results are deterministic and independent of
 *    the launch configuration.
 */
#ifndef MFLA_HIP_H
#define MFLA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFLA_OK 0
#define MFLA_ERR_ARGUMENT (-1)
#define MFLA_ERR_DIMENSION (-2)
#define MFLA_ERR_ALIAS (-3)
#define MFLA_ERR_HIP (-4)
#define MFLA_ERR_FAILED_FACTOR (-5)
#define MFLA_ERR_WORKSPACE (-6)
#define MFLA_ERR_BOUNDS (-7)      /* BoundsError (a reduction range outside the panel) */
/* Positive: the call completed and its outputs are valid, but a stated goal
 * was not met (mfla_refine_solve: the backward-error estimate of the best
 * iterate, which X holds, is above tol). */
#define MFLA_NOT_CONVERGED 1

/* side / uplo / trans / diag symbols (MFLA's :left/:right, :lower/:upper,
 * :N/:T, :nonunit/:unit). MFLA_GENERAL is residual!'s uplo=:general. */
#define MFLA_LEFT 0
#define MFLA_RIGHT 1
#define MFLA_LOWER 0
#define MFLA_UPPER 1
#define MFLA_GENERAL 2
#define MFLA_NOTRANS 0
#define MFLA_TRANS 1
#define MFLA_NONUNIT 0
#define MFLA_UNIT 1

/* gemm_strategy (src/kernels/gemm.jl:27-34) */
#define MFLA_GEMM_AUTO 0
#define MFLA_GEMM_DIRECT 1
#define MFLA_GEMM_PACKED 2
#define MFLA_GEMM_FUSED 3
/* Ozaki-split product on the INT8 matrix cores (no MFLA counterpart): both
 * operands sliced into 7-bit digits, digit products summed exactly in int32,
 * recombined in Float64xN. Normwise accuracy ~ K * 2^-(52N) * max|A(i,:)|
 * * max|B(:,j)| per element (not MFLA's rounding order: parity is within
 * that tolerance, not bitwise). Reduction length <= 8192; gemm! only. */
#define MFLA_GEMM_OZAKI 4
/* ldlt_strategy (src/factorizations/ldlt.jl:329-350) */
#define MFLA_LDLT_AUTO 0
#define MFLA_LDLT_UNBLOCKED 1
#define MFLA_LDLT_BLOCKED 2
/* backward-substitution order */
#define MFLA_TRSM_ORDER_AUTO 0
#define MFLA_TRSM_ORDER_EXACT 1
#define MFLA_TRSM_ORDER_BLOCKED 2

typedef struct mfla_matrix {
    double *data;        /* device pointer: limb 0 of element (0, 0) */
    int64_t rows;
    int64_t cols;
    int64_t ld;          /* column stride within a limb plane (elements) */
    int64_t limb_stride; /* distance between limb planes (elements) */
} mfla_matrix;

/* KernelConfig (src/types.jl:13-27). CPU-only fields (reduction_tile,
 * thread_count, gemm_panel_columns, gemm_micro_columns) are accepted and
 * ignored. gemm_tile selects the GEMM tile shape (-1 = auto). */
typedef struct mfla_config {
    int64_t reduction_tile;
    int64_t column_tile;
    int64_t cholesky_block;
    int64_t lu_block;
    int64_t ldlt_block;            /* 0: MFLA default 16/12/8 by limb count */
    int32_t ldlt_strategy;         /* MFLA_LDLT_* */
    int32_t gemm_strategy;         /* MFLA_GEMM_* */
    int64_t ldlt_blocked_crossover;
    int64_t thread_count;
    int64_t gemm_panel_columns;
    int64_t gemm_micro_columns;
    int32_t trsm_order;            /* MFLA_TRSM_ORDER_*; auto = exact for n <= 256 */
    int32_t gemm_tile;             /* -1 auto (16x16 for lower-triangle outputs), 0: 16x16, 1: 32x32, 2: 32x64, 3: 64x64 (1-2 limbs; else 32x64) output tiles; values outside 0..3 mean 16x16 */
    int64_t gemm_packed_crossover; /* KernelConfig.gemm_packed_crossover (INT64_MAX by default): only gemm_plan reads it */
} mfla_config;

typedef struct mfla_handle_s *mfla_handle;
typedef struct mfla_workspace_s *mfla_workspace;

/* ----------------------------------------------------------- library */
int mfla_version(void);
const char *mfla_last_error(void);                 /* thread-local message */
void mfla_config_default(mfla_config *config);     /* KernelConfig() */

int mfla_handle_create(mfla_handle *out);           /* current device, default stream */
int mfla_handle_destroy(mfla_handle h);
int mfla_handle_set_stream(mfla_handle h, void *hip_stream);
int mfla_synchronize(mfla_handle h);

/* MFWorkspace (src/types.jl:71-123): caller-owned device scratch for LU/LDLT
 * metadata and the blocked-LDLT weighted panel. */
int mfla_workspace_create(mfla_handle h, int limbs, int64_t factor_capacity, int64_t ldlt_block_capacity, mfla_workspace *out);
int mfla_workspace_destroy(mfla_workspace ws);
/* workspace_capacity (types.jl:131-145): out[0] factor, out[1] ldlt_block */
int mfla_workspace_capacity(mfla_workspace ws, int64_t *out2);
/* ensure_workspace_capacity! (types.jl:157-192): grows, never shrinks */
int mfla_workspace_ensure(mfla_workspace ws, int64_t factor_capacity, int64_t ldlt_block_capacity);

/* Elementwise Float64xN networks over planar vectors (limb stride = count):
 * op 0 x+y, 1 x-y, 2 x*y, 3 x/y, 4 sqrt(x), 5 z+x*y, 6 mulacc_x3(z, x, y)
 * (src/kernels/mulacc_x3.jl:61-66, limbs == 3 only), 7 inv(x), 8 z-x*y,
 * 9 abs(x), 10 max(x, y), 11 min(x, y), 12 (x < y). mfla_arith runs the
 * device networks on device pointers; mfla_host_arith runs the same networks
 * on the host (used for O(n) factor diagnostics). These are MultiFloats.jl's
 * scalar operators (+ - * / sqrt inv abs max min <) for the same element
 * types; they have no single MFLA counterpart. */
int mfla_arith(mfla_handle h, int limbs, int op, int64_t count, const double *x, const double *y, const double *z, double *out);
int mfla_host_arith(int limbs, int op, int64_t count, const double *x, const double *y, const double *z, double *out);

/* max |A| over a region (0 all, 1 lower i >= j, 2 upper i <= j) into the host
 * limbs out[0..limbs-1] (_maximum_abs / _lower_maximum_abs, src/types.jl:317-331;
 * the U scan of src/diagnostics.jl:74-81), and the finiteness scans
 * _all_finite / _lower_triangle_finite (src/types.jl:298-315). */
int mfla_absmax(mfla_handle h, int limbs, int region, mfla_matrix A, double *out);
int mfla_all_finite(mfla_handle h, int limbs, int region, mfla_matrix A, int *out);

/* copyto!(dst, src) for limb-planar matrices as a device kernel (usable
 * inside a captured HIP graph, unlike a memcpy node). */
int mfla_copy(mfla_handle h, int limbs, mfla_matrix dst, mfla_matrix src);

/* AoS (Julia Matrix{Float64xN}, limbs interleaved) <-> limb-planar copies */
int mfla_from_interleaved(mfla_handle h, int limbs, mfla_matrix dst, const double *src_dev, int64_t src_ld);
int mfla_to_interleaved(mfla_handle h, int limbs, double *dst_dev, int64_t dst_ld, mfla_matrix src);

/* ------------------------------------------------------------ BLAS 3 */
/* gemm! (src/kernels/gemm.jl:747-774): C = alpha*A*B + beta*C */
int mfla_gemm(mfla_handle h, int limbs, mfla_matrix C, mfla_matrix A, mfla_matrix B, const double *alpha, const double *beta,
              const mfla_config *config);
/* syrk! (src/kernels/syrk.jl:67-108): lower C = alpha*P^T*P + beta*C, P is k x n */
int mfla_syrk(mfla_handle h, int limbs, int uplo, mfla_matrix C, mfla_matrix P, const double *alpha, const double *beta,
              const mfla_config *config);
/* gemmt! (src/kernels/gemmt.jl:69-108): lower C = alpha*L*R^T + beta*C */
int mfla_gemmt(mfla_handle h, int limbs, mfla_matrix C, mfla_matrix L, mfla_matrix R, const double *alpha, const double *beta,
               const mfla_config *config);
/* trsm! (src/kernels/trsm.jl:224-302): op(A) X = alpha B or X op(A) = alpha B, in place */
int mfla_trsm(mfla_handle h, int limbs, int side, int uplo, int trans, int diag, mfla_matrix B, mfla_matrix A, const double *alpha,
              const mfla_config *config);

/* Route plans. Reason codes of GemmPlan / LDLTPlan (MFLA's symbols
 * :forced_direct ... :auto_fused_direct, :forced_unblocked ...). */
#define MFLA_PLAN_FORCED_DIRECT 1
#define MFLA_PLAN_FORCED_PACKED 2
#define MFLA_PLAN_FORCED_FUSED 3
#define MFLA_PLAN_FORCED_OZAKI 4
#define MFLA_PLAN_AUTO_OUTSIDE_CALIBRATED_SHAPE 5
#define MFLA_PLAN_AUTO_REDUCTION_TOO_SMALL 6
#define MFLA_PLAN_AUTO_BELOW_CROSSOVER 7
#define MFLA_PLAN_AUTO_ABOVE_CROSSOVER 8
#define MFLA_PLAN_AUTO_FUSED_DIRECT 9
#define MFLA_PLAN_FORCED_UNBLOCKED 10
#define MFLA_PLAN_FORCED_BLOCKED 11

/* GemmPlan (src/types.jl:210-217) as gemm_plan(T, m, k, n, config)
 * (src/kernels/gemm.jl:44-136) resolves it, plus the device launch gemm!
 * makes for it. strategy is the resolved route (MFLA_GEMM_DIRECT, _PACKED,
 * _FUSED, or _OZAKI when forced; never _AUTO); reason an MFLA_PLAN_* code.
 * panel_columns, micro_columns, workers and packed_elements_per_worker are
 * MFLA's CPU decomposition for config->thread_count (kept so plan-driven
 * caller code sees MFLA's values); the device computes every route with the
 * tile kernel (packed == direct per element: the standard network, ascending
 * k). Device fields: tile (0..3) and its tile_rows x tile_cols outputs per
 * workgroup; network 1 = mulacc_x3, 0 = the standard network; launches = 2
 * when the x3 mulacc route has remainder_rows = m mod 4 rows on the standard
 * network; workgroups over all launches. */
typedef struct mfla_gemm_plan_t {
    int32_t strategy;
    int32_t reason;
    int64_t panel_columns;
    int64_t micro_columns;
    int64_t workers;
    int64_t packed_elements_per_worker;
    int32_t tile;
    int32_t tile_rows;
    int32_t tile_cols;
    int32_t network;
    int64_t launches;
    int64_t workgroups;
    int64_t remainder_rows;
} mfla_gemm_plan_t;
int mfla_gemm_plan(int limbs, int64_t m, int64_t k, int64_t n, const mfla_config *config, mfla_gemm_plan_t *out);

/* LDLTPlan (src/types.jl:219-223) as ldlt_plan(T, n, config)
 * (src/factorizations/ldlt.jl:329-350) resolves it: strategy
 * MFLA_LDLT_UNBLOCKED / _BLOCKED, reason MFLA_PLAN_FORCED_UNBLOCKED,
 * _FORCED_BLOCKED, _AUTO_BELOW_CROSSOVER or _AUTO_ABOVE_CROSSOVER, and the
 * panel width (16 / 12 / 8 for 1-2 / 3 / 4 limbs unless config->ldlt_block,
 * src/factorizations/ldlt_tuning.jl). ldlt! (mfla_sytrf) follows it. */
typedef struct mfla_ldlt_plan_t {
    int32_t strategy;
    int32_t reason;
    int64_t block_size;
} mfla_ldlt_plan_t;
int mfla_ldlt_plan(int limbs, int64_t n, const mfla_config *config, mfla_ldlt_plan_t *out);

/* ------------------------------------------------------------ BLAS 2 */
/* trsv! (src/kernels/trsv.jl:19-73): x is an n x 1 matrix; same order as trsm! */
int mfla_trsv(mfla_handle h, int limbs, int uplo, int trans, int diag, mfla_matrix x, mfla_matrix A, const double *alpha,
              const mfla_config *config);
/* gemv! (src/kernels/gemv.jl:111-167): y = alpha*op(A)*x + beta*y */
int mfla_gemv(mfla_handle h, int limbs, int trans, mfla_matrix y, mfla_matrix A, mfla_matrix x, const double *alpha, const double *beta,
              const mfla_config *config);
/* symv! (src/kernels/symv.jl:53-89): y = alpha*A*x + beta*y from the uplo triangle */
int mfla_symv(mfla_handle h, int limbs, int uplo, mfla_matrix y, mfla_matrix A, mfla_matrix x, const double *alpha, const double *beta,
              const mfla_config *config);

/* ---------------------------------------------------- factorizations */
/* cholesky! (src/factorizations/cholesky.jl:68-107): lower, in place */
int mfla_potrf(mfla_handle h, int limbs, mfla_matrix A, int64_t *info, const mfla_config *config);
/* lu! (src/factorizations/lu.jl:128-195): ipiv_dev has min(m, n) int64 */
int mfla_getrf(mfla_handle h, int limbs, mfla_matrix A, int64_t *ipiv_dev, double *original_maximum, int64_t *info,
               const mfla_config *config, mfla_workspace ws);
/* ldlt! (src/factorizations/ldlt.jl:654-717): dsub is an n x 1 device matrix;
 * pivots_dev (int64) and blocks_dev (uint8) have n entries */
int mfla_sytrf(mfla_handle h, int limbs, mfla_matrix A, mfla_matrix dsub, int64_t *pivots_dev, uint8_t *blocks_dev,
               double *original_maximum, int64_t *info, const mfla_config *config, mfla_workspace ws);

/* ldlt! that allocates nothing. No counterpart in MFLA: for callers that
 * cannot allocate device memory during a call (code that sets up a HIP graph
 * capture, tight serving loops, allocation-free harnesses). Arguments and
 * results are those of mfla_sytrf (bit for bit) without a workspace: the
 * blocked route's weighted panel and the original_maximum column maxima
 * live in a static device array of MFLA_NOALLOC_SCRATCH_BYTES on each GPU
 * (held until the process ends; its address is resolved per device when a
 * handle is created). A call needs limbs * n * (block + 2) * 8 + 256 bytes
 * of it (block = the ldlt_plan panel width; the unblocked route
 * limbs * n * 8 + 256), else it returns MFLA_ERR_WORKSPACE before touching
 * anything: e.g. Float64x4 (block 8) up to n = 26213. Every call on one GPU
 * shares the array and leaves its contents behind: calls that overlap on
 * one GPU (two handles, two host threads) race. */
#define MFLA_NOALLOC_SCRATCH_BYTES (8u * 1024u * 1024u)
int mfla_sytrf_noalloc(mfla_handle h, int limbs, mfla_matrix A, mfla_matrix dsub, int64_t *pivots_dev, uint8_t *blocks_dev,
                       double *original_maximum, int64_t *info, const mfla_config *config);

/* Householder QR. tau is a min(m, n) x 1 device matrix, perm_dev n int64
 * (one-based, A_original[:, perm] = Q R). Non-finite input returns info = -1
 * with tau = 0 and the identity permutation; A is then untouched.
 * qr! (src/factorizations/qr.jl:763-799), no pivoting; the fixed-order form
 * qr!(A, permutation) (qr.jl:813-853) reorders the columns first. */
int mfla_geqrf(mfla_handle h, int limbs, mfla_matrix A, mfla_matrix tau, int64_t *info);
/* rrqr! (src/factorizations/qr.jl:692-743): route 0 = MFLA's choice (the
 * delayed-norm panel route of qr.jl:450-604 when min(m, n) >= 16 and
 * m n >= 16384, else qr.jl:364-398), 1 = unblocked, 2 = blocked */
int mfla_geqp3(mfla_handle h, int limbs, mfla_matrix A, mfla_matrix tau, int64_t *perm_dev, int route, int64_t *info);
/* apply_q! (src/factorizations/qr.jl:938-956): B <- Q B (MFLA_NOTRANS) or
 * Q' B (MFLA_TRANS); B has the factor's row count and must not alias it */
int mfla_ormqr(mfla_handle h, int limbs, int trans, mfla_matrix F, mfla_matrix tau, mfla_matrix B);

/* ------------------------------------------------------------ solves */
/* ldiv!(B, F::MFCholesky) (src/solve.jl:23-83) */
int mfla_potrs(mfla_handle h, int limbs, mfla_matrix F, int64_t info, mfla_matrix B, const mfla_config *config);
/* ldiv!(B, F::MFLU) (src/solve.jl:115-181) */
int mfla_getrs(mfla_handle h, int limbs, mfla_matrix F, const int64_t *ipiv_dev, int64_t info, mfla_matrix B,
               const mfla_config *config);
/* ldiv!(B, F::MFLDLT) (src/solve.jl:298-365) */
int mfla_sytrs(mfla_handle h, int limbs, mfla_matrix F, mfla_matrix dsub, const int64_t *pivots_dev, const uint8_t *blocks_dev,
               int64_t info, mfla_matrix B, const mfla_config *config);

/* ------------------------------------------------ residual / refinement */
/* residual!(R, A, X, B) for matrix right-hand sides (src/residual.jl:82-125,
 * the gemm!/symv! path): R = B - A*X; R may be B itself */
int mfla_residual(mfla_handle h, int limbs, int uplo, mfla_matrix R, mfla_matrix A, mfla_matrix X, mfla_matrix B,
                  const mfla_config *config);
/* residual!(r, A, x, b) for a vector right-hand side (src/residual.jl:55-80,
 * the gemv!/symv! path); r, x, b are n x 1 matrices */
int mfla_residual_vector(mfla_handle h, int limbs, int uplo, mfla_matrix r, mfla_matrix A, mfla_matrix x, mfla_matrix b,
                         const mfla_config *config);
/* residual_mixed! (src/residual.jl:431-496): R (res_limbs) = B - A*X with
 * A, X, B (src_limbs) promoted first. Pairs: 2->3, 2->4, 3->4 (MFLA) and
 * 1->2/3/4 (extension used by FP64-factor refinement). */
int mfla_residual_mixed(mfla_handle h, int src_limbs, int res_limbs, int uplo, mfla_matrix R, mfla_matrix A, mfla_matrix X,
                        mfla_matrix B, const mfla_config *config);
/* X = X + D with D (src_limbs <= dst_limbs) zero-extended to X's precision:
 * the correction-accumulation step of iterative refinement. MFLA leaves the
 * refinement loop to the caller (residual.jl:294-300, approximation .+=
 * correction in test/runtests.jl:2132); this is that update on the device. */
int mfla_add_mixed(mfla_handle h, int dst_limbs, int src_limbs, mfla_matrix X, mfla_matrix D);
/* normwise_backward_error (src/residual.jl:215-269): out is a host array of
 * limbs x nrhs binary64 (limb-planar), one value per right-hand side */
int mfla_backward_error(mfla_handle h, int limbs, int uplo, mfla_matrix A, mfla_matrix X, mfla_matrix B, mfla_matrix R, double *out);

/* ------------------------------------------------- batched small systems */
/* MFLA has no batched API (callers loop cholesky!/ldlt! + ldiv! per system);
 * this runs many uneven systems in one persistent launch. Each system is
 * factored with MFLA's own algorithm (cholesky!: blocked, block
 * config->cholesky_block; ldlt!: the unblocked Bunch-Kaufman route MFLA uses
 * below its crossover) and solved in place on its right-hand sides. A system
 * occupies pool storage at the entry's offsets: A (limbs planes of n x n,
 * limb stride n*n), B (limbs planes of n x nrhs), and for LDLT dsub (limbs
 * planes of n), pivots and blocks (n entries at meta_offset). info[s] gets the
 * MFLA status (or -2 if n > mfla_batched_max_order() or nrhs >
 * mfla_batched_max_rhs(), or n > threads_per_system); origmax[s*limbs..]
 * the LDLT original_maximum. threads_per_system is the workgroup shape: 64,
 * 128 or 256 threads (0 = 256), one thread per row, so a workgroup takes
 * systems with n up to that; the smaller shapes put 16 or 8 systems on a CU
 * instead of 4 (mfla.batched.autotune routes size classes of a batch to
 * shapes). Results do not depend on the shape or on workgroups (0 = a
 * default persistent count). The call allocates nothing and never
 * synchronises: it can be captured in a HIP graph. `counter` is one device
 * unsigned int (reset by the call). */
#define MFLA_BATCH_CHOLESKY 0
#define MFLA_BATCH_LDLT 1
typedef struct mfla_batch_entry {
    int64_t n;
    int32_t kind;        /* MFLA_BATCH_* */
    int32_t nrhs;
    int64_t a_offset;    /* elements into A_pool */
    int64_t b_offset;    /* elements into B_pool */
    int64_t meta_offset; /* elements into dsub/pivot/block pools */
} mfla_batch_entry;

int mfla_batched_solve(mfla_handle h, int limbs, const mfla_batch_entry *entries_dev, int64_t count, double *A_pool,
                       double *B_pool, double *dsub_pool, int64_t *pivot_pool, uint8_t *block_pool, int64_t *info_dev,
                       double *origmax_dev, unsigned int *counter_dev, int64_t workgroups, int threads_per_system,
                       int exact_backward, const mfla_config *config);
int mfla_batched_max_order(void);
int mfla_batched_max_rhs(void);

/* ------------------------------------------- API completion (BLAS-3 extras) */
/* trmm! (src/kernels/trmm.jl:75-131): B = alpha*op(A)*B (side LEFT) or
 * B = alpha*B*op(A) (RIGHT) in place, MFLA's per-element order (ascending
 * inner index, diagonal term first or last as in trmm.jl:1-73). Only the
 * selected triangle of A is used. */
int mfla_trmm(mfla_handle h, int limbs, int side, int uplo, int trans, int diag, mfla_matrix B, mfla_matrix A,
              const double *alpha, const mfla_config *config);
/* mfdot (src/kernels/dot.jl:1-22): sum of x[i]*y[i] over the column-major
 * linear index, one ascending chain; out gets the limbs (host). */
int mfla_dot(mfla_handle h, int limbs, mfla_matrix x, mfla_matrix y, double *out);
/* syrk_packed! (src/kernels/syrk.jl:181-254): packed (a limbs x n(n+1)/2
 * column vector holding the column-major lower triangle) = alpha *
 * P[rf:rl, :]^T P[rf:rl, :] + beta * packed; rf, rl one-based inclusive. */
int mfla_syrk_packed(mfla_handle h, int limbs, mfla_matrix packed, mfla_matrix P, const double *alpha, const double *beta,
                     int64_t reduction_first, int64_t reduction_last, const mfla_config *config);

/* ------------------------------------------ FP64 factor + Float64xN refinement */
/* MFLA supplies the refinement pieces -- residual! (src/residual.jl:55-125),
 * refinement_correction! (one ldiv!, src/residual.jl:302-332) and
 * normwise_backward_error (src/residual.jl:215-269) -- and leaves the loop to
 * the caller. This runs that loop with the factorization done in binary64:
 *   F = cholesky / lu (partial pivoting) of the leading limb of A;
 *   X = 0, R = B; repeat: D = F \ hi(R); X += D; R -= A*D
 * until the normwise backward error estimate max_j |R_j| / (|A| |X_j| + |B_j|)
 * (infinity norms, binary64) is <= tol or stops halving. The product A*D is
 * summed with as many limbs as the size of D relative to X requires to keep
 * the residual accurate to the Float64xN unit roundoff. limbs must be 2..4.
 * `work` is a caller-owned device buffer of at least
 * mfla_refine_workspace_size() doubles. history (host, max_iter entries) gets
 * the estimate after each correction; stage_ms (host, 5 entries, may be NULL)
 * milliseconds spent in factor, solve, update, residual and check.
 * The loop also stops when an estimate is more than half the previous one;
 * a correction that made the estimate worse is taken back, so X is always
 * the best iterate (the smallest entry of history). Returns MFLA_OK when
 * that iterate's estimate is <= tol, else MFLA_NOT_CONVERGED (X still valid:
 * e.g. kappa(A) * u_64 >= 1, or tol below the Float64xN residual floor).
 * *info is the binary64 factorization's status (0, -1 non-finite input, k > 0
 * failing pivot); a failed factorization returns MFLA_ERR_FAILED_FACTOR. */
#define MFLA_REFINE_CHOLESKY 0
#define MFLA_REFINE_LU 1
/* OR into kind: residual updates R -= A*D as Ozaki-split INT8 MFMA products
 * (A split once, n <= 8192) instead of the limb-adaptive multi-double sums. */
#define MFLA_REFINE_OZAKI 16
int mfla_refine_workspace_size(int limbs, int kind, int64_t n, int64_t nrhs, int64_t *doubles);
int mfla_refine_solve(mfla_handle h, int limbs, int kind, int uplo, mfla_matrix A, mfla_matrix B, mfla_matrix X, double *work,
                      int64_t work_len, double tol, int max_iter, int64_t *info, int *iterations, double *history,
                      double *stage_ms);

/* Binary64 building blocks of the refinement, exposed for testing and for
 * callers that run their own loop (no counterpart in MFLA).
 * mfla_dgemm: C = alpha*op(A)*op(B) + beta*C on binary64 matrices (the
 * leading limb plane of a descriptor), op = transpose for MFLA_TRANS.
 * tri_a / tri_b: MFLA_LOWER / MFLA_UPPER declare op(A) / op(B) triangular
 * (the other triangle is not read), MFLA_GENERAL a dense operand.
 * mfla_dpotrf_inverse: binary64 Cholesky of the lower triangle of A in place
 * and the inverse of the factor into the lower triangle of Inv (its strict
 * upper triangle is used as scratch).
 * mfla_dgetrf: binary64 LU with partial pivoting of the square A in place
 * (LAPACK dgetrf semantics: device ipiv of n one-based row indices, info =
 * first zero pivot column, -1 for non-finite input; n <= 32768). */
int mfla_dgemm(mfla_handle h, int trans_a, int trans_b, int tri_a, int tri_b, mfla_matrix C, mfla_matrix A, mfla_matrix B,
               double alpha, double beta);
int mfla_dpotrf_inverse(mfla_handle h, mfla_matrix A, mfla_matrix Inv, int64_t *info);
int mfla_dgetrf(mfla_handle h, mfla_matrix A, int64_t *ipiv, int64_t *info);

/* ----------------------------------------------------- reverse-mode rules */
/* MFLA has no differentiation rules (its dependencies are LinearAlgebra and
 * MultiFloats only, Project.toml:6-8). These are the pullbacks a
 * ChainRulesCore rrule (or a torch.autograd.Function) calls; every one reuses
 * the Cholesky factor F of S = L L^T produced by mfla_potrf (info must be
 * 0) instead of refactoring. Gradients with respect to the symmetric S are
 * returned symmetrized: sym(G) = (G + G^T) / 2, exactly symmetric.
 * mfla_logdet: logdet(S) = 2 sum_i log L_ii into out (host, limbs values).
 * mfla_solve_adjoint: X = S^{-1} B with cotangent Xbar gives
 *   Bbar = S^{-1} Xbar (the solve of src/factorizations/solve.jl:23-80 on
 *   Xbar) and Sbar = -sym(Bbar X^T). Bbar may be the same matrix as Xbar.
 * mfla_logdet_adjoint: Sbar = ybar * S^{-1} (ybar: host limbs).
 * mfla_cholesky_adjoint: L = chol(S) with cotangent Lbar (lower triangle
 *   used) gives Sbar = sym(L^{-T} Phi(L^T Lbar) L^{-1}), Phi = lower triangle
 *   with the diagonal halved; work is an n x n scratch matrix. */
int mfla_logdet(mfla_handle h, int limbs, mfla_matrix F, int64_t info, double *out);
int mfla_solve_adjoint(mfla_handle h, int limbs, mfla_matrix F, int64_t info, mfla_matrix X, mfla_matrix Xbar, mfla_matrix Bbar,
                       mfla_matrix Sbar, const mfla_config *config);
int mfla_logdet_adjoint(mfla_handle h, int limbs, mfla_matrix F, int64_t info, const double *ybar, mfla_matrix Sbar,
                        const mfla_config *config);
int mfla_cholesky_adjoint(mfla_handle h, int limbs, mfla_matrix F, int64_t info, mfla_matrix Lbar, mfla_matrix Sbar,
                          mfla_matrix work, const mfla_config *config);

#ifdef __cplusplus
}
#endif

#endif /* MFLA_HIP_H */
