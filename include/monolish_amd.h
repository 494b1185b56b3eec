/*
 * monolish_amd.h -- C-ABI of the MI355X (gfx950) CRS sparse BLAS and Krylov
 * solve path.  Plain pointers and sizes only; every call returns an int status
 * (MGX_OK = 0; solver outcomes use monolish's MONOLISH_SOLVER_* values;
 * ABI errors are <= -100) and never throws.  mgx_last_error() gives the
 * message of the last failure on the calling thread.
 *
 * Each entry point names the monolish interface it replaces (paths relative
 * to the monolish 0.17.3 source tree).  The drop-in C++ layer
 * (include/monolish_*.hpp) calls these when data is device resident.
 *
 * Threading: one mgx_ctx per host thread (or external locking); all work of a
 * context is ordered on its HIP stream.
 */
#ifndef MONOLISH_AMD_H
#define MONOLISH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
/* solver outcomes: include/monolish/common/monolish_common.hpp:16-21 */
#define MGX_OK 0
#define MGX_SOLVER_SIZE_ERROR -1
#define MGX_SOLVER_MAXITER -2
#define MGX_SOLVER_BREAKDOWN -3
#define MGX_SOLVER_RESIDUAL_NAN -4
#define MGX_SOLVER_NOT_IMPL -10
/* ABI errors */
#define MGX_ERR_INVALID -101  /* bad argument, size or state */
#define MGX_ERR_HIP -102      /* HIP runtime failure */
#define MGX_ERR_NOMEM -103    /* device or host allocation failed */
#define MGX_ERR_BUDGET -104   /* request exceeds the configured HBM budget */
#define MGX_ERR_NOT_FOUND -105
#define MGX_ERR_TIMEOUT -106  /* a wait gave up: a device wait without progress, or a
                                 served request past its deadline */
#define MGX_ERR_CANCELLED -107 /* a served request cancelled by mgx_server_cancel */

typedef enum { MGX_F64 = 0, MGX_F32 = 1 } mgx_dtype;

typedef struct mgx_ctx mgx_ctx;
typedef struct mgx_vec mgx_vec;
typedef struct mgx_csr mgx_csr;

/* ---- context ------------------------------------------------------------ */
const char *mgx_last_error(void);
int mgx_version(int *major, int *minor, int *patch);
int mgx_device_count(int *count);
/* util::set_default_device / device query, src/utils/system/gpu_deivce.cpp:6-34 */
int mgx_init(int device, mgx_ctx **out);
/* releases the context and every vector and matrix of it still alive; their
 * handles must not be used afterwards */
int mgx_finalize(mgx_ctx *ctx);
int mgx_sync(mgx_ctx *ctx);
/* the context's hipStream_t (as void*) for callers that record events */
void *mgx_stream(mgx_ctx *ctx);
/* Queue every later call of the context on the caller's hipStream_t (as
 * void*; NULL: back to the library's own stream).  Waits for the work queued
 * so far.  The stream stays the caller's: it must outlive its use and be
 * replaced (NULL) before it is destroyed; not while a server runs on the
 * context.  (monolish's GPU calls run on the default stream and synchronise;
 * this is the per-call stream of the C-ABI, set per context.) */
int mgx_set_stream(mgx_ctx *ctx, void *stream);
/* name (NUL-terminated, truncated to len), compute units, HBM bytes */
int mgx_device_info(mgx_ctx *ctx, char *name, size_t len, int *compute_units,
                    size_t *hbm_bytes);
/* SpMV tuning: rows per block limit, entries per block limit, long-row
 * threshold, XCD-aware block order.  Applies to CRS created afterwards. */
int mgx_set_spmv_tuning(mgx_ctx *ctx, int rows_per_block, int nnz_per_block,
                        int long_row, int xcd_remap);
/* Autotuner of the SpMV block plan (no monolish counterpart: monolish hands
 * CRS SpMV to cuSPARSE): times A's SpMV under candidate plans -- rows per
 * block {64, 128, 256} x entries per block {1024, 2048, 4096} x XCD-aware
 * order on/off with products staged in LDS in entry order; rows per block
 * {128, 256} x entries {1024, 2048} x XCD order with the row walk -- then,
 * for matrices whose x does not fit an XCD's L2 and whose rows list columns
 * in ascending order, the best of these with 4, 8, 12 or 16 column panels (the
 * short rows' entries stored a second time, panel-major, so each pass
 * gathers x from an L2-sized slice), and, where 64-row slices pad the
 * entries by at most 30 %, the sliced layout (entries stored a second time
 * in 64-row slices of 16-byte tuples, one lane per row, no LDS) in launch
 * order, with one range of row blocks per XCD, and with runs of 16 row
 * blocks dealt to the XCDs in turn -- `reps` launches each; keeps the
 * fastest for A and reports it
 * (within 2 %, an XCD-ordered block plan).  The long-row threshold is
 * unchanged, so results do not depend on the choice.  The column-panel and
 * L2-panel copies are dropped by mgx_csr_update_values; the sliced copy is
 * rewritten. */
typedef struct mgx_spmv_tuning {
  int32_t rows_per_block, nnz_per_block, long_row, xcd_remap;
  double us;           /* time of one SpMV under the chosen plan */
  int32_t candidates;  /* plans timed */
  int32_t column_panels; /* 0, or the column panels of the chosen plan */
  int32_t row_walk;    /* 0 products staged in LDS in entry order; 1 row
                          walk (lane per row over LDS-staged val / col_ind;
                          coalesced gathers on banded rows); 2 the sliced
                          layout; 3 L2 panels (row sums kept in LDS while the
                          blocks walk column panels of about 0.5-2 MiB of x,
                          column_panels of them) */
  int32_t pad;
} mgx_spmv_tuning;
int mgx_csr_autotune(mgx_csr *A, int reps, mgx_spmv_tuning *out);
/* Sets A's SpMV plan explicitly (the candidates the autotuner times):
 * rows_per_block in [1, 256], nnz_per_block in [64, 8192] (<= 4096 with
 * row_walk), xcd_remap 0 (launch order), 1 (one range of blocks per XCD)
 * or a run length in [2, 1024] (runs of that many consecutive blocks dealt
 * to the XCDs in turn), row_walk 0/1 (block kind), 2 (the sliced
 * layout, over a product-staged block plan kept for other values; refused
 * when its padding exceeds 30 %) or 3 (L2 panels: column_panels in
 * [1, 65536], at most 2^20 columns each, ascending columns within rows),
 * column_panels 0 or 2..64 otherwise (needs ascending columns within rows;
 * not with row_walk 2): MGX_ERR_INVALID otherwise.
 * The long-row threshold stays as it is. */
int mgx_csr_set_plan(mgx_csr *A, int rows_per_block, int nnz_per_block, int xcd_remap,
                     int row_walk, int column_panels);
/* 1: fused solvers and batches replay each group of iterations from a
 * captured hipGraph (re-captured when the set of running systems changes);
 * 0 (default): direct kernel launches */
int mgx_set_graphs(mgx_ctx *ctx, int enable);
/* 1: a fused CG solve of one system small enough for the two-kernel
 * iteration runs its iterations inside one cooperative launch (grid barriers
 * between the stages); 0 (default, measured faster on MI355X): one launch per
 * stage; a value above 1 enables it with at most that many blocks (each
 * block then owns several 128-row quarters) */
int mgx_set_persistent(mgx_ctx *ctx, int enable);

/* ---- vectors: vector<T>::send/recv/nonfree_recv/device_free ------------- */
/* vector<T>::send (src/utils/system/gpu_comm.cpp:8-24): device copy of n
 * elements; host may be NULL for a zero-filled vector */
int mgx_vec_create(mgx_ctx *ctx, int64_t n, mgx_dtype dtype, const void *host,
                   mgx_vec **out);
/* view1D<vector<T>,T>(vec, offset, n) (include/monolish/common/
 * monolish_view1D.hpp:50): shares the base vector's device memory */
int mgx_vec_view(mgx_vec *base, int64_t offset, int64_t n, mgx_vec **out);
/* send() on a resident vector: host -> device update (gpu_comm.cpp:14-16) */
int mgx_vec_upload(mgx_vec *v, const void *host);
/* nonfree_recv (gpu_comm.cpp:45-57): device -> host, stays resident */
int mgx_vec_download(mgx_vec *v, void *host);
/* device_free (gpu_comm.cpp:60-73); recv = download + destroy */
int mgx_vec_destroy(mgx_vec *v);
/* vector(N, value) / internal::vbroadcast (double_arithmetic.cpp:255-275) */
int mgx_vec_fill(mgx_vec *v, double value);
int64_t mgx_vec_size(const mgx_vec *v);
void *mgx_vec_device_ptr(mgx_vec *v);

/* ---- CRS: matrix::CRS<T> residency (gpu_comm.cpp:89-171) ---------------- */
/* CRS<T>::send: uploads row_ptr[rows+1], col_ind[nnz], val[nnz] (int32,
 * 0-based, monolish_crs.hpp:76-103) and builds the SpMV row-block plan.
 * Validates indices (row_ptr non-decreasing except the all-zero tail that
 * CRS(COO&) leaves after the last non-empty row; 0 <= col < cols). */
int mgx_csr_create(mgx_ctx *ctx, int64_t rows, int64_t cols, int64_t nnz,
                   const int32_t *row_ptr, const int32_t *col_ind,
                   const void *val, mgx_dtype dtype, mgx_csr **out);
/* CRS<T>(COO<T>&) (src/utils/convert/convert_crs.cpp:9-43) from host COO
 * triplets (0-based int32 row, col; values of dtype), built on the device.
 * sort = 0: the triplets are taken in the given order, which must list rows
 * in ascending order, as CRS(COO&) reads them (MGX_ERR_INVALID otherwise);
 * sort = 1: first ordered by (row, col) like COO::sort(false)
 * (src/utils/at_insert_sort/at_insert_sort_coo.cpp:119-144), stably, so
 * equal (row, col) keep their input order; sort = 2: ordered and merged
 * like COO::sort(true): each run of equal (row, col) collapses to one entry
 * holding the value of the last one in input order.  Indices outside
 * rows x cols are rejected.  row_ptr of rows after the last non-empty one
 * reads back as 0, as CRS(COO&) leaves it. */
int mgx_csr_from_coo(mgx_ctx *ctx, int64_t rows, int64_t cols, int64_t nnz,
                     const int32_t *row, const int32_t *col, const void *val,
                     mgx_dtype dtype, int sort, mgx_csr **out);
/* COO<T>::input_mm (src/utils/IO/IO_coo.cpp:69-153) followed by
 * mgx_csr_from_coo: a MatrixMarket "matrix coordinate real general" file
 * (other banners rejected), '%' comment lines, the size line, then nnz
 * lines "i j value" with 1-based indices; values parsed as dtype (FP32 text
 * rounded once to float, as istream >> float).  The text is parsed by host
 * threads; sorting, merging and row_ptr run on the device. */
int mgx_csr_read_mm(mgx_ctx *ctx, const char *path, mgx_dtype dtype, int sort,
                    mgx_csr **out);
/* COO<T>::output_mm (src/utils/IO/IO_coo.cpp:50-66) of COO(A): the same
 * banner, size line and "i j value" lines (row order, 1-based, values in
 * scientific notation with max_digits10 digits), written by host threads. */
int mgx_csr_write_mm(mgx_csr *A, const char *path);
/* CRS<T>::send on a resident matrix: value update (gpu_comm.cpp:95-99) */
int mgx_csr_update_values(mgx_csr *A, const void *val);
/* CRS<T>::nonfree_recv (gpu_comm.cpp:134-150); any pointer may be NULL */
int mgx_csr_download(mgx_csr *A, int32_t *row_ptr, int32_t *col_ind,
                     void *val);
/* CRS<T>::device_free (gpu_comm.cpp:153-171) */
int mgx_csr_destroy(mgx_csr *A);
/* CRS<T>::diag (src/utils/getvec/crs_getvec.cpp:8-55) */
int mgx_csr_diag(mgx_csr *A, mgx_vec *d);

typedef struct mgx_csr_stats {
  int64_t rows, cols, nnz;
  int32_t max_row, min_row;
  double mean_row;
  int64_t stream_blocks, long_rows; /* SpMV plan */
  /* distinct 128-byte lines of x gathered by 64 consecutive entries
     (sampled): above 24, fused CG forms p before its SpMV and gathers it
     once per entry instead of gathering z and p */
  double gather_lines;
} mgx_csr_stats;
int mgx_csr_stats_get(mgx_csr *A, mgx_csr_stats *out);

/* ---- sparse BLAS -------------------------------------------------------- */
/* blas::matvec(A, x, y) / matvec_N (src/blas/matrix/matvec/
 * matvec_blas.cpp:3849-3860, crs_matvec.hpp:8-114): y = A x */
int mgx_matvec(mgx_csr *A, mgx_vec *x, mgx_vec *y);
/* blas::matvec(a, A, x, b, y): y = a A x + b y (the device path's meaning,
 * crs_matvec.hpp:37-65) */
int mgx_matvec_ab(double alpha, mgx_csr *A, mgx_vec *x, double beta,
                  mgx_vec *y);
/* blas::matvec_T (crs_matvec.hpp:88-98): y = A^T x */
int mgx_matvec_t(mgx_csr *A, mgx_vec *x, mgx_vec *y);

/* ---- vector BLAS / VML -------------------------------------------------- */
/* blas::dot (src/blas/vector/dot.hpp:6-78): blocking, scalar to host */
int mgx_dot(mgx_vec *x, mgx_vec *y, double *out);
/* blas::nrm2 (src/blas/vector/nrm2.hpp:6-67): sqrt(dot(x,x)) */
int mgx_nrm2(mgx_vec *x, double *out);
/* blas::dot / blas::nrm2 with the result left in device memory: *d_out (a
 * double in device memory) receives the same bits mgx_dot / mgx_nrm2 return,
 * queued on the context's stream without a host round trip. */
int mgx_dot_device(mgx_vec *x, mgx_vec *y, double *d_out);
int mgx_nrm2_device(mgx_vec *x, double *d_out);
/* No counterpart in the reference.  mgx_dot_device for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops that
 * forbid hipMalloc): the same bits, with the reduction's counters and
 * partials in a static device array of about 1 MiB per device (held until
 * the process ends) instead of library-allocated buffers; allocates nothing.
 * At most 2^26 elements (larger inputs: MGX_ERR_INVALID).  Every call on a
 * device shares that device's array, so calls that overlap on one device
 * (other streams, other threads) race. */
int mgx_dot_device_noalloc(mgx_vec *x, mgx_vec *y, double *d_out);
/* blas::axpy (axpy.hpp:7-61): y = a x + y */
int mgx_axpy(double a, mgx_vec *x, mgx_vec *y);
/* blas::xpay (xpay.hpp:7-66): y = x + a y */
int mgx_xpay(double a, mgx_vec *x, mgx_vec *y);
/* blas::scal (scal.hpp:6-49): x = a x */
int mgx_scal(double a, mgx_vec *x);
/* blas::copy (copy.hpp:5-13): y = x */
int mgx_copy(mgx_vec *x, mgx_vec *y);
/* blas::axpyz (axpyz.hpp:7-66): z = a x + y */
int mgx_axpyz(double a, mgx_vec *x, mgx_vec *y, mgx_vec *z);
/* vml::add / sub / mul (src/internal/arithmetic/double_arithmetic.cpp:
 * 112-195): y = a op b */
int mgx_vadd(mgx_vec *a, mgx_vec *b, mgx_vec *y);
int mgx_vsub(mgx_vec *a, mgx_vec *b, mgx_vec *y);
int mgx_vmul(mgx_vec *a, mgx_vec *b, mgx_vec *y);
/* vml::reciprocal (double_arithmetic.cpp:308-330): y = 1 / a */
int mgx_vreciprocal(mgx_vec *a, mgx_vec *y);

/* ---- Krylov solvers ----------------------------------------------------- */
#define MGX_PRECOND_DEFAULT 0 /* no preconditioner bound: z = r */
#define MGX_PRECOND_NONE 1    /* equation::none (src/equation/none.cpp) */
#define MGX_PRECOND_JACOBI 2  /* equation::Jacobi (src/equation/jacobi.cpp) */
#define MGX_PRECOND_SOR 3     /* equation::SOR (src/equation/sor.cpp): forward and
                                 backward sweeps with 1/(omega diag(A)), run as
                                 level-scheduled kernels with the serial loop's
                                 bits; solvers using it take the reference call
                                 sequence in either mode */
#define MGX_PRECOND_IC 4      /* equation::IC (src/equation/ic.cpp): IC(0) of the
                                 lower pattern, L L^T; square, sorted columns,
                                 A(i,i) stored, SPD (MGX_ERR_INVALID naming the
                                 row otherwise); factor and both triangular
                                 solves level-scheduled on the device */
#define MGX_PRECOND_ILU 5     /* equation::ILU (src/equation/ilu.cpp): ILU(0), L U
                                 on A's pattern; same requirements, a zero pivot
                                 is an error.  Solvers with IC / ILU take the
                                 reference call sequence in either mode */

/* how the iteration is executed; every mode follows the reference's
 * arithmetic, so FP64 results of REFERENCE and FUSED are bit-identical */
#define MGX_MODE_REFERENCE 0 /* one kernel per BLAS call, scalars to host */
#define MGX_MODE_FUSED 1     /* fused kernels, device scalars (CG, BiCGSTAB) */
#define MGX_MODE_MIXED_IR 2  /* CG of an FP64 system by FP32-refined iterations
                                (mgx_ir_cg_solve with inner_tol 1e-5, at most 100
                                refinement steps; served requests: the server's
                                ir_inner_tol / ir_max_outer) */

typedef struct mgx_solve_params {
  double tol;          /* solver::set_tol (monolish_solver.hpp:86-215) */
  int64_t maxiter;     /* solver::set_maxiter; <= 0 means unlimited */
  int64_t miniter;     /* solver::set_miniter */
  int32_t precond;     /* MGX_PRECOND_* */
  int32_t mode;        /* MGX_MODE_* */
  int32_t tol_is_relative; /* 1: the threshold is tol * ||b|| */
  int32_t check_every; /* FUSED: iterations per replayed graph (0: auto) */
  double *rhistory;    /* optional host array [rhistory_len] of residuals */
  int64_t rhistory_len;
  double omega;        /* solver::set_omega: SOR relaxation (default 1.9) */
} mgx_solve_params;

typedef struct mgx_solve_result {
  int32_t status;      /* MONOLISH_SOLVER_* */
  int32_t pad;
  int64_t final_iter;  /* solver::get_final_iter */
  double final_resid;  /* solver::get_final_residual */
  double seconds;      /* wall time of the solve call */
  int64_t inner_iter;  /* mixed precision: total inner iterations */
  int64_t outer_iter;  /* mixed precision: refinement steps */
} mgx_solve_result;

void mgx_solve_params_default(mgx_solve_params *p);

/* equation::CG<CRS<T>,T>::solve (src/equation/cg.cpp:12-116) */
int mgx_cg_solve(mgx_csr *A, mgx_vec *x, mgx_vec *b, const mgx_solve_params *p,
                 mgx_solve_result *res);
/* equation::BiCGSTAB<CRS<T>,T>::solve (src/equation/bicgstab.cpp:12-166) */
int mgx_bicgstab_solve(mgx_csr *A, mgx_vec *x, mgx_vec *b,
                       const mgx_solve_params *p, mgx_solve_result *res);
/* equation::Jacobi<CRS<T>,T>::solve (src/equation/jacobi.cpp:83-183) */
int mgx_jacobi_solve(mgx_csr *A, mgx_vec *x, mgx_vec *b,
                     const mgx_solve_params *p, mgx_solve_result *res);
/* equation::SOR<CRS<T>,T>::solve (src/equation/sor.cpp:91-164): x += (D/omega
 * + L)^-1 (b - A M x) with the bound preconditioner M, relative stopping
 * test ||b - A M x|| / ||b|| < tol; the lower sweep (sor_kernel_lower,
 * src/equation/kernel/sor_kernel.hpp:5-22) runs level by level on the device
 * with the serial loop's bits */
int mgx_sor_solve(mgx_csr *A, mgx_vec *x, mgx_vec *b,
                  const mgx_solve_params *p, mgx_solve_result *res);
/* equation::IC<CRS<T>,T>::solve / equation::ILU<CRS<T>,T>::solve
 * (src/equation/ic.cpp:140-235, ilu.cpp:145-240): factor a copy of A's
 * values (IC(0) / ILU(0), as MGX_PRECOND_IC / _ILU) and apply it once,
 * x = (L L^T)^-1 b or (L U)^-1 b.  p->precond is ignored; res reports
 * MGX_OK with final_iter 0.  Errors as for the preconditioners */
int mgx_ic_solve(mgx_csr *A, mgx_vec *x, mgx_vec *b,
                 const mgx_solve_params *p, mgx_solve_result *res);
int mgx_ilu_solve(mgx_csr *A, mgx_vec *x, mgx_vec *b,
                  const mgx_solve_params *p, mgx_solve_result *res);
/* standard_eigen::LOBPCG<CRS<T>,T>::solve (src/eigen/lobpcg_standard.cpp:
 * 10-300) with initvec_scheme::USER: the m smallest eigenpairs of symmetric
 * A (1 <= m <= 10, 3m <= n).  X holds m * n entries, row i = vector i: the
 * initial vectors on entry, the eigenvectors on return; lambda (m host
 * doubles) receives the eigenvalues.  Stops when the largest residual norm
 * ||M (A x_i - lambda_i x_i)|| (M the preconditioner p->precond) is below
 * p->tol (absolute) after p->miniter iterations; p->rhistory receives that
 * largest residual per iteration.  The Gram matrices and the basis updates
 * run on the device, the 3m x 3m eigenproblem on the host (the reference
 * calls LAPACK sygvd there).  A Gram matrix that is not positive definite
 * within the {w, x} block is an error (the reference loops) */
int mgx_lobpcg_solve(mgx_csr *A, mgx_vec *X, int64_t m, double *lambda,
                     const mgx_solve_params *p, mgx_solve_result *res);
/* util::get_residual_l2 (src/utils/equation_utils.cpp:45-56): ||b - A x|| */
int mgx_residual_l2(mgx_csr *A, mgx_vec *x, mgx_vec *b, double *out);

/* Mixed-precision iterative refinement around Jacobi-PCG: FP64 A, x, b; the
 * inner solves run on FP32 values and vectors with FP32 accumulation, the
 * outer residual r = b - A x, the update and the stopping test
 * ||r|| < tol (or tol*||b||) are FP64.  No monolish counterpart: it is the
 * precision form of equation::CG (cg.cpp) for FP64 CRS matrices.
 * inner_tol is the relative reduction each inner solve aims for. */
int mgx_ir_cg_solve(mgx_csr *A, mgx_vec *x, mgx_vec *b,
                    const mgx_solve_params *p, double inner_tol,
                    int64_t max_outer, mgx_solve_result *res);

/* Mixed-precision Jacobi-PCG with FP64 residual replacement: the CG
 * iterations run on FP32 values and vectors with FP32 accumulation in the
 * SpMV; for systems of at least 2^19 rows with a sliced layout, their dot
 * products multiply and add in FP64 and q = A p adds a per-row correction
 * of the FP32-rounded diagonal (FP32 accumulation otherwise); each
 * time the FP32 residual has dropped by the factor delta since the last
 * replacement, the FP32 correction is added to the FP64 x and the FP64
 * residual r = b - A x replaces the FP32 one, keeping the search direction.
 * The stopping test ||r|| < tol (or tol*||b||) is made on that FP64
 * residual.  No monolish counterpart: a precision form of equation::CG
 * (cg.cpp).  res: inner_iter = FP32 iterations, outer_iter = replacements. */
int mgx_mixed_cg_solve(mgx_csr *A, mgx_vec *x, mgx_vec *b,
                       const mgx_solve_params *p, double delta,
                       int64_t max_updates, mgx_solve_result *res);

/* ---- batched solves ------------------------------------------------------ */
#define MGX_METHOD_CG 0
#define MGX_METHOD_BICGSTAB 1

/* equation::CG / BiCGSTAB ::solve over nsys independent systems at once
 * (cg.cpp:12-86, bicgstab.cpp:12-135): A[i] x[i] = b[i], one fused launch
 * per kernel stage covers every running system.  Each system follows exactly
 * the iteration it would follow alone (per-system reductions in tile order),
 * params apply to each system (tol_is_relative: tol * ||b[i]||) and
 * results[i] receives its outcome.  Returns MGX_OK once every system has
 * stopped (successfully or not); ABI errors as usual. */
int mgx_batch_solve(mgx_ctx *ctx, int method, int64_t nsys, mgx_csr *const *A,
                    mgx_vec *const *x, mgx_vec *const *b,
                    const mgx_solve_params *p, mgx_solve_result *results);

/* FP32-refined form of mgx_batch_solve for CG on FP64 systems (no monolish
 * counterpart: the precision form of equation::CG, cg.cpp:12-86, over many
 * systems): iterative refinement with FP64 residuals r = b - A x, updates
 * and stopping tests (||r|| < tol, or tol * ||b|| when tol_is_relative),
 * and the corrections of all running systems solved together by the
 * batched FP32 PCG (p->precond) to the accuracy inner_tol.  Each system
 * follows the arithmetic of mgx_ir_cg_solve.  results[i]: final_iter and
 * inner_iter = FP32 iterations, outer_iter = refinement steps, final_resid
 * = FP64 ||r||; p->maxiter bounds the inner iterations per system. */
int mgx_batch_ir_cg_solve(mgx_ctx *ctx, int64_t nsys, mgx_csr *const *A, mgx_vec *const *x,
                          mgx_vec *const *b, const mgx_solve_params *p, double inner_tol,
                          int64_t max_outer, mgx_solve_result *results);

/* ---- solve server: continuous batching of arriving requests -------------- */
typedef struct mgx_server mgx_server;
typedef struct mgx_server_config {
  size_t hbm_budget;        /* bytes for resident requests (0: 3/4 of free) */
  int32_t max_active;       /* systems iterated together (0: 1024) */
  int32_t iters_per_poll;   /* fused iterations between status checks (0: 8) */
  double ir_inner_tol;      /* FP32-refined requests: accuracy of each FP32
                               correction (0: 1e-5) */
  int32_t ir_max_outer;     /* FP32-refined requests: refinement steps (0: 100) */
  int32_t pad_;
} mgx_server_config;

typedef struct mgx_server_stats {
  int64_t submitted, completed, active, queued;
  double p50_latency_s, p99_latency_s, mean_latency_s, max_latency_s;
  double mean_queue_s;      /* submit -> admission */
  double solves_per_s;      /* completed / (last completion - first submit) */
  size_t peak_bytes;        /* peak resident bytes (matrices + work) */
  int64_t cancelled;        /* requests ended by mgx_server_cancel */
  int64_t expired;          /* requests ended by their deadline */
} mgx_server_stats;

/* A worker thread on the context's stream admits requests in arrival order
 * while slots and the HBM budget allow, iterates all resident systems
 * together, and retires each as soon as it stops.  The context must not be
 * used by other calls while the server exists. */
int mgx_server_create(mgx_ctx *ctx, const mgx_server_config *cfg,
                      mgx_server **out);
/* copies the FP64 system (CRS arrays, b, optional x0) and queues it;
 * method MGX_METHOD_CG or MGX_METHOD_BICGSTAB; *ticket identifies it.
 * p->mode MGX_MODE_MIXED_IR (CG only) serves it by FP32-refined iterations:
 * FP64 residual, stop test and update per refinement step, the FP32
 * correction iterated together with the other requests' corrections; each
 * such request follows mgx_ir_cg_solve's arithmetic (result inner_iter /
 * outer_iter as there). */
int mgx_server_submit(mgx_server *s, int method, int64_t n, int64_t nnz,
                      const int32_t *row_ptr, const int32_t *col_ind,
                      const double *val, const double *b, const double *x0,
                      const mgx_solve_params *p, int64_t *ticket);
/* queues a solve of a device-resident FP64 system without copying it: A
 * and b belong to the caller, may be shared by many requests and must stay
 * alive and unchanged until the request completes.  x (optional, not shared
 * with another pending request) is the initial guess and receives the
 * solution; without x the server starts from zero and keeps only the
 * statistics.  Admission charges the work vectors (and x when the server
 * allocates it) against the HBM budget. */
int mgx_server_submit_resident(mgx_server *s, int method, mgx_csr *A, mgx_vec *b, mgx_vec *x,
                               const mgx_solve_params *p, int64_t *ticket);
/* waits up to timeout_s (< 0: forever) for a request; MGX_ERR_NOT_FOUND on
 * timeout; copies the solution into x (n doubles) when x != NULL and the
 * request was submitted with its arrays (mgx_server_submit), and releases
 * the ticket */
int mgx_server_wait(mgx_server *s, int64_t ticket, double timeout_s,
                    mgx_solve_result *res, double *x);
/* Per-request control (no monolish counterpart: the reference solves one
 * system per call).  mgx_server_cancel ends a queued request at once and a
 * running one at the worker's next poll (status MGX_ERR_CANCELLED; a running
 * request's x holds its last iterate); a request that finished first keeps
 * its result.  mgx_server_set_deadline(s, t, sec) ends the request with
 * MGX_ERR_TIMEOUT if it has not finished sec seconds after its submission
 * (0: no deadline).  mgx_server_poll gives MGX_REQ_QUEUED / _ACTIVE / _DONE.
 * Ended requests are still collected with mgx_server_wait; latency
 * statistics count completed requests only.  Unknown tickets:
 * MGX_ERR_NOT_FOUND. */
#define MGX_REQ_QUEUED 0
#define MGX_REQ_ACTIVE 1
#define MGX_REQ_DONE 2
int mgx_server_cancel(mgx_server *s, int64_t ticket);
int mgx_server_set_deadline(mgx_server *s, int64_t ticket, double seconds);
int mgx_server_poll(mgx_server *s, int64_t ticket, int32_t *state);
int mgx_server_stats_get(mgx_server *s, mgx_server_stats *out);
/* finishes queued and resident requests, stops the worker */
int mgx_server_destroy(mgx_server *s);

/* ---- workload generators (host) ----------------------------------------- */
/* 3D stencil on an nx*ny*nz grid, row = (k*ny + j)*nx + i, ascending columns,
 * Dirichlet truncation.  points = 7: diag 6, off-diag -1; points = 27: diag
 * 26, off-diag -1.  shift is added to the diagonal.  Pass NULL arrays to get
 * nnz only. */
int mgx_gen_stencil3d(int64_t nx, int64_t ny, int64_t nz, int points,
                      double shift, int64_t *nnz, int32_t *row_ptr,
                      int32_t *col_ind, double *val);
/* random CRS: row lengths round(exp(N(mu,sigma))) clipped to [1,maxlen],
 * mu = ln(mean) - sigma^2/2; distinct uniform columns sorted ascending;
 * values U[-1,1).  Two calls: row_ptr only (col/val NULL) gives nnz, then
 * the full fill.  Deterministic in seed and independent of thread count. */
int mgx_gen_random_lognormal(int64_t rows, int64_t cols, double mean,
                             double sigma, int32_t maxlen, uint64_t seed,
                             int32_t *row_ptr, int32_t *col_ind, double *val);
/* util::laplacian_matrix_2D_5p(m, n) (src/utils/create_matrix/
 * laplacian_2D5p.cpp:6-50) directly as CRS; NULL arrays give nnz only */
int mgx_gen_laplacian2d(int64_t m, int64_t n, int64_t *nnz, int32_t *row_ptr,
                        int32_t *col_ind, double *val);
/* 2D convection-diffusion -lap(u) + peclet du/dx on an m x m grid, upwind:
 * diag 4 + c, west -1 - c, east/south/north -1, c = peclet / (m + 1) */
int mgx_gen_convdiff2d(int64_t m, double peclet, int64_t *nnz, int32_t *row_ptr,
                       int32_t *col_ind, double *val);
/* SPD M-matrix: kmin..kmax random off-diagonals U[-1,0) per row,
 * symmetrised (A + A^T)/2, diag = sum |off-diagonal| + shift.  Two calls as
 * for mgx_gen_random_lognormal. */
int mgx_gen_mmatrix(int64_t n, int32_t kmin, int32_t kmax, double shift,
                    uint64_t seed, int32_t *row_ptr, int32_t *col_ind,
                    double *val);

/* vector<T>(N, min, max, seed) / util::random_vector
 * (src/utils/vector_constructor.cpp:78-113): std::mt19937(seed) and
 * std::uniform_real_distribution<double>(min, max) drawn in order into the
 * host array out (FP32: float bounds, each draw rounded to float). */
int mgx_gen_random_vector(int64_t n, double min, double max, uint32_t seed,
                          mgx_dtype dtype, void *out);

#ifdef __cplusplus
}
#endif
#endif
