/*
 * oracle.h -- CPU restatement of monolish's reference (OSS: OpenMP + CBLAS)
 * path for the CRS sparse BLAS and Krylov loop.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (monolish_amd/, include/)
 * may link, load or call this code.  It is used by tests/ as the checker, by
 * __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline leg.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the monolish 0.17.3 source tree).
 *
 * Pinning (how each function is checked against something other than itself;
 * the checks live in tests/test_oracle.py):
 *   or_mt_*, or_random_vector_*   MT19937 known answer (10000th output of the
 *                                 default seed, fixed by the C++ standard) and
 *                                 libstdc++ std::uniform_real_distribution
 *                                 outputs (tests/golden/std_random.json).
 *   or_coo_*, or_crs_from_coo     reference KATs (test/matrix/matrix_common.cpp
 *                                 duplicate-merge, 3x3 arrays), closed form of
 *                                 the 5-point Laplacian, dense brute force.
 *   or_input_mm                   reference fixtures test/test.mtx, test2.mtx.
 *   or_spmv_*                     reference KATs y={60,90,210}, y={10,0,10};
 *                                 dense triple loop (bit-exact for ascending
 *                                 columns); forward-error bound vs exact sums.
 *   or_dot_* / or_nrm2_*          exact (fsum) sums + gamma_n error bound;
 *                                 special cases n<=2 exact; reductions on
 *                                 integer-valued data exact.  Bit-level order
 *                                 of the reference (external CBLAS ddot) is
 *                                 PARITY UNPINNED: the order used here is the
 *                                 documented canonical tile order below.
 *   vector ops                    textbook definitions on special values.
 *   or_cg/or_bicgstab/or_jacobi   reference solver KAT (test.mtx, b=A*1 ->
 *                                 x=1 within 10*tol), direct-solve answers
 *                                 (scipy spsolve) on Poisson / convection-
 *                                 diffusion, finite-termination in <= n steps,
 *                                 error paths (NaN, maxiter, breakdown).
 *
 * Canonical reduction order ("tile order"), used for every dot/nrm2 here and
 * by the GPU kernels, so that both sides produce identical bits:
 *   p_i = x_i*y_i (rounded, no FMA).  Tiles of 512 elements; in tile k lane
 *   t (0..255) forms a_t = p[512k+2t] + p[512k+2t+1] (missing terms are 0).
 *   Lanes are grouped in 4 waves of 64; within a wave, for off=32,16,..,1:
 *   v[l] += v[l+off] for l<off.  Tile partial P_k = (W0+W1)+(W2+W3).
 *   Stage 2: lane t sums P_t, P_{t+256}, ... sequentially from 0, then the same
 *   wave tree and (W0+W1)+(W2+W3) combine gives the result.
 */
#ifndef MONOLISH_ORACLE_H
#define MONOLISH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes, reference include/monolish/common/monolish_common.hpp:16-21 */
#define OR_SUCCESS 0
#define OR_SIZE_ERROR -1
#define OR_MAXITER -2
#define OR_BREAKDOWN -3
#define OR_RESIDUAL_NAN -4
#define OR_NOT_IMPL -10
/* IC / ILU set-up failed: unsorted or duplicate columns, a missing
 * diagonal entry or a zero (IC: non-positive) pivot; the reference throws */
#define OR_FACTOR_ERROR -11

#define OR_PRECOND_DEFAULT 0 /* solver without a bound preconditioner: z = r */
#define OR_PRECOND_NONE 1    /* equation::none: z = copy(r) */
#define OR_PRECOND_JACOBI 2  /* equation::Jacobi: z = (1/diag(A)) o r */
#define OR_PRECOND_SOR 3     /* equation::SOR: forward and backward sweeps */
#define OR_PRECOND_IC 4      /* equation::IC: IC(0), L L^T */
#define OR_PRECOND_ILU 5     /* equation::ILU: ILU(0), L U */

/* relaxation coefficient of the SOR preconditioner and solver (the
 * reference's solver default, include/monolish_solver.hpp:48: 1.9) */
void or_set_sor_omega(double w);
double or_get_sor_omega(void);
/* row at which the last IC / ILU set-up failed (-1: none) */
int64_t or_factor_bad_row(void);

/* ---- random numbers: std::mt19937 + std::uniform_real_distribution<double> */
typedef struct {
  uint32_t mt[624];
  int idx;
} or_mt19937;
void or_mt_seed(or_mt19937 *s, uint32_t seed);
uint32_t or_mt_next(or_mt19937 *s);
double or_mt_uniform(or_mt19937 *s, double a, double b);
void or_random_vector_d(double *v, int64_t n, double min, double max,
                        uint32_t seed);
void or_random_vector_s(float *v, int64_t n, float min, float max,
                        uint32_t seed);

/* ---- COO (values held as double; float matrices hold float-exact values) */
typedef struct {
  int64_t rows, cols, nnz, alloc;
  int32_t *row, *col;
  double *val;
} or_coo;
void or_coo_init(or_coo *m, int64_t rows, int64_t cols);
void or_coo_free(or_coo *m);
void or_coo_insert(or_coo *m, int64_t i, int64_t j, double v);
void or_coo_sort(or_coo *m, int merge);
double or_coo_at(const or_coo *m, int64_t i, int64_t j);
/* laplacian_matrix_2D_5p(m, n) */
void or_laplacian_2d5p(or_coo *out, int m, int n);
/* COO::input_mm; is_float selects float parsing of values.  0 on success,
 * negative on a rejected banner / unreadable file. */
int or_input_mm(or_coo *out, const char *path, int is_float);

int64_t or_coo_sort_arrays(int64_t nnz, int32_t *row, int32_t *col,
                           double *val, int merge);

/* COO -> CRS (CRS::convert(COO&)); row_ptr has rows+1 entries */
void or_crs_from_coo(const or_coo *coo, int32_t *row_ptr, int32_t *col_ind,
                     double *val);

/* ---- generic (double "_d" / float "_s") kernels ----------------------- */
#define OR_DECL(T, S)                                                          \
  void or_crs_diag_##S(int64_t rows, int64_t cols, const int32_t *rp,          \
                       const int32_t *ci, const T *val, T *d);                 \
  void or_spmv_##S(int64_t rows, const int32_t *rp, const int32_t *ci,         \
                   const T *val, const T *x, T *y);                            \
  void or_spmv_t_##S(int64_t rows, int64_t cols, const int32_t *rp,            \
                     const int32_t *ci, const T *val, const T *x, T *y);       \
  T or_dot_##S(int64_t n, const T *x, const T *y);                             \
  T or_nrm2_##S(int64_t n, const T *x);                                        \
  void or_axpy_##S(int64_t n, T a, const T *x, T *y);                          \
  void or_xpay_##S(int64_t n, T a, const T *x, T *y);                          \
  void or_axpyz_##S(int64_t n, T a, const T *x, const T *y, T *z);             \
  void or_scal_##S(int64_t n, T a, T *x);                                      \
  void or_copy_##S(int64_t n, const T *x, T *y);                               \
  void or_vadd_##S(int64_t n, const T *a, const T *b, T *y);                   \
  void or_vsub_##S(int64_t n, const T *a, const T *b, T *y);                   \
  void or_vmul_##S(int64_t n, const T *a, const T *b, T *y);                   \
  void or_vreciprocal_##S(int64_t n, const T *a, T *y);                        \
  T or_residual_l2_##S(int64_t n, const int32_t *rp, const int32_t *ci,        \
                       const T *val, const T *x, const T *b);                  \
  int or_cg_##S(int64_t n, const int32_t *rp, const int32_t *ci,               \
                const T *val, T *x, const T *b, int precond, double tol,       \
                int64_t miniter, int64_t maxiter, int64_t *final_iter,         \
                double *final_resid, double *rhist);                           \
  int or_bicgstab_##S(int64_t n, const int32_t *rp, const int32_t *ci,         \
                      const T *val, T *x, const T *b, int precond,             \
                      double tol, int64_t miniter, int64_t maxiter,            \
                      int64_t *final_iter, double *final_resid,                \
                      double *rhist);                                          \
  int or_jacobi_##S(int64_t n, const int32_t *rp, const int32_t *ci,           \
                    const T *val, T *x, const T *b, int precond, double tol,   \
                    int64_t miniter, int64_t maxiter, int64_t *final_iter,     \
                    double *final_resid, double *rhist);                       \
  int or_sor_##S(int64_t n, const int32_t *rp, const int32_t *ci,              \
                 const T *val, T *x, const T *b, int precond, double tol,      \
                 int64_t miniter, int64_t maxiter, int64_t *final_iter,        \
                 double *final_resid, double *rhist);                          \
  void or_sor_precond_##S(int64_t n, const int32_t *rp, const int32_t *ci,     \
                          const T *val, const T *r, T *z);                     \
  int or_ilu_factor_##S(int64_t n, const int32_t *rp, const int32_t *ci,       \
                        const T *val, int kind, T *f);                         \
  int or_ilu_precond_##S(int64_t n, const int32_t *rp, const int32_t *ci,      \
                         const T *val, int kind, const T *r, T *z);

OR_DECL(double, d)
OR_DECL(float, s)
#undef OR_DECL

/* number of OpenMP threads the oracle kernels run with */
int or_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
