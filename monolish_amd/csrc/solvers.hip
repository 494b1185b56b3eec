// solvers.hip -- CG, BiCGSTAB and Jacobi solve loops (src/equation/cg.cpp,
// bicgstab.cpp, jacobi.cpp) and mixed-precision refinement on gfx950.
//
// MGX_MODE_REFERENCE runs the reference's call sequence, one kernel per BLAS
// call, every dot/nrm2 returned to the host (the plain GPU solver).
// MGX_MODE_FUSED runs CG and BiCGSTAB on the fused multi-system engine
// (batch.hip) as a batch of one: CG takes two kernels per iteration and
// BiCGSTAB three, all scalars stay on the device, and the host checks the
// solve status every check_every iterations.  Every value is rounded exactly
// as in the reference loop (same operations, same order, tile-order dots), so
// both modes give identical bits.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

#include "mgx_batch.hpp"
#include "mgx_device.hpp"

namespace mgx {

namespace {

template <class T> T *tp(void *p) { return static_cast<T *>(p); }
template <class T> const T *tp(const void *p) { return static_cast<const T *>(p); }

template <class T> struct DT;
template <> struct DT<double> { static constexpr mgx_dtype v = MGX_F64; };
template <> struct DT<float> { static constexpr mgx_dtype v = MGX_F32; };

template <class T> T host_sqrt(T v) { return std::sqrt(v); }

}  // namespace

// row chunks for the fused kernels: <= budget entries, never crossing a
// 128-row quarter of a 512-row tile (one fused-SpMV block per quarter);
// rows longer than the budget become long chunks.  d_tile_chunk[q] is the
// first chunk of quarter q (4 per tile, empty past the last row).
void ensure_val32(Csr *A) {
  if (A->val32) return;
  Ctx *c = A->ctx;
  // padded like col_ind/val: vector loads may read past the last entry
  A->val32 = (float *)dev_alloc((size_t)(A->nnz + 4) * 4);
  MGX_HIP(hipMemsetAsync(A->val32 + A->nnz, 0, 16, c->stream));
  launch_convert(c, MGX_F64, MGX_F32, A->nnz, A->val, A->val32);
  spmv_sell_values(A);  // the sliced copy of the FP32 values, if A has one
  diag_corr_build(A);
}

namespace {
// lane per row: the sum over the row's diagonal entries of a - (double)(float)a
__global__ void k_diag_corr(int64_t n, const int32_t *rp, const int32_t *ci, const double *val,
                            float *dc) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    double s = 0;
    for (int32_t e = rp[i]; e < rp[i + 1]; e++)
      if (ci[e] == (int32_t)i) s += val[e] - (double)(float)val[e];
    dc[i] = (float)s;
  }
}
}  // namespace

void diag_corr_build(Csr *A) {
  if (A->dtype != MGX_F64) return;
  const int64_t n = A->rows;
  if (!A->dcorr32) A->dcorr32 = (float *)dev_alloc((size_t)std::max<int64_t>(n, 1) * 4);
  if (n == 0) return;
  const int64_t g = std::min<int64_t>((n + kBlock - 1) / kBlock, (int64_t)A->ctx->num_cus * 16);
  k_diag_corr<<<(unsigned)g, kBlock, 0, A->ctx->stream>>>(
      n, A->rp, A->ci, static_cast<const double *>(A->val), A->dcorr32);
  MGX_HIP(hipGetLastError());
}

void fused_plan_build(Csr *A, const int32_t *rp_host) {
  if (A->d_chunks) return;
  const int64_t n = A->rows;
  std::vector<int32_t> rp_copy;
  const int32_t *rp = rp_host;
  if (!rp) {
    rp_copy.resize(n + 1);
    MGX_HIP(hipMemcpy(rp_copy.data(), A->rp, (n + 1) * 4, hipMemcpyDeviceToHost));
    rp = rp_copy.data();
  }
  const int budget = kFusedNnz;
  const int64_t nq = 4 * ntiles_of(n);
  std::vector<int32_t> chunks, tile_chunk(nq + 1);
  for (int64_t k = 0; k < nq; k++) {
    tile_chunk[k] = (int32_t)chunks.size();
    const int64_t R1 = std::min(n, (k + 1) * (int64_t)kQuarter);
    int64_t r = std::min(n, k * kQuarter);
    while (r < R1) {
      int32_t len = rp[r + 1] - rp[r];
      if (len > budget) {
        chunks.push_back((int32_t)r | (int32_t)0x80000000u);
        r++;
        continue;
      }
      const int64_t s = r;
      int64_t acc = 0;
      while (r < R1) {
        len = rp[r + 1] - rp[r];
        if (len > budget || r - s >= kBlock || acc + len > budget) break;
        acc += len;
        r++;
      }
      chunks.push_back((int32_t)s);
    }
  }
  tile_chunk[nq] = (int32_t)chunks.size();
  chunks.push_back((int32_t)n);
  spmv_engine_offsets(A, rp);
  A->d_chunks = (int32_t *)dev_alloc(chunks.size() * 4);
  A->d_tile_chunk = (int32_t *)dev_alloc(tile_chunk.size() * 4);
  MGX_HIP(hipMemcpy(A->d_chunks, chunks.data(), chunks.size() * 4, hipMemcpyHostToDevice));
  MGX_HIP(hipMemcpy(A->d_tile_chunk, tile_chunk.data(), tile_chunk.size() * 4,
                    hipMemcpyHostToDevice));
  A->fused_nnz = budget;
}

namespace {

struct SolveOut {
  int status = MGX_SOLVER_MAXITER;
  int64_t iters = 0;
  double resid = 0;
};

// a preconditioner choice: MGX_PRECOND_* and the SOR relaxation coefficient
struct PrecSpec {
  int kind;
  double omega;
};

template <class T> struct Prec {
  int kind = MGX_PRECOND_DEFAULT;
  T *M = nullptr;  // Jacobi: 1/diag(A); SOR: 1/(omega diag(A)); IC / ILU: the factor (nnz)
  T *t = nullptr;  // IC / ILU: intermediate vector of the two triangular solves
  Csr *A = nullptr;
  TriOp fwd, bwd;  // SOR / IC / ILU: the forward and backward sweeps
};

// Jacobi::create_precond (jacobi.cpp:12-29): M = 1/diag(A); SOR::
// create_precond (sor.cpp:12-35): M = 1/(omega diag(A)) (omega in T);
// IC / ILU::create_precond (ic.cpp:12-80, ilu.cpp:12-85): M = a copy of
// the values, factored in place; none / default: nothing
template <class T>
Prec<T> make_precond(Csr *A, const void *val, PrecSpec spec) {
  Prec<T> P;
  P.kind = spec.kind;
  P.A = A;
  Ctx *c = A->ctx;
  const mgx_dtype dt = DT<T>::v;
  if (spec.kind == MGX_PRECOND_IC || spec.kind == MGX_PRECOND_ILU) {
    P.M = (T *)ws_get(c, (size_t)std::max<int64_t>(A->nnz, 1) * sizeof(T));
    P.t = (T *)ws_get(c, (size_t)std::max<int64_t>(A->rows, 1) * sizeof(T));
    launch_copy(c, dt, A->nnz, val, P.M);
    ilu_factor(A, dt, spec.kind, P.M);
    const int32_t *diag = ilu_diag(A);
    if (spec.kind == MGX_PRECOND_IC) {
      P.fwd = tri_prepare(A, tri_plan(A, kTriLower), dt, kTriDiv, P.M, P.M, diag);
      P.bwd = tri_prepare(A, A->tri[kTriLt], dt, kTriDiv, P.M, P.M, diag);
    } else {
      P.fwd = tri_prepare(A, tri_plan(A, kTriLower), dt, kTriUnit, P.M, P.M, diag);
      P.bwd = tri_prepare(A, tri_plan(A, kTriUpper), dt, kTriDiv, P.M, P.M, diag);
    }
    return P;
  }
  if (spec.kind != MGX_PRECOND_JACOBI && spec.kind != MGX_PRECOND_SOR) return P;
  P.M = (T *)ws_get(c, (size_t)A->rows * sizeof(T));
  launch_diag_v(A, val, DT<T>::v, P.M);
  if (spec.kind == MGX_PRECOND_SOR)
    launch_scal(c, DT<T>::v, A->rows, (double)(T)spec.omega, P.M);
  launch_reciprocal(c, DT<T>::v, A->rows, P.M, P.M);
  if (spec.kind == MGX_PRECOND_SOR) {
    P.fwd = tri_prepare(A, tri_plan(A, kTriLower), dt, kTriMul, val, P.M, nullptr);
    P.bwd = tri_prepare(A, tri_plan(A, kTriUpper), dt, kTriSorBack, val, P.M, nullptr);
  }
  return P;
}

// apply_precond: Jacobi z = M o r (jacobi.cpp:53-58); SOR the forward and
// backward sweeps of sor_kernel_precond (sor_kernel.hpp:41-73); IC / ILU
// the two triangular solves (ic.cpp:100-140, ilu.cpp:105-145); none
// copy(r, z) (none.cpp:40-46); default z = r (monolish_solver.hpp:246-251)
template <class T>
void apply_precond(Ctx *c, const Prec<T> &P, int64_t n, const T *r, T *z) {
  if (P.kind == MGX_PRECOND_JACOBI) {
    launch_vv(c, DT<T>::v, 2, n, P.M, r, z);
  } else if (P.kind == MGX_PRECOND_SOR) {
    tri_run(P.A, P.fwd, DT<T>::v, r, z);
    tri_run(P.A, P.bwd, DT<T>::v, z, z);
  } else if (P.kind == MGX_PRECOND_IC || P.kind == MGX_PRECOND_ILU) {
    tri_run(P.A, P.fwd, DT<T>::v, r, P.t);
    tri_run(P.A, P.bwd, DT<T>::v, P.t, z);
  } else {
    launch_copy(c, DT<T>::v, n, r, z);
  }
}

template <class T> T dot_t(Ctx *c, int64_t n, const T *x, const T *y) {
  return (T)dot_host(c, DT<T>::v, n, x, y);
}

// equation::CG, reference call sequence (cg.cpp:27-85)
template <class T>
SolveOut cg_reference(Csr *A, const void *val, T *x, const T *b, PrecSpec precond,
                      double tol, int64_t miniter, int64_t maxiter,
                      double *rhist, int64_t rhist_len) {
  Ctx *c = A->ctx;
  const int64_t n = A->rows;
  const mgx_dtype dt = DT<T>::v;
  const size_t bytes = (size_t)std::max<int64_t>(n, 1) * sizeof(T);
  T *r = (T *)ws_get(c, bytes), *p = (T *)ws_get(c, bytes),
    *q = (T *)ws_get(c, bytes), *z = (T *)ws_get(c, bytes);
  for (T *v : {r, p, q, z}) launch_fill(c, dt, v, n, 0.0);
  const Prec<T> M = make_precond<T>(A, val, precond);
  SolveOut o;
  launch_spmv(A, val, dt, x, q, 1, 0, false);  // cg.cpp:39
  launch_vv(c, dt, 1, n, b, q, r);             // cg.cpp:40
  apply_precond(c, M, n, r, z);                // cg.cpp:43
  launch_copy(c, dt, n, z, p);                 // cg.cpp:44
  for (int64_t iter = 0; iter < maxiter; iter++) {
    launch_spmv(A, val, dt, p, q, 1, 0, false);  // cg.cpp:47
    T tmp = dot_t(c, n, z, r);                   // cg.cpp:49
    T pq = dot_t(c, n, p, q);
    T alpha = tmp / pq;                          // cg.cpp:50
    launch_axpy(c, dt, n, alpha, p, x);          // cg.cpp:52
    launch_axpy(c, dt, n, -alpha, q, r);         // cg.cpp:54
    apply_precond(c, M, n, r, z);                // cg.cpp:56
    T zr = dot_t(c, n, z, r);
    T beta = zr / tmp;                           // cg.cpp:57
    launch_xpay(c, dt, n, beta, z, p);           // cg.cpp:59
    T resid = host_sqrt(dot_t(c, n, r, r));      // cg.cpp:61
    if (rhist && iter < rhist_len) rhist[iter] = (double)resid;
    if ((double)resid < tol && miniter <= iter + 1) {  // cg.cpp:67
      o.status = MGX_OK;
      o.iters = iter + 1;
      o.resid = (double)resid;
      return o;
    }
    if (std::isnan(resid)) {  // cg.cpp:74
      o.status = MGX_SOLVER_RESIDUAL_NAN;
      o.iters = iter + 1;
      o.resid = (double)resid;
      return o;
    }
  }
  o.status = MGX_SOLVER_MAXITER;  // cg.cpp:82
  o.iters = maxiter;
  o.resid = (double)host_sqrt(dot_t(c, n, r, r));
  return o;
}

// one-system run of the fused engine (CG or BiCGSTAB); identical arithmetic
// to the reference-shaped loops
template <class T>
SolveOut fused_solve(int method, Csr *A, const void *val, T *x, const T *b, int precond,
                     double tol, int64_t miniter, int64_t maxiter, int check_every,
                     double *rhist_host, int64_t rhist_len) {
  Ctx *c = A->ctx;
  Engine<T> &eng = ctx_engine<T>(c);
  double *rh = nullptr;
  const int64_t rlen = rhist_host ? std::min(rhist_len, maxiter) : 0;
  if (rlen > 0) rh = (double *)ws_get(c, (size_t)rlen * sizeof(double));
  SolveCfg cfg;
  cfg.method = method;
  cfg.precond = precond;
  cfg.tol = tol;
  cfg.miniter = miniter;
  cfg.maxiter = maxiter;
  cfg.rhist = rh;
  cfg.rhist_len = rlen;
  const int slot = eng.add(A, val, x, b, cfg);
  MGX_REQUIRE(slot >= 0, "fused solve: out of device memory for the work vectors");
  if (eng.persistent_ok(slot)) {
    eng.run_persistent(slot);
  } else {
    const int every = check_every > 0 ? check_every : 32;
    for (;;) {
      eng.step(every);
      if (!eng.poll().empty()) break;
    }
  }
  SolveState hs = eng.state(slot);
  eng.release(slot);
  if (rlen > 0) {
    const int64_t got = std::min<int64_t>(hs.iter, rlen);
    if (got > 0) MGX_HIP(hipMemcpy(rhist_host, rh, (size_t)got * 8, hipMemcpyDeviceToHost));
  }
  SolveOut o;
  o.status = hs.status;
  o.iters = hs.iter;
  o.resid = hs.resid;
  return o;
}

template <class T>
SolveOut cg_fused(Csr *A, const void *val, T *x, const T *b, int precond, double tol,
                  int64_t miniter, int64_t maxiter, int check_every, double *rhist_host,
                  int64_t rhist_len) {
  if (A->rows == 0 || maxiter <= 0)
    return cg_reference<T>(A, val, x, b, PrecSpec{precond, 1.9}, tol, miniter, maxiter,
                           rhist_host, rhist_len);
  return fused_solve<T>(kMethodCG, A, val, x, b, precond, tol, miniter, maxiter, check_every,
                        rhist_host, rhist_len);
}

// equation::BiCGSTAB, reference call sequence (bicgstab.cpp:27-134)
template <class T>
SolveOut bicgstab_reference(Csr *A, T *x, const T *b, PrecSpec precond, double tol,
                            int64_t miniter, int64_t maxiter, double *rhist,
                            int64_t rhist_len) {
  Ctx *c = A->ctx;
  const int64_t n = A->rows;
  const mgx_dtype dt = DT<T>::v;
  const size_t bytes = (size_t)std::max<int64_t>(n, 1) * sizeof(T);
  T *r = (T *)ws_get(c, bytes), *r0 = (T *)ws_get(c, bytes),
    *p = (T *)ws_get(c, bytes), *phat = (T *)ws_get(c, bytes),
    *s = (T *)ws_get(c, bytes), *shat = (T *)ws_get(c, bytes),
    *v = (T *)ws_get(c, bytes), *t = (T *)ws_get(c, bytes);
  for (T *w : {r, r0, p, phat, s, shat, v, t}) launch_fill(c, dt, w, n, 0.0);
  T rho_old = 1, rho = 1, alpha = 1, beta, omega = 1;
  const Prec<T> M = make_precond<T>(A, A->val, precond);
  SolveOut o;
  auto nrm = [&](const T *w) { return host_sqrt(dot_t(c, n, w, w)); };
  launch_spmv(A, A->val, dt, x, r, 1, 0, false);  // :47
  launch_vv(c, dt, 1, n, b, r, r);                // :48
  launch_copy(c, dt, n, r, r0);                   // :51
  for (int64_t iter = 0; iter < maxiter; iter++) {
    rho = dot_t(c, n, r, r0);  // :56
    if (rho == T(0)) {         // :58
      o.status = MGX_SOLVER_BREAKDOWN;
      o.iters = iter + 1;
      o.resid = (double)nrm(r);
      return o;
    }
    if (iter == 0) {
      launch_copy(c, dt, n, r, p);  // :66
    } else {
      beta = (rho / rho_old) * (alpha / omega);  // :69
      launch_axpy(c, dt, n, -omega, v, p);       // :72
      launch_xpay(c, dt, n, beta, r, p);         // :73
    }
    apply_precond(c, M, n, p, phat);                 // :77
    launch_spmv(A, A->val, dt, phat, v, 1, 0, false);  // :79
    alpha = rho / dot_t(c, n, v, r0);                // :80
    launch_axpyz(c, dt, n, -alpha, v, r, s);         // :83
    apply_precond(c, M, n, s, shat);                 // :86
    launch_spmv(A, A->val, dt, shat, t, 1, 0, false);  // :88
    T ts = dot_t(c, n, t, s);
    T tt = dot_t(c, n, t, t);
    omega = ts / tt;            // :91
    if (omega == T(0)) {        // :93
      o.status = MGX_SOLVER_BREAKDOWN;
      o.iters = iter + 1;
      o.resid = (double)nrm(r);
      return o;
    }
    launch_axpy(c, dt, n, alpha, phat, x);     // :101
    launch_axpy(c, dt, n, omega, shat, x);     // :102
    launch_axpyz(c, dt, n, -omega, t, s, r);   // :105
    T resid = nrm(r);                          // :108
    if (rhist && iter < rhist_len) rhist[iter] = (double)resid;
    if ((double)resid < tol && miniter <= iter + 1) {  // :114
      o.status = MGX_OK;
      o.iters = iter + 1;
      o.resid = (double)resid;
      return o;
    }
    if (std::isnan(resid)) {  // :121
      o.status = MGX_SOLVER_RESIDUAL_NAN;
      o.iters = iter + 1;
      o.resid = (double)resid;
      return o;
    }
    rho_old = rho;  // :128
  }
  o.status = MGX_SOLVER_MAXITER;
  o.iters = maxiter;
  o.resid = (double)nrm(r);
  return o;
}

// equation::Jacobi solver, reference call sequence (jacobi.cpp:96-149)
template <class T>
SolveOut jacobi_reference(Csr *A, T *x, const T *b, PrecSpec precond, double tol,
                          int64_t miniter, int64_t maxiter, double *rhist,
                          int64_t rhist_len) {
  Ctx *c = A->ctx;
  const int64_t n = A->rows;
  const mgx_dtype dt = DT<T>::v;
  const size_t bytes = (size_t)std::max<int64_t>(n, 1) * sizeof(T);
  T *r = (T *)ws_get(c, bytes), *t = (T *)ws_get(c, bytes),
    *s = (T *)ws_get(c, bytes), *d = (T *)ws_get(c, bytes);
  for (T *w : {r, t, s, d}) launch_fill(c, dt, w, n, 0.0);
  auto nrm = [&](const T *w) { return host_sqrt(dot_t(c, n, w, w)); };
  T bnrm2 = nrm(b);                       // :104
  bnrm2 = (T)(1.0 / (double)bnrm2);       // :105
  T nrm2 = 0;
  launch_diag(A, d);                      // :108
  launch_reciprocal(c, dt, n, d, d);      // :109
  const Prec<T> M = make_precond<T>(A, A->val, precond);  // :111
  SolveOut o;
  for (int64_t iter = 0; iter < maxiter; iter++) {
    apply_precond(c, M, n, x, s);                 // :116
    launch_copy(c, dt, n, x, s);                  // :117
    launch_spmv(A, A->val, dt, s, t, 1, 0, false);  // :118
    launch_axpyz(c, dt, n, -1.0, t, b, r);        // :119
    nrm2 = nrm(r);                                // :120
    launch_vv(c, dt, 2, n, r, d, r);              // :121
    launch_vv(c, dt, 0, n, x, r, x);              // :122
    nrm2 = nrm2 * bnrm2;                          // :124
    if (rhist && iter < rhist_len) rhist[iter] = (double)nrm2;
    if ((double)nrm2 < tol && miniter <= iter + 1) {  // :131
      o.status = MGX_OK;
      o.iters = iter + 1;
      o.resid = (double)nrm(r);
      return o;
    }
    if (std::isnan(nrm2)) {  // :138
      o.status = MGX_SOLVER_RESIDUAL_NAN;
      o.iters = iter + 1;
      o.resid = (double)nrm(r);
      return o;
    }
  }
  o.status = MGX_SOLVER_MAXITER;
  o.iters = maxiter;
  o.resid = (double)nrm(r);
  return o;
}

// equation::SOR solver, reference call sequence (sor.cpp:91-164): d =
// 1/(diag(A)/omega); per iteration s = M x, t = A s, r = b - t, ||r||, the
// lower sweep t = (D/omega + L)^-1 r (sor_kernel_lower), x += t; relative test
template <class T>
SolveOut sor_reference(Csr *A, T *x, const T *b, PrecSpec precond, double tol, int64_t miniter,
                       int64_t maxiter, double *rhist, int64_t rhist_len) {
  Ctx *c = A->ctx;
  const int64_t n = A->rows;
  const mgx_dtype dt = DT<T>::v;
  const size_t bytes = (size_t)std::max<int64_t>(n, 1) * sizeof(T);
  T *r = (T *)ws_get(c, bytes), *t = (T *)ws_get(c, bytes), *s = (T *)ws_get(c, bytes),
    *d = (T *)ws_get(c, bytes);
  for (T *w : {r, t, s, d}) launch_fill(c, dt, w, n, 0.0);
  auto nrm = [&](const T *w) { return host_sqrt(dot_t(c, n, w, w)); };
  const T w = (T)(1.0 / (double)(T)precond.omega);  // :112
  launch_diag(A, d);                                 // :113
  launch_scal(c, dt, n, (double)w, d);               // :114
  launch_reciprocal(c, dt, n, d, d);                 // :115
  T bnrm2 = nrm(b);                                  // :118
  bnrm2 = (T)(1.0 / (double)bnrm2);                  // :119
  T nrm2 = 0;
  const Prec<T> M = make_precond<T>(A, A->val, precond);  // :122
  const TriOp lower = tri_prepare(A, tri_plan(A, kTriLowerBreak), dt, kTriMul, A->val, d, nullptr);
  SolveOut o;
  for (int64_t iter = 0; iter < maxiter; iter++) {
    apply_precond(c, M, n, x, s);                   // :126
    launch_spmv(A, A->val, dt, s, t, 1, 0, false);  // :127
    launch_axpyz(c, dt, n, -1.0, t, b, r);          // :128
    nrm2 = nrm(r);                                  // :129
    tri_run(A, lower, dt, r, t);                    // :133
    launch_axpy(c, dt, n, 1.0, t, x);               // :137
    nrm2 = nrm2 * bnrm2;                            // :139
    if (rhist && iter < rhist_len) rhist[iter] = (double)nrm2;
    if ((double)nrm2 < tol && miniter <= iter + 1) {  // :145
      o.status = MGX_OK;
      o.iters = iter + 1;
      o.resid = (double)nrm(r);
      return o;
    }
    if (std::isnan(nrm2)) {  // :151
      o.status = MGX_SOLVER_RESIDUAL_NAN;
      o.iters = iter + 1;
      o.resid = (double)nrm(r);
      return o;
    }
  }
  o.status = MGX_SOLVER_MAXITER;
  o.iters = maxiter;
  o.resid = (double)nrm(r);
  return o;
}

int64_t maxiter_of(const mgx_solve_params *p) {
  return p->maxiter <= 0 ? INT64_MAX / 4 : p->maxiter;
}

void check_solve_args(mgx_csr *A, mgx_vec *x, mgx_vec *b, const mgx_solve_params *p,
                      mgx_solve_result *res, const char *w) {
  MGX_REQUIRE(A && x && b && p && res, "%s: NULL argument", w);
  if (A->rows != A->cols) {
    set_error("%s: error, A.row != A.col", w);
    throw Error{MGX_ERR_INVALID};
  }
  MGX_REQUIRE(x->ctx == A->ctx && b->ctx == A->ctx, "%s: context mismatch", w);
  MGX_REQUIRE(x->dtype == A->dtype && b->dtype == A->dtype, "%s: dtype mismatch", w);
  MGX_REQUIRE(x->n == A->rows && b->n == A->rows, "%s: size mismatch", w);
  MGX_REQUIRE(p->precond >= MGX_PRECOND_DEFAULT && p->precond <= MGX_PRECOND_ILU,
              "%s: unknown preconditioner %d", w, p->precond);
  MGX_REQUIRE(p->precond != MGX_PRECOND_SOR || (p->omega > 0 && p->omega < 2),
              "%s: SOR omega must be in (0, 2)", w);
  MGX_REQUIRE(p->mode == MGX_MODE_REFERENCE || p->mode == MGX_MODE_FUSED,
              "%s: unknown mode %d", w, p->mode);
}

double threshold(Ctx *c, const mgx_solve_params *p, mgx_vec *b) {
  if (!p->tol_is_relative) return p->tol;
  double bb = dot_host(c, b->dtype, b->n, b->d, b->d);
  double bn = b->dtype == MGX_F64 ? std::sqrt(bb) : (double)std::sqrt((float)bb);
  return p->tol * bn;
}

template <class F>
int run_solver(mgx_csr *A, mgx_solve_result *res, F body) {
  Ctx *c = A->ctx;
  const double t0 = seconds_now();
  SolveOut o;
  try {
    o = body();
  } catch (...) {
    ws_release_all(c);
    throw;
  }
  ws_release_all(c);
  MGX_HIP(hipStreamSynchronize(c->stream));
  tri_check_stall(A);
  res->status = o.status;
  res->final_iter = o.iters;
  res->final_resid = o.resid;
  res->seconds = seconds_now() - t0;
  res->inner_iter = 0;
  res->outer_iter = 0;
  return o.status;
}

// ---- mixed-precision iterative refinement ---------------------------------
__global__ void k_to_f32_scaled(int64_t n, const double *r, double scale,
                                float *r32) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    r32[i] = (float)(r[i] * scale);
}

__global__ void k_add_scaled_f32(int64_t n, double s, const float *d,
                                 double *x) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    x[i] = x[i] + s * (double)d[i];
}

int stream_grid(Ctx *c, int64_t n) {
  int64_t g = (n + kBlock * 4 - 1) / (kBlock * 4);
  g = std::min<int64_t>(std::max<int64_t>(g, 1), (int64_t)c->num_cus * 16);
  return (int)g;
}

// Outer loop in FP64: r = b - A x, ||r|| checked against thr; inner FP32
// Jacobi-PCG solves A32 d = r / ||r|| to a relative inner_tol; x += ||r|| d.
SolveOut ir_cg(Csr *A, double *x, const double *b, int precond, double thr,
               int64_t max_inner_total, double inner_tol, int64_t max_outer,
               int check_every, int64_t *inner_total, int64_t *outer_steps) {
  Ctx *c = A->ctx;
  const int64_t n = A->rows;
  ensure_val32(A);
  double *r = (double *)ws_get(c, (size_t)std::max<int64_t>(n, 1) * 8);
  float *r32 = (float *)ws_get(c, (size_t)std::max<int64_t>(n, 1) * 4);
  float *d32 = (float *)ws_get(c, (size_t)std::max<int64_t>(n, 1) * 4);
  const int g = stream_grid(c, n);
  auto residual = [&]() {
    launch_spmv(A, A->val, MGX_F64, x, r, 1, 0, false);
    launch_vv(c, MGX_F64, 1, n, b, r, r);
    return std::sqrt(dot_host(c, MGX_F64, n, r, r));
  };
  SolveOut o;
  *inner_total = 0;
  *outer_steps = 0;
  double rn = residual();
  for (;;) {
    o.resid = rn;
    if (rn < thr) {
      o.status = MGX_OK;
      break;
    }
    if (std::isnan(rn)) {
      o.status = MGX_SOLVER_RESIDUAL_NAN;
      break;
    }
    if (*outer_steps >= max_outer || *inner_total >= max_inner_total) {
      o.status = MGX_SOLVER_MAXITER;
      break;
    }
    const size_t mark = c->ws_used.size();
    if (n > 0) {
      k_to_f32_scaled<<<g, kBlock, 0, c->stream>>>(n, r, 1.0 / rn, r32);
      MGX_HIP(hipGetLastError());
    }
    launch_fill(c, MGX_F32, d32, n, 0.0);
    // the scaled residual has unit norm: inner_tol is its absolute target
    SolveOut in = cg_fused<float>(A, A->val32, d32, r32, precond, inner_tol, 0,
                                  max_inner_total - *inner_total, check_every,
                                  nullptr, 0);
    ws_release_to(c, mark);
    *inner_total += in.iters;
    if (n > 0) {
      k_add_scaled_f32<<<g, kBlock, 0, c->stream>>>(n, rn, d32, x);
      MGX_HIP(hipGetLastError());
    }
    (*outer_steps)++;
    rn = residual();
  }
  o.iters = *inner_total;
  return o;
}

// r32 = (float) r; z = M r32 (Jacobi) or r32: the FP32 CG's residual and
// preconditioned residual after a residual replacement
__global__ void k_mp_replace(int64_t n, const double *r, const float *M, float *r32,
                             float *z) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    const float v = (float)r[i];
    r32[i] = v;
    z[i] = M ? M[i] * v : v;
  }
}

// the FP32 CG's state after a residual replacement, written on the device:
// beta = zr_k / zr_{k-1} and zr = zr_k were of the replaced residual, so
// beta is rescaled to the new <z, r>; zr, rr and the residual of the new
// residual; the next tolerance; running, resumed with beta and zr given
__global__ void k_mp_resume(SolveState *st, const double *zr_new, const double *rr_new,
                            double tol) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const double zr = *zr_new, rr = *rr_new;
  st->beta = st->beta * (zr / st->zr);
  st->zr = zr;
  st->rr = rr;
  st->resid = __builtin_sqrt(rr);
  st->tol = tol;
  st->status = kStatRunning;
  st->fresh = 2;
  st->ia = st->iter;
}

// Mixed-precision PCG with FP64 residual replacement ("reliable updates",
// Sleijpen & van der Vorst, Computing 56 (1996); the FP32/FP64 form of
// Clark et al., Comput. Phys. Commun. 181 (2010), alg. 3).  The FP32 PCG
// iterates a correction d32 from the residual r32.  Once its residual has
// dropped by the factor delta since the last replacement, the correction is
// folded into the FP64 x, the FP64 residual r = b - A x replaces r32 and z =
// M r32, and the iteration continues with the same search direction p:
// beta = <z,r>_new / <z,r>_{k-1}.  Unlike a restarted refinement the Krylov
// direction survives each replacement, so the FP32 iteration count stays
// close to the FP64 solver's.
SolveOut mp_cg(Csr *A, double *x, const double *b, int precond, double thr,
               int64_t max_inner_total, double delta, int64_t max_updates, int check_every,
               int64_t *inner_total, int64_t *updates) {
  Ctx *c = A->ctx;
  const int64_t n = A->rows;
  ensure_val32(A);
  double *r = (double *)ws_get(c, (size_t)std::max<int64_t>(n, 1) * 8);
  float *r32 = (float *)ws_get(c, (size_t)std::max<int64_t>(n, 1) * 4);
  float *d32 = (float *)ws_get(c, (size_t)std::max<int64_t>(n, 1) * 4);
  const int g = stream_grid(c, n);
  auto residual = [&]() {
    launch_spmv(A, A->val, MGX_F64, x, r, 1, 0, false);
    launch_vv(c, MGX_F64, 1, n, b, r, r);
    return std::sqrt(dot_host(c, MGX_F64, n, r, r));
  };
  SolveOut o;
  *inner_total = 0;
  *updates = 0;
  double rn = residual();
  o.resid = rn;
  if (rn < thr) {
    o.status = MGX_OK;
    return o;
  }
  if (std::isnan(rn)) {
    o.status = MGX_SOLVER_RESIDUAL_NAN;
    return o;
  }
  if (max_inner_total <= 0) {
    o.status = MGX_SOLVER_MAXITER;
    return o;
  }
  k_to_f32_scaled<<<g, kBlock, 0, c->stream>>>(n, r, 1.0, r32);
  MGX_HIP(hipGetLastError());
  launch_fill(c, MGX_F32, d32, n, 0.0);
  Engine<float> &eng = ctx_engine<float>(c);
  SolveCfg cfg;
  cfg.method = kMethodCG;
  cfg.precond = precond;
  cfg.tol = std::max(delta * rn, thr);
  cfg.miniter = 0;
  cfg.maxiter = max_inner_total;
  // FP64 reductions and the FP64 diagonal: with FP32 dot products, or with
  // the FP32-rounded diagonal (a systematic shift of every eigenvalue on a
  // constant-diagonal stencil), the FP32 iteration needs about 9 % more
  // iterations than the FP64 solver on the 27-point 128^3 system; with
  // both, the same count (CPU simulation of this iteration)
  cfg.acc64 = 1;
  cfg.dcorr = A->dcorr32;
  const int slot = eng.add(A, A->val32, d32, r32, cfg);
  MGX_REQUIRE(slot >= 0, "mixed CG: out of device memory for the work vectors");
  const int every = check_every > 0 ? check_every : 8;
  try {
    for (;;) {
      for (;;) {
        eng.step(every);
        if (!eng.poll().empty()) break;
      }
      const SolveState hs = eng.state(slot);
      *inner_total = hs.iter;
      k_add_scaled_f32<<<g, kBlock, 0, c->stream>>>(n, 1.0, d32, x);
      MGX_HIP(hipGetLastError());
      (*updates)++;
      rn = residual();
      o.resid = rn;
      if (rn < thr) {
        o.status = MGX_OK;
        break;
      }
      if (std::isnan(rn) || hs.status == MGX_SOLVER_RESIDUAL_NAN) {
        o.status = MGX_SOLVER_RESIDUAL_NAN;
        break;
      }
      if (hs.status != MGX_OK || hs.iter >= max_inner_total || *updates >= max_updates) {
        o.status = MGX_SOLVER_MAXITER;
        break;
      }
      // replace the FP32 residual; d32 restarts from 0, p is kept.  The new
      // state is formed on the device (no host round trip)
      const SysDesc<float> &d = eng.desc(slot);
      launch_fill(c, MGX_F32, d32, n, 0.0);
      k_mp_replace<<<g, kBlock, 0, c->stream>>>(n, r, d.M, d.r, d.z);
      MGX_HIP(hipGetLastError());
      launch_dot(c, MGX_F32, n, d.z, d.r, &c->d_state->zr, c->d_tick);
      launch_dot(c, MGX_F32, n, d.r, d.r, &c->d_state->rr, c->d_tick);
      k_mp_resume<<<1, 64, 0, c->stream>>>(eng.device_state(slot), &c->d_state->zr,
                                            &c->d_state->rr, std::max(delta * rn, thr));
      MGX_HIP(hipGetLastError());
      eng.resume_device(slot);
    }
  } catch (...) {
    eng.release(slot);
    throw;
  }
  eng.release(slot);
  o.iters = *inner_total;
  return o;
}

}  // namespace

}  // namespace mgx

using namespace mgx;

namespace {

// ---- LOBPCG (src/eigen/lobpcg_standard.cpp:10-300) ------------------------
// The small generalised eigenproblem Sa v = lambda Sb v (k <= 30) of each
// iteration is solved on the host, as the reference does with LAPACK's
// sygvd (jobz 'V', uplo 'L', itype 1): Cholesky Sb = L L^T, then the
// eigenpairs of C = L^-1 Sa L^-T by Householder tridiagonalisation and the
// implicit QL method (Wilkinson & Reinsch, Handbook for Automatic
// Computation II, tred2 / tql2; the cyclic Jacobi method, Golub & Van Loan,
// Matrix Computations, 4th ed., Alg. 8.5.3, if QL does not converge),
// eigenvalues ascending, v = L^-T q so that v^T Sb v = 1.  Each eigenvector
// is signed so that its largest component is positive.

// in-place Cholesky of the lower triangle of B (row-major k x k); returns 0,
// or the order of the first leading minor that is not positive definite
int chol_lower(int k, std::vector<double> &B) {
  for (int j = 0; j < k; j++) {
    double d = B[j * k + j];
    for (int q = 0; q < j; q++) d -= B[j * k + q] * B[j * k + q];
    if (!(d > 0.0)) return j + 1;
    d = std::sqrt(d);
    B[j * k + j] = d;
    for (int i = j + 1; i < k; i++) {
      double s = B[i * k + j];
      for (int q = 0; q < j; q++) s -= B[i * k + q] * B[j * k + q];
      B[i * k + j] = s / d;
    }
  }
  return 0;
}

// eigenpairs of the symmetric C (full, row-major); w ascending, V row i =
// eigenvector i
void jacobi_eig(int k, std::vector<double> C, std::vector<double> &w, std::vector<double> &V) {
  std::vector<double> Q(k * k, 0.0);
  for (int i = 0; i < k; i++) Q[i * k + i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < k; i++)
      for (int j = 0; j < k; j++) {
        tot += C[i * k + j] * C[i * k + j];
        if (i != j) off += C[i * k + j] * C[i * k + j];
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < k - 1; p++)
      for (int q = p + 1; q < k; q++) {
        const double apq = C[p * k + q];
        if (apq == 0.0) continue;
        const double theta = (C[q * k + q] - C[p * k + p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
        for (int r = 0; r < k; r++) {  // C = J^T C J
          const double crp = C[r * k + p], crq = C[r * k + q];
          C[r * k + p] = cs * crp - sn * crq;
          C[r * k + q] = sn * crp + cs * crq;
        }
        for (int r = 0; r < k; r++) {
          const double cpr = C[p * k + r], cqr = C[q * k + r];
          C[p * k + r] = cs * cpr - sn * cqr;
          C[q * k + r] = sn * cpr + cs * cqr;
        }
        for (int r = 0; r < k; r++) {  // Q = Q J
          const double qrp = Q[r * k + p], qrq = Q[r * k + q];
          Q[r * k + p] = cs * qrp - sn * qrq;
          Q[r * k + q] = sn * qrp + cs * qrq;
        }
      }
  }
  std::vector<int> idx(k);
  for (int i = 0; i < k; i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(),
                   [&](int a, int b) { return C[a * k + a] < C[b * k + b]; });
  w.resize(k);
  V.assign(k * k, 0.0);
  for (int i = 0; i < k; i++) {
    w[i] = C[idx[i] * k + idx[i]];
    for (int r = 0; r < k; r++) V[i * k + r] = Q[r * k + idx[i]];
  }
}

// Householder reduction of the symmetric a (row-major k x k, overwritten by
// the orthogonal Q) to tridiagonal d (diagonal), e (e[i]: entry (i, i-1))
void householder_tridiag(int k, std::vector<double> &a, std::vector<double> &d,
                         std::vector<double> &e) {
  auto A = [&](int i, int j) -> double & { return a[i * k + j]; };
  d.assign(k, 0.0);
  e.assign(k, 0.0);
  for (int i = k - 1; i > 0; i--) {
    const int l = i - 1;
    double h = 0.0;
    if (l > 0) {
      double scale = 0.0;
      for (int q = 0; q <= l; q++) scale += std::fabs(A(i, q));
      if (scale == 0.0) {
        e[i] = A(i, l);
      } else {
        for (int q = 0; q <= l; q++) {
          A(i, q) /= scale;
          h += A(i, q) * A(i, q);
        }
        double f = A(i, l);
        double g = f >= 0.0 ? -std::sqrt(h) : std::sqrt(h);
        e[i] = scale * g;
        h -= f * g;
        A(i, l) = f - g;
        f = 0.0;
        for (int j = 0; j <= l; j++) {
          A(j, i) = A(i, j) / h;
          g = 0.0;
          for (int q = 0; q <= j; q++) g += A(j, q) * A(i, q);
          for (int q = j + 1; q <= l; q++) g += A(q, j) * A(i, q);
          e[j] = g / h;
          f += e[j] * A(i, j);
        }
        const double hh = f / (h + h);
        for (int j = 0; j <= l; j++) {
          f = A(i, j);
          e[j] = g = e[j] - hh * f;
          for (int q = 0; q <= j; q++) A(j, q) -= f * e[q] + g * A(i, q);
        }
      }
    } else {
      e[i] = A(i, l);
    }
    d[i] = h;
  }
  d[0] = 0.0;
  e[0] = 0.0;
  for (int i = 0; i < k; i++) {
    const int l = i - 1;
    if (d[i] != 0.0) {
      for (int j = 0; j <= l; j++) {
        double g = 0.0;
        for (int q = 0; q <= l; q++) g += A(i, q) * A(q, j);
        for (int q = 0; q <= l; q++) A(q, j) -= g * A(q, i);
      }
    }
    d[i] = A(i, i);
    A(i, i) = 1.0;
    for (int j = 0; j <= l; j++) A(j, i) = A(i, j) = 0.0;
  }
}

// implicit QL on the tridiagonal (d, e from householder_tridiag), the
// rotations accumulated into z (columns: eigenvectors); false if an
// eigenvalue needs more than 60 iterations
bool tridiag_ql(int k, std::vector<double> &d, std::vector<double> &e, std::vector<double> &z) {
  auto Z = [&](int i, int j) -> double & { return z[i * k + j]; };
  for (int i = 1; i < k; i++) e[i - 1] = e[i];
  e[k - 1] = 0.0;
  for (int l = 0; l < k; l++) {
    int it = 0, m;
    do {
      for (m = l; m < k - 1; m++) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 1e-16 * dd) break;
      }
      if (m != l) {
        if (it++ == 60) return false;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; i--) {
          double f = s * e[i];
          const double b = c * e[i];
          e[i + 1] = (r = std::hypot(f, g));
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          d[i + 1] = g + (p = s * r);
          g = c * r - b;
          for (int q = 0; q < k; q++) {
            f = Z(q, i + 1);
            Z(q, i + 1) = s * Z(q, i) + c * f;
            Z(q, i) = c * Z(q, i) - s * f;
          }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  return true;
}

// eigenpairs of the symmetric C (full, row-major); w ascending, V row i =
// eigenvector i
bool sym_eig(int k, std::vector<double> C, std::vector<double> &w, std::vector<double> &V) {
  std::vector<double> d, e;
  householder_tridiag(k, C, d, e);
  if (!tridiag_ql(k, d, e, C)) return false;
  std::vector<int> idx(k);
  for (int i = 0; i < k; i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
  w.resize(k);
  V.assign(k * k, 0.0);
  for (int i = 0; i < k; i++) {
    w[i] = d[idx[i]];
    for (int r = 0; r < k; r++) V[i * k + r] = C[r * k + idx[i]];
  }
  return true;
}

// Sa v = lambda Sb v from the lower triangles; returns 0 or the Cholesky
// failure order of Sb
int sygv_lower(int k, const std::vector<double> &Sa, std::vector<double> Sb,
               std::vector<double> &w, std::vector<double> &V) {
  const int q = chol_lower(k, Sb);
  if (q) return q;
  // C = L^-1 Sa L^-T with Sa symmetric from its lower triangle
  std::vector<double> S(k * k), Y(k * k), C(k * k);
  for (int i = 0; i < k; i++)
    for (int j = 0; j < k; j++) S[i * k + j] = i >= j ? Sa[i * k + j] : Sa[j * k + i];
  for (int c = 0; c < k; c++)  // Y = L^-1 S (column by column)
    for (int i = 0; i < k; i++) {
      double s = S[i * k + c];
      for (int r = 0; r < i; r++) s -= Sb[i * k + r] * Y[r * k + c];
      Y[i * k + c] = s / Sb[i * k + i];
    }
  for (int r = 0; r < k; r++)  // C^T = L^-1 Y^T, C symmetric
    for (int i = 0; i < k; i++) {
      double s = Y[r * k + i];
      for (int q2 = 0; q2 < i; q2++) s -= Sb[i * k + q2] * C[r * k + q2];
      C[r * k + i] = s / Sb[i * k + i];
    }
  for (int i = 0; i < k; i++)
    for (int j = 0; j < i; j++) C[i * k + j] = C[j * k + i] = 0.5 * (C[i * k + j] + C[j * k + i]);
  std::vector<double> Qv;
  if (!sym_eig(k, C, w, Qv)) jacobi_eig(k, C, w, Qv);
  V.assign(k * k, 0.0);
  for (int e = 0; e < k; e++) {  // v = L^-T q
    for (int i = k - 1; i >= 0; i--) {
      double s = Qv[e * k + i];
      for (int r = i + 1; r < k; r++) s -= Sb[r * k + i] * V[e * k + r];
      V[e * k + i] = s / Sb[i * k + i];
    }
    int big = 0;
    for (int i = 1; i < k; i++)
      if (std::fabs(V[e * k + i]) > std::fabs(V[e * k + big])) big = i;
    if (V[e * k + big] < 0)
      for (int i = 0; i < k; i++) V[e * k + i] = -V[e * k + i];
  }
  return 0;
}

template <class T>
SolveOut lobpcg_run(Csr *A, T *Xio, int64_t m, double *lambda, PrecSpec precond, double tol,
                    int64_t miniter, int64_t maxiter, double *rhist, int64_t rhist_len) {
  Ctx *c = A->ctx;
  const int64_t n = A->rows;
  const mgx_dtype dt = DT<T>::v;
  const int k3 = (int)(3 * m);
  const size_t blk = (size_t)k3 * n * sizeof(T);
  T *wxp = (T *)ws_get(c, blk), *WXP = (T *)ws_get(c, blk), *wxp_p = (T *)ws_get(c, blk),
    *WXP_p = (T *)ws_get(c, blk);
  T *v1 = (T *)ws_get(c, (size_t)n * sizeof(T)), *v2 = (T *)ws_get(c, (size_t)n * sizeof(T));
  const int64_t ntile = (n + 127) / 128;
  double *part = (double *)ws_get(c, (size_t)ntile * k3 * (k3 + 1) * 8);
  double *gout = (double *)ws_get(c, (size_t)k3 * (k3 + 1) * 8);
  T *coef = (T *)ws_get(c, (size_t)m * k3 * sizeof(T));
  for (T *b : {wxp, WXP, wxp_p, WXP_p}) launch_fill(c, dt, b, (int64_t)k3 * n, 0.0);
  const Prec<T> M = make_precond<T>(A, A->val, precond);  // :24
  auto nrm = [&](const T *v) { return host_sqrt(dot_t(c, n, v, v)); };
  auto row = [&](T *base, int64_t r) { return base + r * n; };
  // initvec_scheme::USER: x = the caller's vectors (:93-97)
  launch_copy(c, dt, m * n, Xio, row(wxp, m));
  for (int64_t i = 0; i < m; i++) {  // :103-113
    T *x = row(wxp, m + i), *X = row(WXP, m + i), *w = row(wxp, i), *W = row(WXP, i);
    launch_spmv(A, A->val, dt, x, X, 1, 0, false);
    const T mu = dot_t(c, n, x, X);
    launch_axpyz(c, dt, n, -(double)mu, x, X, w);
    const T norm = nrm(w);
    launch_scal(c, dt, n, (double)(T)(1.0 / (double)norm), w);
    launch_spmv(A, A->val, dt, w, W, 1, 0, false);
  }
  auto out_x = [&] { launch_copy(c, dt, m * n, row(wxp, m), Xio); };
  // nrm2 of rows first .. first+cnt-1 of wxp: the canonical dots of those
  // rows in one launch and one host read
  const int64_t pst = ntiles_of(n) + 2 * kRedGroups;
  double *d_dots = (double *)ws_get(c, (size_t)2 * m * 8);
  void *d_part = ws_get(c, (size_t)2 * m * pst * sizeof(T));
  uint32_t *d_ticks = (uint32_t *)ws_get(c, (size_t)2 * m * kTickWords * 4);
  MGX_HIP(hipMemsetAsync(d_ticks, 0, (size_t)2 * m * kTickWords * 4, c->stream));
  std::vector<double> h_dots(2 * m);
  auto norms = [&](int64_t first, int64_t cnt) {
    launch_mnrm2sq(c, dt, n, (int)cnt, row(wxp, first), n, d_part, d_ticks, d_dots);
    MGX_HIP(hipMemcpyAsync(h_dots.data(), d_dots, (size_t)cnt * 8, hipMemcpyDeviceToHost,
                           c->stream));
    MGX_HIP(hipStreamSynchronize(c->stream));
    std::vector<T> r(cnt);
    for (int64_t i = 0; i < cnt; i++) r[i] = host_sqrt((T)h_dots[i]);
    return r;
  };
  // the Jacobi / none / default preconditioners are applied inside the
  // fused x, w pass; the triangular ones run their sweeps per pair
  const bool fused_prec = precond.kind <= MGX_PRECOND_JACOBI;
  bool singular = false;
  std::vector<double> g(k3 * (k3 + 1)), w, V;
  std::vector<T> hcoef(m * k3);
  std::vector<T> residual(m);
  T resid = 0;
  SolveOut o;
  for (int64_t iter = 0; iter < maxiter; iter++) {
    const int k = (iter == 0 || singular) ? (int)(2 * m) : k3;  // :136-149
    lobpcg_gram(c, dt, n, k, wxp, WXP, part, gout);              // :156-160
    MGX_HIP(hipMemcpyAsync(g.data(), gout, (size_t)k * (k + 1) * 8, hipMemcpyDeviceToHost,
                           c->stream));
    MGX_HIP(hipStreamSynchronize(c->stream));
    const int npair = k * (k + 1) / 2;
    std::vector<double> Sa(k * k, 0.0), Sb(k * k, 0.0);
    for (int i = 0; i < k; i++)
      for (int j = 0; j <= i; j++) {
        Sa[i * k + j] = g[i * (i + 1) / 2 + j];
        Sb[i * k + j] = g[npair + i * (i + 1) / 2 + j];
      }
    const int q = sygv_lower(k, Sa, Sb, w, V);  // :170-172
    if (q > 0) {
      // a leading minor of Sb past the {w, x} block: the p block lost rank;
      // restart the step without it (:177-186)
      MGX_REQUIRE(q > k - m && k == k3,
                  "LOBPCG: the basis Gram matrix is not positive definite (order %d of %d)", q,
                  k);
      singular = true;
      iter--;
      continue;
    }
    // :197-198 copy the basis to the previous-step buffers; every row of
    // the current buffers is rewritten below (x, p by the combine, w and W
    // after it), so exchanging the buffers gives the same values
    std::swap(wxp, wxp_p);
    std::swap(WXP, WXP_p);
    for (int64_t i = 0; i < m; i++) {
      lambda[i] = (double)(T)w[i];  // :203
      for (int j = 0; j < k3; j++) hcoef[i * k3 + j] = j < k ? (T)V[i * k + j] : (T)0;
    }
    MGX_HIP(hipMemcpyAsync(coef, hcoef.data(), hcoef.size() * sizeof(T), hipMemcpyHostToDevice,
                           c->stream));
    lobpcg_combine(c, dt, n, (int)m, k3, k == k3, coef, wxp_p, WXP_p, wxp, WXP);  // :220-238
    // The pairs are independent within an iteration, so each step below
    // runs for all m pairs at once (one launch, one host read per batch of
    // norms); every value is the one the reference's pair-by-pair order
    // computes.  (On a NaN residual the x of later pairs is already
    // updated; the reference leaves them at the previous step.)
    const std::vector<T> nxp = norms(m, 2 * m);  // x rows, then p rows (:240-246)
    std::vector<double> sp(m), sx(m), nl(m), sw(m);
    for (int64_t i = 0; i < m; i++) {
      sx[i] = (double)(T)(1.0 / (double)nxp[i]);
      sp[i] = (double)(T)(1.0 / (double)nxp[m + i]);
      nl[i] = -(double)(T)lambda[i];
    }
    lobpcg_scale_rows(c, dt, n, (int)m, sp.data(), row(wxp, 2 * m), row(WXP, 2 * m));
    if (fused_prec) {  // :246-253
      lobpcg_xw(c, dt, n, (int)m, sx.data(), nl.data(), row(wxp, m), row(WXP, m), wxp,
                precond.kind == MGX_PRECOND_JACOBI ? M.M : nullptr);
    } else {
      for (int64_t i = 0; i < m; i++) {
        T *x = row(wxp, m + i), *X = row(WXP, m + i), *wv = row(wxp, i);
        launch_scal(c, dt, n, sx[i], x);
        launch_scal(c, dt, n, sx[i], X);
        launch_axpyz(c, dt, n, nl[i], x, X, wv);  // :248
        launch_copy(c, dt, n, wv, v2);            // :250-252
        apply_precond(c, M, n, v2, v1);
        launch_copy(c, dt, n, v1, wv);
      }
    }
    residual = norms(0, m);  // :255
    for (int64_t i = 0; i < m; i++) {
      if (std::isnan(residual[i])) {  // :260-270
        out_x();
        o.status = MGX_SOLVER_RESIDUAL_NAN;
        o.iters = iter + 1;
        o.resid = (double)residual[i];
        return o;
      }
      sw[i] = (double)(T)(1.0 / (double)residual[i]);
    }
    lobpcg_scale_rows(c, dt, n, (int)m, sw.data(), wxp, nullptr);  // :272
    if (!lobpcg_spmm(A, dt, (int)m, wxp, WXP))                        // :274
      for (int64_t i = 0; i < m; i++)
        launch_spmv(A, A->val, dt, row(wxp, i), row(WXP, i), 1, 0, false);
    resid = *std::max_element(residual.begin(), residual.end());  // :277
    if (rhist && iter < rhist_len) rhist[iter] = (double)resid;
    if ((double)resid < tol && miniter < iter + 1) {  // :279-287
      out_x();
      o.status = MGX_OK;
      o.iters = iter + 1;
      o.resid = (double)resid;
      return o;
    }
    singular = false;
  }
  out_x();
  o.status = MGX_SOLVER_MAXITER;
  o.iters = maxiter;
  o.resid = (double)resid;
  return o;
}

}  // namespace

extern "C" {

void mgx_solve_params_default(mgx_solve_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  p->tol = 1.0e-8;  // monolish_solver.hpp:36
  p->maxiter = 0;   // unlimited (SIZE_MAX in the reference)
  p->miniter = 0;
  p->precond = MGX_PRECOND_DEFAULT;
  p->mode = MGX_MODE_FUSED;
  p->omega = 1.9;  // monolish_solver.hpp:48
}

int mgx_cg_solve(mgx_csr *A, mgx_vec *x, mgx_vec *b, const mgx_solve_params *p,
                 mgx_solve_result *res) {
  if (p && p->mode == MGX_MODE_MIXED_IR) return mgx_ir_cg_solve(A, x, b, p, 1e-5, 100, res);
  return guard([&] {
    check_solve_args(A, x, b, p, res, "mgx_cg_solve");
    activate(A->ctx);
    const double thr = threshold(A->ctx, p, b);
    const int64_t mi = maxiter_of(p);
    // the SOR / IC / ILU preconditioners' sweeps run as their own kernels:
    // the CG with them follows the reference call sequence in either mode
    const bool fused = p->mode == MGX_MODE_FUSED && p->precond <= MGX_PRECOND_JACOBI;
    const PrecSpec ps{p->precond, p->omega};
    return run_solver(A, res, [&] {
      if (A->dtype == MGX_F64) {
        if (fused)
          return cg_fused<double>(A, A->val, (double *)x->d, (const double *)b->d, p->precond,
                                  thr, p->miniter, mi, p->check_every, p->rhistory,
                                  p->rhistory_len);
        return cg_reference<double>(A, A->val, (double *)x->d, (const double *)b->d, ps, thr,
                                    p->miniter, mi, p->rhistory, p->rhistory_len);
      }
      if (fused)
        return cg_fused<float>(A, A->val, (float *)x->d, (const float *)b->d, p->precond, thr,
                               p->miniter, mi, p->check_every, p->rhistory, p->rhistory_len);
      return cg_reference<float>(A, A->val, (float *)x->d, (const float *)b->d, ps, thr,
                                 p->miniter, mi, p->rhistory, p->rhistory_len);
    });
  });
}

int mgx_bicgstab_solve(mgx_csr *A, mgx_vec *x, mgx_vec *b, const mgx_solve_params *p,
                       mgx_solve_result *res) {
  return guard([&] {
    check_solve_args(A, x, b, p, res, "mgx_bicgstab_solve");
    activate(A->ctx);
    const double thr = threshold(A->ctx, p, b);
    const int64_t mi = maxiter_of(p);
    const bool fused =
        p->mode == MGX_MODE_FUSED && A->rows > 0 && mi > 0 && p->precond <= MGX_PRECOND_JACOBI;
    const PrecSpec ps{p->precond, p->omega};
    return run_solver(A, res, [&] {
      if (fused && A->dtype == MGX_F64)
        return fused_solve<double>(kMethodBiCGSTAB, A, A->val, (double *)x->d,
                                   (const double *)b->d, p->precond, thr, p->miniter, mi,
                                   p->check_every, p->rhistory, p->rhistory_len);
      if (fused)
        return fused_solve<float>(kMethodBiCGSTAB, A, A->val, (float *)x->d,
                                  (const float *)b->d, p->precond, thr, p->miniter, mi,
                                  p->check_every, p->rhistory, p->rhistory_len);
      if (A->dtype == MGX_F64)
        return bicgstab_reference<double>(A, (double *)x->d, (const double *)b->d, ps, thr,
                                          p->miniter, mi, p->rhistory, p->rhistory_len);
      return bicgstab_reference<float>(A, (float *)x->d, (const float *)b->d, ps, thr,
                                       p->miniter, mi, p->rhistory, p->rhistory_len);
    });
  });
}

int mgx_jacobi_solve(mgx_csr *A, mgx_vec *x, mgx_vec *b, const mgx_solve_params *p,
                     mgx_solve_result *res) {
  return guard([&] {
    check_solve_args(A, x, b, p, res, "mgx_jacobi_solve");
    activate(A->ctx);
    // the Jacobi iteration's own test is relative (jacobi.cpp:124,131)
    const double thr = p->tol;
    const int64_t mi = maxiter_of(p);
    const PrecSpec ps{p->precond, p->omega};
    return run_solver(A, res, [&] {
      if (A->dtype == MGX_F64)
        return jacobi_reference<double>(A, (double *)x->d, (const double *)b->d, ps, thr,
                                        p->miniter, mi, p->rhistory, p->rhistory_len);
      return jacobi_reference<float>(A, (float *)x->d, (const float *)b->d, ps, thr,
                                     p->miniter, mi, p->rhistory, p->rhistory_len);
    });
  });
}

namespace {

// IC::solve / ILU::solve (ic.cpp:140-235, ilu.cpp:145-240): factor a copy of
// the values and apply the preconditioner once, x = (L U)^-1 b
int factor_solve(mgx_csr *A, mgx_vec *x, mgx_vec *b, const mgx_solve_params *p,
                 mgx_solve_result *res, int kind, const char *who) {
  return guard([&] {
    check_solve_args(A, x, b, p, res, who);
    activate(A->ctx);
    return run_solver(A, res, [&] {
      const PrecSpec ps{kind, p->omega};
      if (A->dtype == MGX_F64) {
        const Prec<double> P = make_precond<double>(A, A->val, ps);
        apply_precond(A->ctx, P, A->rows, (const double *)b->d, (double *)x->d);
      } else {
        const Prec<float> P = make_precond<float>(A, A->val, ps);
        apply_precond(A->ctx, P, A->rows, (const float *)b->d, (float *)x->d);
      }
      SolveOut o;
      o.status = MGX_OK;
      return o;
    });
  });
}

}  // namespace

int mgx_ic_solve(mgx_csr *A, mgx_vec *x, mgx_vec *b, const mgx_solve_params *p,
                 mgx_solve_result *res) {
  return factor_solve(A, x, b, p, res, MGX_PRECOND_IC, "mgx_ic_solve");
}

int mgx_ilu_solve(mgx_csr *A, mgx_vec *x, mgx_vec *b, const mgx_solve_params *p,
                  mgx_solve_result *res) {
  return factor_solve(A, x, b, p, res, MGX_PRECOND_ILU, "mgx_ilu_solve");
}

int mgx_lobpcg_solve(mgx_csr *A, mgx_vec *X, int64_t m, double *lambda,
                     const mgx_solve_params *p, mgx_solve_result *res) {
  return guard([&] {
    MGX_REQUIRE(A && X && lambda && p && res, "mgx_lobpcg_solve: NULL argument");
    MGX_REQUIRE(A->rows == A->cols, "mgx_lobpcg_solve: error, A.row != A.col");
    MGX_REQUIRE(m >= 1 && m <= 10, "mgx_lobpcg_solve: 1 <= m <= 10 eigenpairs");
    MGX_REQUIRE(3 * m <= A->rows, "mgx_lobpcg_solve: needs 3 m <= n");
    MGX_REQUIRE(X->ctx == A->ctx && X->dtype == A->dtype && X->n == m * A->rows,
                "mgx_lobpcg_solve: X must hold m * n entries of A's dtype");
    MGX_REQUIRE(p->precond >= MGX_PRECOND_DEFAULT && p->precond <= MGX_PRECOND_ILU,
                "mgx_lobpcg_solve: unknown preconditioner %d", p->precond);
    MGX_REQUIRE(p->precond != MGX_PRECOND_SOR || (p->omega > 0 && p->omega < 2),
                "mgx_lobpcg_solve: SOR omega must be in (0, 2)");
    activate(A->ctx);
    const int64_t mi = maxiter_of(p);
    const PrecSpec ps{p->precond, p->omega};
    return run_solver(A, res, [&] {
      if (A->dtype == MGX_F64)
        return lobpcg_run<double>(A, (double *)X->d, m, lambda, ps, p->tol, p->miniter, mi,
                                  p->rhistory, p->rhistory_len);
      return lobpcg_run<float>(A, (float *)X->d, m, lambda, ps, p->tol, p->miniter, mi,
                               p->rhistory, p->rhistory_len);
    });
  });
}

int mgx_sor_solve(mgx_csr *A, mgx_vec *x, mgx_vec *b, const mgx_solve_params *p,
                  mgx_solve_result *res) {
  return guard([&] {
    check_solve_args(A, x, b, p, res, "mgx_sor_solve");
    activate(A->ctx);
    // the SOR iteration's test is relative (sor.cpp:139,145)
    const double thr = p->tol;
    const int64_t mi = maxiter_of(p);
    const PrecSpec ps{p->precond, p->omega};
    return run_solver(A, res, [&] {
      if (A->dtype == MGX_F64)
        return sor_reference<double>(A, (double *)x->d, (const double *)b->d, ps, thr,
                                     p->miniter, mi, p->rhistory, p->rhistory_len);
      return sor_reference<float>(A, (float *)x->d, (const float *)b->d, ps, thr,
                                  p->miniter, mi, p->rhistory, p->rhistory_len);
    });
  });
}

int mgx_residual_l2(mgx_csr *A, mgx_vec *x, mgx_vec *b, double *out) {
  return guard([&] {
    MGX_REQUIRE(A && x && b && out, "mgx_residual_l2: NULL argument");
    MGX_REQUIRE(x->n == A->cols && b->n == A->rows && x->dtype == A->dtype &&
                    b->dtype == A->dtype,
                "mgx_residual_l2: size/dtype mismatch");
    Ctx *c = A->ctx;
    activate(c);
    void *tmp = ws_get(c, (size_t)std::max<int64_t>(A->rows, 1) * A->elem());
    launch_spmv(A, A->val, A->dtype, x->d, tmp, 1, 0, false);
    launch_vv(c, A->dtype, 1, A->rows, b->d, tmp, tmp);
    double d = dot_host(c, A->dtype, A->rows, tmp, tmp);
    ws_release_all(c);
    *out = A->dtype == MGX_F64 ? std::sqrt(d) : (double)std::sqrt((float)d);
    return MGX_OK;
  });
}

int mgx_ir_cg_solve(mgx_csr *A, mgx_vec *x, mgx_vec *b, const mgx_solve_params *p,
                    double inner_tol, int64_t max_outer, mgx_solve_result *res) {
  return guard([&] {
    check_solve_args(A, x, b, p, res, "mgx_ir_cg_solve");
    MGX_REQUIRE(p->precond <= MGX_PRECOND_JACOBI,
                "mgx_ir_cg_solve: Jacobi or none (the SOR / IC / ILU sweeps are not part of "
                "the fused engine)");
    MGX_REQUIRE(A->dtype == MGX_F64, "mgx_ir_cg_solve: needs an FP64 matrix");
    MGX_REQUIRE(inner_tol > 0 && inner_tol < 1, "mgx_ir_cg_solve: inner_tol in (0,1)");
    activate(A->ctx);
    const double thr = threshold(A->ctx, p, b);
    int64_t inner = 0, outer = 0;
    int st = run_solver(A, res, [&] {
      return ir_cg(A, (double *)x->d, (const double *)b->d, p->precond, thr, maxiter_of(p),
                   inner_tol, max_outer > 0 ? max_outer : 100, p->check_every, &inner,
                   &outer);
    });
    res->inner_iter = inner;
    res->outer_iter = outer;
    return st;
  });
}

int mgx_mixed_cg_solve(mgx_csr *A, mgx_vec *x, mgx_vec *b, const mgx_solve_params *p,
                       double delta, int64_t max_updates, mgx_solve_result *res) {
  return guard([&] {
    check_solve_args(A, x, b, p, res, "mgx_mixed_cg_solve");
    MGX_REQUIRE(p->precond <= MGX_PRECOND_JACOBI,
                "mgx_mixed_cg_solve: Jacobi or none (the SOR / IC / ILU sweeps are not part "
                "of the fused engine)");
    MGX_REQUIRE(A->dtype == MGX_F64, "mgx_mixed_cg_solve: needs an FP64 matrix");
    MGX_REQUIRE(delta > 0 && delta < 1, "mgx_mixed_cg_solve: delta in (0,1)");
    activate(A->ctx);
    const double thr = threshold(A->ctx, p, b);
    int64_t inner = 0, updates = 0;
    int st = run_solver(A, res, [&] {
      return mp_cg(A, (double *)x->d, (const double *)b->d, p->precond, thr, maxiter_of(p),
                   delta, max_updates > 0 ? max_updates : 1000, p->check_every, &inner,
                   &updates);
    });
    res->inner_iter = inner;
    res->outer_iter = updates;
    return st;
  });
}

}  // extern "C"

