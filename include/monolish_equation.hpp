// monolish_equation.hpp -- drop-in monolish::equation solvers for
// matrix::CRS<T> on the device: none, Jacobi and SOR (solvers and
// preconditioners), CG, BiCGSTAB (include/monolish_equation.hpp:40-257 of
// monolish 0.17.3).
//
// With none / Jacobi / no preconditioner bound, CG and BiCGSTAB run the fused
// device loop (mgx_cg_solve / mgx_bicgstab_solve); any other preconditioner
// object runs the reference loop through its own callbacks with one device
// BLAS call per step.  Both follow src/equation/{cg,bicgstab}.cpp step by
// step, with the same return codes, stopping rule (absolute nrm2(r) < tol
// after miniter) and final_iter / final_resid.
#pragma once

#include "monolish_blas.hpp"
#include "monolish_solver.hpp"
#include "monolish_vml.hpp"

namespace monolish {
namespace equation {

namespace detail_eq {
// scoped Logger::solver_in / solver_out (an "IN" line now, "OUT" with the time at scope exit)
class SolverLog {
 public:
  explicit SolverLog(const std::string &name) { Logger::get_instance().solver_in(name); }
  ~SolverLog() { Logger::get_instance().solver_out(); }
  SolverLog(const SolverLog &) = delete;
  SolverLog &operator=(const SolverLog &) = delete;
};
template <class MATRIX, class Float>
void check_system(const MATRIX &A, const vector<Float> &x, const vector<Float> &b,
                  const char *who) {
  if (A.get_row() != A.get_col()) throw std::runtime_error("error, A.row != A.col");
  if (A.get_device_mem_stat() != x.get_device_mem_stat() &&
      A.get_device_mem_stat() != b.get_device_mem_stat())
    throw std::runtime_error(
        "error, A.get_device_mem_stat != x.get_device_mem_stat != b.get_device_mem_stat");
  detail::need_device(A.get_device_mem_stat() && x.get_device_mem_stat() &&
                          b.get_device_mem_stat(),
                      who);
}
}  // namespace detail_eq

// ------------------------------------------------------------------- none
template <typename MATRIX, typename Float>
class none : public monolish::solver::solver<MATRIX, Float> {
 public:
  void create_precond(MATRIX &A) { (void)A; }
  // none.cpp:40-46: z = r
  void apply_precond(const vector<Float> &r, vector<Float> &z) { blas::copy(r, z); }
  [[nodiscard]] int solve(MATRIX &A, vector<Float> &x, vector<Float> &b) {
    (void)A, (void)x, (void)b;
    return MONOLISH_SOLVER_SUCCESS;
  }
  [[nodiscard]] std::string name() const { return "monolish::equation::none"; }
  [[nodiscard]] std::string solver_name() const { return "none"; }
};

// ----------------------------------------------------------------- Jacobi
template <typename MATRIX, typename Float>
class Jacobi : public monolish::solver::solver<MATRIX, Float> {
 public:
  // jacobi.cpp:12-29: M = 1 / diag(A), resident when A is
  void create_precond(MATRIX &A) {
    detail_eq::SolverLog log_(monolish_func);
    if (A.get_row() != A.get_col()) throw std::runtime_error("error A.row != A.col");
    auto &M = this->precond.M;
    M.device_free();
    M.resize(A.get_row());
    if (A.get_device_mem_stat()) M.send();
    A.diag(M);
    vml::reciprocal(M, M);
  }
  // jacobi.cpp:53-58: z = M o r
  void apply_precond(const vector<Float> &r, vector<Float> &z) {
    detail_eq::SolverLog log_(monolish_func);
    vml::mul(this->precond.M, r, z);
  }
  // jacobi.cpp:83-150 (relative stopping test)
  [[nodiscard]] int solve(MATRIX &A, vector<Float> &x, vector<Float> &b) {
    Logger::get_instance().solver_in(monolish_func);
    int ret = 0;
    if (this->lib == 0) {
      detail_eq::check_system(A, x, b, "equation::Jacobi");
      int kind = this->device_precond();
      if (kind >= 0) this->precond.create_precond(A);
      ret = this->run_device(mgx_jacobi_solve, A, x, b, kind >= 0 ? kind : MGX_PRECOND_DEFAULT,
                             MGX_MODE_REFERENCE);
    }
    Logger::get_instance().solver_out();
    return ret;
  }
  [[nodiscard]] std::string name() const { return "monolish::equation::Jacobi"; }
  [[nodiscard]] std::string solver_name() const { return "Jacobi"; }
};

// -------------------------------------------------------------------- SOR
// As a preconditioner (sor.cpp:12-82) it is bound to CG / BiCGSTAB /
// Jacobi / SOR solvers, which then run the device sweeps (the forward and
// backward loops of sor_kernel_precond with M = 1/(omega diag(A))); as a
// solver (sor.cpp:91-164) it iterates x += (D/omega + L)^-1 (b - A M x) on
// the device with its own omega (a bound SOR preconditioner uses it too).
template <typename MATRIX, typename Float>
class SOR : public monolish::solver::solver<MATRIX, Float> {
 public:
  void create_precond(MATRIX &A) {
    if (A.get_row() != A.get_col()) throw std::runtime_error("error A.row != A.col");
    this->precond.A = &A;
  }
  void apply_precond(const vector<Float> &r, vector<Float> &z) {
    (void)r, (void)z;
    throw std::runtime_error(
        "equation::SOR::apply_precond: bind the SOR object to a solver (the sweeps run there)");
  }
  [[nodiscard]] int solve(MATRIX &A, vector<Float> &x, vector<Float> &b) {
    Logger::get_instance().solver_in(monolish_func);
    int ret = 0;
    if (this->lib == 0) {
      detail_eq::check_system(A, x, b, "equation::SOR");
      int kind = this->device_precond();
      if (kind >= 0) this->precond.create_precond(A);
      ret = this->run_device(mgx_sor_solve, A, x, b, kind >= 0 ? kind : MGX_PRECOND_DEFAULT,
                             MGX_MODE_REFERENCE);
    }
    Logger::get_instance().solver_out();
    return ret;
  }
  [[nodiscard]] std::string name() const { return "monolish::equation::SOR"; }
  [[nodiscard]] std::string solver_name() const { return "SOR"; }
};

// ------------------------------------------------------------- IC / ILU
// As preconditioners (ic.cpp:12-140, ilu.cpp:12-145) they are bound to CG /
// BiCGSTAB / Jacobi / SOR solvers, which factor a copy of A's values on the
// device (IC(0) of the lower pattern / ILU(0), no fill) and apply the two
// level-scheduled triangular solves; as solvers (ic.cpp:140-235,
// ilu.cpp:145-240) they apply the factor once, x = M^-1 b.  The reference
// requires device-resident data (it throws on the CPU); so does this one.
template <typename MATRIX, typename Float>
class IC : public monolish::solver::solver<MATRIX, Float> {
 public:
  void create_precond(MATRIX &A) {
    if (A.get_row() != A.get_col()) throw std::runtime_error("error A.row != A.col");
    this->precond.A = &A;
  }
  void apply_precond(const vector<Float> &r, vector<Float> &z) {
    (void)r, (void)z;
    throw std::runtime_error(
        "equation::IC::apply_precond: bind the IC object to a solver (the solves run there)");
  }
  [[nodiscard]] int solve(MATRIX &A, vector<Float> &x, vector<Float> &b) {
    Logger::get_instance().solver_in(monolish_func);
    int ret = 0;
    if (this->lib == 0) {
      detail_eq::check_system(A, x, b, "equation::IC");
      ret = this->run_device(mgx_ic_solve, A, x, b, MGX_PRECOND_DEFAULT, MGX_MODE_REFERENCE);
    }
    Logger::get_instance().solver_out();
    return ret;
  }
  [[nodiscard]] std::string name() const { return "monolish::equation::IC"; }
  [[nodiscard]] std::string solver_name() const { return "IC"; }
};

template <typename MATRIX, typename Float>
class ILU : public monolish::solver::solver<MATRIX, Float> {
 public:
  void create_precond(MATRIX &A) {
    if (A.get_row() != A.get_col()) throw std::runtime_error("error A.row != A.col");
    this->precond.A = &A;
  }
  void apply_precond(const vector<Float> &r, vector<Float> &z) {
    (void)r, (void)z;
    throw std::runtime_error(
        "equation::ILU::apply_precond: bind the ILU object to a solver (the solves run there)");
  }
  [[nodiscard]] int solve(MATRIX &A, vector<Float> &x, vector<Float> &b) {
    Logger::get_instance().solver_in(monolish_func);
    int ret = 0;
    if (this->lib == 0) {
      detail_eq::check_system(A, x, b, "equation::ILU");
      ret = this->run_device(mgx_ilu_solve, A, x, b, MGX_PRECOND_DEFAULT, MGX_MODE_REFERENCE);
    }
    Logger::get_instance().solver_out();
    return ret;
  }
  [[nodiscard]] std::string name() const { return "monolish::equation::ILU"; }
  [[nodiscard]] std::string solver_name() const { return "ILU"; }
};

// --------------------------------------------------------------------- CG
template <typename MATRIX, typename Float>
class CG : public monolish::solver::solver<MATRIX, Float> {
 public:
  [[nodiscard]] int solve(MATRIX &A, vector<Float> &x, vector<Float> &b) {
    Logger::get_instance().solver_in(monolish_func);
    int ret = 0;
    if (this->lib == 0) ret = monolish_CG(A, x, b);
    Logger::get_instance().solver_out();
    return ret;
  }
  void create_precond(MATRIX &A) {
    (void)A;
    throw std::runtime_error("this precond. is not impl.");
  }
  void apply_precond(const vector<Float> &r, vector<Float> &z) {
    (void)r, (void)z;
    throw std::runtime_error("this precond. is not impl.");
  }
  [[nodiscard]] std::string name() const { return "monolish::equation::CG"; }
  [[nodiscard]] std::string solver_name() const { return "CG"; }

 private:
  int monolish_CG(MATRIX &A, vector<Float> &x, vector<Float> &b) {
    detail_eq::SolverLog log_(monolish_func);
    detail_eq::check_system(A, x, b, "equation::CG");
    if (this->resid_method != 0) throw std::runtime_error("error vector size is not same");
    const int kind = this->device_precond();
    if (kind >= 0) {
      this->precond.create_precond(A);
      return this->run_device(mgx_cg_solve, A, x, b, kind, MGX_MODE_FUSED);
    }
    // cg.cpp:27-85 with the bound callbacks
    const size_t n = A.get_row();
    vector<Float> r(n, 0.0), p(n, 0.0), q(n, 0.0), z(n, 0.0);
    util::send(r, p, q, z);
    this->precond.create_precond(A);
    blas::matvec(A, x, q);
    vml::sub(b, q, r);
    this->precond.apply_precond(r, z);
    blas::copy(z, p);
    for (size_t iter = 0; iter < this->get_maxiter(); iter++) {
      blas::matvec(A, p, q);
      auto tmp = blas::dot(z, r);
      auto alpha = tmp / blas::dot(p, q);
      blas::axpy(alpha, p, x);
      blas::axpy(-alpha, q, r);
      this->precond.apply_precond(r, z);
      auto beta = blas::dot(z, r) / tmp;
      blas::xpay(beta, z, p);
      Float resid = blas::nrm2(r);
      if (this->get_print_rhistory())
        *this->rhistory_stream << iter + 1 << "\t" << std::scientific << resid << std::endl;
      if (resid < this->get_tol() && this->get_miniter() <= iter + 1) {
        this->final_iter = iter + 1;
        this->final_resid = blas::nrm2(r);
        return MONOLISH_SOLVER_SUCCESS;
      }
      if (std::isnan(resid)) {
        this->final_iter = iter + 1;
        this->final_resid = blas::nrm2(r);
        return MONOLISH_SOLVER_RESIDUAL_NAN;
      }
    }
    this->final_iter = this->maxiter;
    this->final_resid = blas::nrm2(r);
    return MONOLISH_SOLVER_MAXITER;
  }
};

// --------------------------------------------------------------- BiCGSTAB
template <typename MATRIX, typename Float>
class BiCGSTAB : public monolish::solver::solver<MATRIX, Float> {
 public:
  [[nodiscard]] int solve(MATRIX &A, vector<Float> &x, vector<Float> &b) {
    Logger::get_instance().solver_in(monolish_func);
    int ret = 0;
    if (this->lib == 0) ret = monolish_BiCGSTAB(A, x, b);
    Logger::get_instance().solver_out();
    return ret;
  }
  void create_precond(MATRIX &A) {
    (void)A;
    throw std::runtime_error("this precond. is not impl.");
  }
  void apply_precond(const vector<Float> &r, vector<Float> &z) {
    (void)r, (void)z;
    throw std::runtime_error("this precond. is not impl.");
  }
  [[nodiscard]] std::string name() const { return "monolish::equation::BiCGSTAB"; }
  [[nodiscard]] std::string solver_name() const { return "BiCGSTAB"; }

 private:
  int monolish_BiCGSTAB(MATRIX &A, vector<Float> &x, vector<Float> &b) {
    detail_eq::SolverLog log_(monolish_func);
    detail_eq::check_system(A, x, b, "equation::BiCGSTAB");
    if (this->resid_method != 0) throw std::runtime_error("error vector size is not same");
    const int kind = this->device_precond();
    if (kind >= 0) {
      this->precond.create_precond(A);
      return this->run_device(mgx_bicgstab_solve, A, x, b, kind, MGX_MODE_FUSED);
    }
    // bicgstab.cpp:27-134 with the bound callbacks
    const size_t n = A.get_row();
    vector<Float> r(n, 0.0), r0(n, 0.0), p(n, 0.0), phat(n, 0.0), s(n, 0.0), shat(n, 0.0),
        v(n, 0.0), t(n, 0.0);
    util::send(r, r0, p, phat, s, shat, v, t);
    Float rho_old = 1, rho = 1, alpha = 1, beta, omega = 1;
    this->precond.create_precond(A);
    blas::matvec(A, x, r);
    vml::sub(b, r, r);
    blas::copy(r, r0);
    for (size_t iter = 0; iter < this->get_maxiter(); iter++) {
      rho = blas::dot(r, r0);
      if (rho == 0.0) {
        this->final_iter = iter + 1;
        this->final_resid = blas::nrm2(r);
        return MONOLISH_SOLVER_BREAKDOWN;
      }
      if (iter == 0) {
        blas::copy(r, p);
      } else {
        beta = (rho / rho_old) * (alpha / omega);
        blas::axpy(-omega, v, p);
        blas::xpay(beta, r, p);
      }
      this->precond.apply_precond(p, phat);
      blas::matvec(A, phat, v);
      alpha = rho / blas::dot(v, r0);
      blas::axpyz(-alpha, v, r, s);
      this->precond.apply_precond(s, shat);
      blas::matvec(A, shat, t);
      omega = blas::dot(t, s) / blas::dot(t, t);
      if (omega == 0.0) {
        this->final_iter = iter + 1;
        this->final_resid = blas::nrm2(r);
        return MONOLISH_SOLVER_BREAKDOWN;
      }
      blas::axpy(alpha, phat, x);
      blas::axpy(omega, shat, x);
      blas::axpyz(-omega, t, s, r);
      auto resid = blas::nrm2(r);
      if (this->get_print_rhistory())
        *this->rhistory_stream << iter + 1 << "\t" << std::scientific << resid << std::endl;
      if (resid < this->get_tol() && this->get_miniter() <= iter + 1) {
        this->final_iter = iter + 1;
        this->final_resid = blas::nrm2(r);
        return MONOLISH_SOLVER_SUCCESS;
      }
      if (std::isnan(resid)) {
        this->final_iter = iter + 1;
        this->final_resid = blas::nrm2(r);
        return MONOLISH_SOLVER_RESIDUAL_NAN;
      }
      rho_old = rho;
    }
    this->final_iter = this->maxiter;
    this->final_resid = blas::nrm2(r);
    return MONOLISH_SOLVER_MAXITER;
  }
};

}  // namespace equation
}  // namespace monolish
