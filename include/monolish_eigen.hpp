// monolish_eigen.hpp -- drop-in standard eigensolver (include/monolish_eigen.hpp
// and include/monolish/eigen/monolish_standard_eigen.hpp of monolish 0.17.3).
//
// standard_eigen::LOBPCG<CRS<T>, T>::solve(A, l, x) computes the l.size()
// smallest eigenpairs of symmetric A on the device (mgx_lobpcg_solve): x is
// an l.size() x n Dense matrix whose rows are the eigenvectors.  The
// preconditioner bound with set_create_precond / set_apply_precond must be
// one of the library's (none, Jacobi, SOR, IC, ILU).  initvec_scheme::USER
// starts from x's rows; RANDOM (the default) from unit vectors drawn with
// std::random_device, as the reference's unseeded util::random_vector.
#pragma once

#include "monolish_equation.hpp"

namespace monolish {
namespace standard_eigen {

template <typename MATRIX, typename Float>
class LOBPCG : public monolish::solver::solver<MATRIX, Float> {
 public:
  // src/eigen/lobpcg_standard.cpp:10-300
  [[nodiscard]] int solve(MATRIX &A, vector<Float> &l, matrix::Dense<Float> &x) {
    Logger::get_instance().solver_in(monolish_func);
    if (A.get_col() != x.get_col()) throw std::runtime_error("error, A.col != x.col");
    if (l.size() != x.get_row()) throw std::runtime_error("error, l.size != x.row");
    detail::need_device(A.get_device_mem_stat(), "matrix");
    const int kind = this->device_precond();
    if (kind < 0)
      throw std::runtime_error(
          "standard_eigen::LOBPCG: bind none, Jacobi, SOR, IC or ILU (the iteration runs on the "
          "device)");
    this->precond.create_precond(A);
    const size_t m = l.size(), n = A.get_col();
    if (this->get_initvec_scheme() == monolish::solver::initvec_scheme::RANDOM) {
      // lobpcg_standard.cpp:82-91: random (0, 1) vectors, normalised
      if (x.get_device_mem_stat()) x.device_free();
      std::random_device rd;
      std::mt19937 mt(rd());
      std::uniform_real_distribution<> u(0.0, 1.0);
      for (size_t i = 0; i < m; i++) {
        double s = 0;
        std::vector<double> r(n);
        for (auto &v : r) {
          v = u(mt);
          s += v * v;
        }
        s = std::sqrt(s);
        for (size_t j = 0; j < n; j++) x.insert(i, j, (Float)(r[j] / s));
      }
    }
    if (!x.get_device_mem_stat()) x.send();
    mgx_solve_params p;
    mgx_solve_params_default(&p);
    p.tol = this->tol;
    p.maxiter = this->maxiter >= (size_t)INT64_MAX ? 0 : (int64_t)this->maxiter;
    p.miniter = (int64_t)this->miniter;
    p.precond = kind;
    p.omega = kind == MGX_PRECOND_SOR && this->precond.omega ? this->precond.omega()
                                                             : (double)this->omega;
    std::vector<double> hist;
    if (this->print_rhistory) {
      hist.assign(std::min<size_t>(this->maxiter, (size_t)1 << 22), 0.0);
      p.rhistory = hist.data();
      p.rhistory_len = (int64_t)hist.size();
    }
    std::vector<double> lam(m);
    mgx_solve_result r;
    const int rc = mgx_lobpcg_solve(A.dev(), x.dev(), (int64_t)m, lam.data(), &p, &r);
    detail::check(rc);
    this->final_iter = (size_t)r.final_iter;
    this->final_resid = r.final_resid;
    if (l.get_device_mem_stat()) {
      std::vector<Float> lv(lam.begin(), lam.end());
      detail::check(mgx_vec_upload(l.dev(), lv.data()));
    } else {
      for (size_t i = 0; i < m; i++) l.data()[i] = (Float)lam[i];
    }
    if (this->print_rhistory)
      for (size_t i = 0; i < std::min<size_t>(this->final_iter, hist.size()); i++)
        *this->rhistory_stream << i + 1 << "\t" << std::scientific << (Float)hist[i] << std::endl;
    Logger::get_instance().solver_out();
    return rc;
  }
  void create_precond(MATRIX &A) {
    (void)A;
    throw std::runtime_error("this precond. is not impl.");
  }
  void apply_precond(const vector<Float> &r, vector<Float> &z) {
    (void)r, (void)z;
    throw std::runtime_error("this precond. is not impl.");
  }
  [[nodiscard]] std::string name() const { return "monolish::standard_eigen::LOBPCG"; }
  [[nodiscard]] std::string solver_name() const { return "LOBPCG"; }
};

}  // namespace standard_eigen

namespace eigen = standard_eigen;

}  // namespace monolish
