// monolish_solver.hpp -- drop-in solver base class and preconditioner
// plumbing (include/monolish_solver.hpp:32-252 of monolish 0.17.3).
#pragma once

#include <vector>

#include "monolish/common/monolish_common.hpp"

namespace monolish {
namespace equation {
template <typename MATRIX, typename Float> class none;
template <typename MATRIX, typename Float> class Jacobi;
template <typename MATRIX, typename Float> class SOR;
template <typename MATRIX, typename Float> class IC;
template <typename MATRIX, typename Float> class ILU;
}  // namespace equation

namespace solver {

enum class initvec_scheme { RANDOM, USER };

namespace detail_solver {
// which device preconditioner a bound object corresponds to; anything else
// runs through its own std::function callbacks
template <class P> struct precond_kind { static constexpr int value = -1; };
template <class M, class F> struct precond_kind<equation::none<M, F>> {
  static constexpr int value = MGX_PRECOND_NONE;
};
template <class M, class F> struct precond_kind<equation::Jacobi<M, F>> {
  static constexpr int value = MGX_PRECOND_JACOBI;
};
template <class M, class F> struct precond_kind<equation::SOR<M, F>> {
  static constexpr int value = MGX_PRECOND_SOR;
};
template <class M, class F> struct precond_kind<equation::IC<M, F>> {
  static constexpr int value = MGX_PRECOND_IC;
};
template <class M, class F> struct precond_kind<equation::ILU<M, F>> {
  static constexpr int value = MGX_PRECOND_ILU;
};
// the relaxation coefficient of a bound SOR object (read at solve time)
template <class P> std::function<double()> omega_of(P &) { return nullptr; }
template <class M, class F> std::function<double()> omega_of(equation::SOR<M, F> &p) {
  return [&p] { return (double)p.get_omega(); };
}
}  // namespace detail_solver

template <typename MATRIX, typename Float> class precondition {
 public:
  vector<Float> M;
  MATRIX *A = nullptr;
  std::function<void(MATRIX &)> create_precond;
  std::function<void(const vector<Float> &r, vector<Float> &z)> apply_precond;
  // MGX_PRECOND_* of the bound create/apply objects (-1: custom callbacks)
  int create_kind = MGX_PRECOND_DEFAULT, apply_kind = MGX_PRECOND_DEFAULT;
  std::function<double()> omega;  // a bound SOR object's omega

  std::function<void(void)> get_precond();
  void set_precond_data(vector<Float> &m) { M = m; }
  vector<Float> get_precond_data() { return M; }

  precondition() {
    create_precond = [](MATRIX &) {};
    apply_precond = [](const vector<Float> &r, vector<Float> &z) { z = r; };
  }
};

template <typename MATRIX, typename Float> class solver {
 protected:
  int lib = 0;
  double tol = 1.0e-8;
  size_t miniter = 0;
  size_t maxiter = SIZE_MAX;
  size_t resid_method = 0;
  bool print_rhistory = false;
  std::string rhistory_file;
  std::ostream *rhistory_stream = &std::cout;
  initvec_scheme initvecscheme = initvec_scheme::RANDOM;
  double final_resid = 0;
  size_t final_iter = 0;
  Float omega = 1.9;
  int singularity = 0;
  int reorder = 3;
  precondition<MATRIX, Float> precond;

  // the precondition kind both callbacks agree on, or -1
  int device_precond() const {
    if (precond.create_kind == precond.apply_kind && precond.create_kind >= 0)
      return precond.create_kind;
    return -1;
  }

  // one call of a C-ABI solver with this object's settings; prints the
  // residual history when requested (cg.cpp:62-65)
  template <class MAT>
  int run_device(int (*fn)(mgx_csr *, mgx_vec *, mgx_vec *, const mgx_solve_params *,
                           mgx_solve_result *),
                 MAT &A, vector<Float> &x, vector<Float> &b, int kind, int mode) {
    if (maxiter == 0) {  // the loop never runs: MAXITER with the initial residual
      final_iter = 0;
      double rr = 0;
      detail::check(mgx_residual_l2(A.dev(), x.dev(), b.dev(), &rr));
      final_resid = rr;
      return MONOLISH_SOLVER_MAXITER;
    }
    mgx_solve_params p;
    mgx_solve_params_default(&p);
    p.tol = tol;
    p.maxiter = maxiter >= (size_t)INT64_MAX ? 0 : (int64_t)maxiter;
    p.miniter = (int64_t)miniter;
    p.precond = kind;
    p.mode = mode;
    // the SOR solver uses its own omega (also for a bound SOR preconditioner)
    p.omega = fn != mgx_sor_solve && kind == MGX_PRECOND_SOR && precond.omega ? precond.omega()
                                                                             : (double)omega;
    std::vector<double> hist;
    if (print_rhistory) {
      hist.assign(std::min<size_t>(maxiter, (size_t)1 << 22), 0.0);
      p.rhistory = hist.data();
      p.rhistory_len = (int64_t)hist.size();
    }
    mgx_solve_result r;
    const int rc = fn(A.dev(), x.dev(), b.dev(), &p, &r);
    detail::check(rc);
    final_iter = (size_t)r.final_iter;
    final_resid = r.final_resid;
    if (print_rhistory) {
      size_t lines = std::min<size_t>(final_iter, hist.size());
      if (rc == MONOLISH_SOLVER_BREAKDOWN && lines > 0) lines = final_iter - 1;
      for (size_t i = 0; i < lines; i++)
        *rhistory_stream << i + 1 << "\t" << std::scientific << (Float)hist[i] << std::endl;
    }
    return rc;
  }

 public:
  solver() = default;
  ~solver() {
    if (rhistory_stream != &std::cout && !rhistory_file.empty()) delete rhistory_stream;
  }

  template <class PRECOND> void set_create_precond(PRECOND &p) {
    precond.create_precond = std::bind(&PRECOND::create_precond, &p, std::placeholders::_1);
    precond.create_kind = detail_solver::precond_kind<PRECOND>::value;
    precond.omega = detail_solver::omega_of(p);
  }
  template <class PRECOND> void set_apply_precond(PRECOND &p) {
    precond.apply_precond = std::bind(&PRECOND::apply_precond, &p, std::placeholders::_1,
                                      std::placeholders::_2);
    precond.apply_kind = detail_solver::precond_kind<PRECOND>::value;
  }

  void set_lib(int l) { lib = l; }
  void set_tol(double t) { tol = t; }
  void set_maxiter(size_t max) { maxiter = max; }
  void set_miniter(size_t min) { miniter = min; }
  void set_residual_method(size_t r) { resid_method = r; }
  void set_print_rhistory(bool flag) {
    print_rhistory = flag;
    rhistory_stream = &std::cout;
  }
  void set_rhistory_filename(std::string file) {
    rhistory_file = file;
    rhistory_stream = new std::ofstream(rhistory_file);
    if (rhistory_stream->fail()) throw std::runtime_error("error bad filename");
  }
  void set_initvec_scheme(initvec_scheme scheme) { initvecscheme = scheme; }
  void set_omega(Float w) { omega = w; }
  void set_reorder(int r) { reorder = r; }

  [[nodiscard]] int get_lib() const { return lib; }
  [[nodiscard]] double get_tol() const { return tol; }
  [[nodiscard]] size_t get_maxiter() const { return maxiter; }
  [[nodiscard]] size_t get_miniter() const { return miniter; }
  [[nodiscard]] size_t get_residual_method() const { return resid_method; }
  bool get_print_rhistory() const { return print_rhistory; }
  initvec_scheme get_initvec_scheme() const { return initvecscheme; }
  Float get_omega() { return omega; }
  int get_reorder() { return reorder; }
  int get_singularity() { return singularity; }
  double get_final_residual() { return final_resid; }
  size_t get_final_iter() { return final_iter; }
};

}  // namespace solver
}  // namespace monolish
