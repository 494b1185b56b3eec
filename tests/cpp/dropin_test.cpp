// Drop-in check of the monolish C++ API on the device path, written the way
// monolish's own tests are (test/matrix/matrix_common.cpp,
// test/vector/blas/*.hpp, test/equation/iterative/equation_kernel.hpp):
// known answers, host loops as the answer, the solver harness on test.mtx.
// Prints "PASS" and returns 0 when every check holds.
#include <cmath>
#include <cstring>
#include <iostream>
#include <random>

#include "monolish_blas.hpp"
#include "monolish_eigen.hpp"
#include "monolish_equation.hpp"
#include "monolish_vml.hpp"

static std::mt19937 test_random_engine(1234);
static int failures = 0;

#define CHECK(cond)                                                          \
  do {                                                                       \
    if (!(cond)) {                                                           \
      std::cout << "FAIL " << __LINE__ << ": " #cond << std::endl;           \
      failures++;                                                            \
    }                                                                        \
  } while (0)

template <typename T> bool ans_check(const T *result, const T *ans, size_t n, double tol) {
  for (size_t i = 0; i < n; i++) {
    double err = ans[i] <= tol ? std::abs(result[i] - ans[i])
                               : std::abs((result[i] - ans[i]) / ans[i]);
    if (!(err < tol)) return false;
  }
  return true;
}

template <typename T> void matvec_known_answers(const std::string &dir) {
  const int row[] = {0, 0, 0, 1, 1, 2, 2, 2}, col[] = {0, 1, 2, 0, 2, 0, 1, 2};
  const T val[] = {1, 2, 3, 4, 5, 6, 7, 8};
  monolish::matrix::COO<T> addr_COO(3, 3, 8, row, col, val);
  monolish::matrix::CRS<T> addr_CRS(addr_COO);
  monolish::matrix::COO<T> file_COO(dir + "/test.mtx");
  monolish::matrix::CRS<T> file_CRS(file_COO);
  monolish::vector<T> x(3, 10);
  monolish::vector<T> filey(3), addry(3);
  monolish::util::send(x, filey, addry, file_CRS, addr_CRS);
  monolish::blas::matvec(file_CRS, x, filey);
  monolish::blas::matvec(addr_CRS, x, addry);
  monolish::util::recv(addry, filey);
  CHECK(addry[0] == 60 && addry[1] == 90 && addry[2] == 210);
  CHECK(filey[0] == 10 && filey[1] == 0 && filey[2] == 10);

  auto expanded = addr_COO;
  expanded.set_row(4);
  expanded.set_col(4);
  expanded.insert(3, 3, 1.0);
  expanded.insert(0, 0, 3.0);
  CHECK(expanded.get_nnz() == 10);
  expanded.sort(true);
  CHECK(expanded.at(0, 0) == 3.0 && expanded.get_nnz() == 9);
}

template <typename T> void vector_blas(size_t n) {
  monolish::vector<T> x(n, 0.0, 1.0, test_random_engine());
  monolish::vector<T> y(n, 0.0, 1.0, test_random_engine());
  std::vector<T> hx(x.begin(), x.end()), hy(y.begin(), y.end());
  monolish::util::send(x, y);
  const T a = 123.0;
  monolish::view1D<monolish::vector<T>, T> xv(x, 5, n / 2), yv(y, 5, n / 2);
  monolish::blas::axpy(a, xv, yv);  // y[5:5+n/2] += a x[5:5+n/2]
  monolish::blas::xpay(a, x, y);    // y = x + a y
  monolish::blas::scal(a, x);       // x = a x
  y.recv();
  x.recv();
  for (size_t i = 5; i < 5 + n / 2; i++) {
    T t = a * hx[i];
    hy[i] = t + hy[i];
  }
  bool ok = true;
  for (size_t i = 0; i < n; i++) {
    T t = a * hy[i];
    T yy = hx[i] + t;
    ok = ok && y[i] == yy && x[i] == (T)(a * hx[i]);
  }
  CHECK(ok);
  // dot against a tolerance as in test/vector/vector_blas.cpp
  monolish::vector<T> u(n, 1.0), w(n, 2.0);
  monolish::util::send(u, w);
  T d = monolish::blas::dot(u, w);
  CHECK(d == (T)(2.0 * n));
  CHECK(std::abs(monolish::blas::nrm2(w) - std::sqrt((T)(4.0 * n))) <= 1e-5 * std::sqrt(4.0 * n));
  monolish::vector<T> z(n);
  z.send();
  monolish::vml::sub(w, u, z);  // 1
  monolish::vml::mul(z, w, z);  // 2
  monolish::vml::reciprocal(z, z);  // 0.5
  monolish::blas::axpyz(a, u, z, w);  // 123.5
  w.recv();
  bool ok2 = true;
  for (size_t i = 0; i < n; i++) ok2 = ok2 && w[i] == (T)123.5;
  CHECK(ok2);
}

template <class MATRIX, class T, class SOLVER, class PRECOND>
bool solver_harness(const std::string &file, T tol) {
  monolish::matrix::COO<T> COO(file);
  MATRIX A(COO);
  monolish::vector<T> ans(A.get_row(), 1.0);
  monolish::vector<T> b(A.get_row(), 0.0);
  monolish::vector<T> x(A.get_row(), 0.0, 1.0, test_random_engine());
  monolish::util::send(A, x, b, ans);
  monolish::blas::matvec(A, ans, b);
  SOLVER solver;
  solver.set_tol(tol);
  solver.set_lib(0);
  solver.set_miniter(0);
  solver.set_maxiter(10000);
  PRECOND precond;
  precond.set_omega(1.9);  // test/equation/iterative/equation_kernel.hpp:29
  solver.set_create_precond(precond);
  solver.set_apply_precond(precond);
  int rc = solver.solve(A, x, b);
  ans.recv();
  x.recv();
  return rc == MONOLISH_SOLVER_SUCCESS && ans_check<T>(x.data(), ans.data(), x.size(), tol * 10);
}

// a user preconditioner unknown to the library: Jacobi written by hand, so
// the solver runs the reference loop through the callbacks
template <class MATRIX, class T> struct MyJacobi {
  monolish::vector<T> M;
  void create_precond(MATRIX &A) {
    M.device_free();
    M.resize(A.get_row());
    M.send();
    A.diag(M);
    monolish::vml::reciprocal(M, M);
  }
  void apply_precond(const monolish::vector<T> &r, monolish::vector<T> &z) {
    monolish::vml::mul(M, r, z);
  }
};

template <typename T> void fused_and_callback_paths_agree() {
  using MAT = monolish::matrix::CRS<T>;
  auto coo = monolish::util::laplacian_matrix_2D_5p<T>(32, 32);
  MAT A(coo);
  monolish::vector<T> b(A.get_row(), -1.0, 1.0, 1001);
  monolish::vector<T> x1(A.get_row(), 0.0), x2(A.get_row(), 0.0);
  monolish::util::send(A, b, x1, x2);
  monolish::equation::CG<MAT, T> s1, s2;
  monolish::equation::Jacobi<MAT, T> j;
  MyJacobi<MAT, T> mine;
  s1.set_create_precond(j);
  s1.set_apply_precond(j);
  s2.set_create_precond(mine);
  s2.set_apply_precond(mine);
  const T tol = sizeof(T) == 8 ? 1e-10 * monolish::blas::nrm2(b) : 1e-5 * monolish::blas::nrm2(b);
  s1.set_tol(tol);
  s2.set_tol(tol);
  CHECK(s1.solve(A, x1, b) == 0);
  CHECK(s2.solve(A, x2, b) == 0);
  CHECK(s1.get_final_iter() == s2.get_final_iter());
  CHECK(s1.get_final_residual() == s2.get_final_residual());
  x1.recv();
  x2.recv();
  CHECK(std::memcmp(x1.data(), x2.data(), sizeof(T) * x1.size()) == 0);
  CHECK(monolish::util::solver_check(MONOLISH_SOLVER_MAXITER) == false);
}

// test/eigen/LOBPCG/lobpcg.cpp:91-130, 268-282: tridiagonal Toeplitz (100,
// 10, -1), 2 pairs, x(i, 7i mod n) = 1 (USER) or random (RANDOM), residual
// tol 1e-5 * 100, maxiter 100, eigenvalues within 1e-3 of the closed form
template <typename T, typename PRECOND> bool lobpcg_harness(bool user) {
  using MAT = monolish::matrix::CRS<T>;
  const int dim = 100;
  auto coo = monolish::util::tridiagonal_toeplitz_matrix<T>(dim, 10.0, -1.0);
  MAT A(coo);
  monolish::vector<T> lambda(2);
  monolish::matrix::Dense<T> x(2, A.get_row());
  for (size_t i = 0; i < 2; i++) x.insert(i, (i * 7) % A.get_row(), 1.0);
  monolish::util::send(A);
  monolish::eigen::LOBPCG<MAT, T> solver;
  solver.set_tol(1.0e-5 * dim);
  solver.set_lib(0);
  solver.set_miniter(0);
  solver.set_maxiter(dim);
  if (user) solver.set_initvec_scheme(monolish::solver::initvec_scheme::USER);
  PRECOND precond;
  solver.set_create_precond(precond);
  solver.set_apply_precond(precond);
  if (monolish::util::solver_check(solver.solve(A, lambda, x))) return false;
  for (int i = 0; i < 2; i++) {
    const double ex = monolish::util::tridiagonal_toeplitz_matrix_eigenvalue<T>(dim, i, 10.0, -1.0);
    if (std::abs((lambda.data()[i] - ex) / ex) >= 1e-3) return false;
  }
  monolish::util::recv(x);
  return x.get_row() == 2 && !x.get_device_mem_stat();
}

// The Logger hooks alone, with no device work: the nesting, the level gate
// and the line format of logger_utils.cpp:7-178.  Also util::is_same_structure
// and is_same_size on host objects.
int log_hooks(size_t level, const std::string &file) {
  monolish::util::set_log_level(level);
  monolish::util::set_log_filename(file);
  monolish::Logger &l = monolish::Logger::get_instance();
  l.solver_in("solve");
  l.solver_in("monolish_CG");
  l.func_in("Ddot_core");
  l.util_in("send");
  l.util_out();
  l.func_out();
  l.solver_out();
  l.solver_out();
  monolish::util::set_log_level(0);
  using MAT = monolish::matrix::CRS<double>;
  auto ca = monolish::util::laplacian_matrix_2D_5p<double>(4, 4);
  auto cc = monolish::util::laplacian_matrix_2D_5p<double>(2, 8);  // 16 rows, fewer entries
  MAT A(ca), B(ca), C(cc);
  monolish::vector<double> u(16), v(16), w(15);
  CHECK(monolish::util::is_same_structure(A, B));
  CHECK(!monolish::util::is_same_structure(A, C));
  CHECK(monolish::util::is_same_size(A, C));
  CHECK(monolish::util::is_same_structure(u, v));
  CHECK(!monolish::util::is_same_structure(u, v, w));
  CHECK(!monolish::util::is_same_structure(A, u));
  return failures == 0 ? 0 : 1;
}

// The Logger's output for one drop-in CG solve with a Jacobi preconditioner
// on test.mtx (device resident, as the reference's GPU runs).
int log_cg_solve(const std::string &dir, size_t level, const std::string &file) {
  monolish::util::set_log_level(level);
  monolish::util::set_log_filename(file);
  using MAT = monolish::matrix::CRS<double>;
  monolish::matrix::COO<double> coo(dir + "/test.mtx");
  MAT A(coo);
  monolish::vector<double> x(A.get_row(), 0.0), b(A.get_row(), 1.0);
  monolish::util::send(A, x, b);
  monolish::equation::CG<MAT, double> solver;
  monolish::equation::Jacobi<MAT, double> precond;
  solver.set_create_precond(precond);
  solver.set_apply_precond(precond);
  solver.set_tol(1e-12);
  const int rc = solver.solve(A, x, b);
  monolish::util::recv(x);
  monolish::util::set_log_level(0);
  return rc;
}

int main(int argc, char **argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  using namespace monolish;
  if (argc > 4 && std::string(argv[2]) == "loghooks")
    return log_hooks(std::stoul(argv[3]), argv[4]);
  if (argc > 4 && std::string(argv[2]) == "logcg")
    return log_cg_solve(dir, std::stoul(argv[3]), argv[4]);
  matvec_known_answers<double>(dir);
  matvec_known_answers<float>(dir);
  vector_blas<double>(4321);
  vector_blas<float>(1234);
  const std::string mtx = dir + "/test.mtx", mtx2 = dir + "/test2.mtx";
#define HARNESS(S, P, T, TOL, F)                                                              \
  CHECK((solver_harness<matrix::CRS<T>, T, equation::S<matrix::CRS<T>, T>,                    \
                        equation::P<matrix::CRS<T>, T>>(F, TOL)))
  HARNESS(CG, none, double, 1e-8, mtx);
  HARNESS(CG, Jacobi, double, 1e-8, mtx);
  HARNESS(CG, none, float, 1e-4, mtx);
  HARNESS(CG, Jacobi, float, 1e-4, mtx);
  HARNESS(BiCGSTAB, none, double, 1e-8, mtx);
  HARNESS(BiCGSTAB, Jacobi, double, 1e-8, mtx2);
  HARNESS(BiCGSTAB, Jacobi, float, 1e-4, mtx2);
  HARNESS(Jacobi, none, double, 1e-8, mtx);
  HARNESS(Jacobi, Jacobi, float, 1e-4, mtx);
  HARNESS(CG, SOR, double, 1e-8, mtx);
  HARNESS(CG, SOR, float, 1e-4, mtx);
  HARNESS(BiCGSTAB, SOR, double, 1e-8, mtx);
  HARNESS(Jacobi, SOR, double, 1e-8, mtx);
  HARNESS(SOR, none, double, 1e-8, mtx);
  HARNESS(SOR, none, float, 1e-4, mtx);
  HARNESS(CG, IC, double, 1e-8, mtx);
  HARNESS(CG, IC, float, 1e-4, mtx);
  HARNESS(CG, ILU, double, 1e-8, mtx);
  HARNESS(CG, ILU, float, 1e-4, mtx);
  HARNESS(BiCGSTAB, ILU, double, 1e-8, mtx);
  HARNESS(BiCGSTAB, IC, float, 1e-4, mtx);
  HARNESS(Jacobi, ILU, double, 1e-8, mtx);
  HARNESS(IC, none, double, 1e-8, mtx);
  HARNESS(IC, none, float, 1e-4, mtx);
  HARNESS(ILU, none, double, 1e-8, mtx);
  HARNESS(ILU, none, float, 1e-4, mtx);
  CHECK((lobpcg_harness<double, equation::none<matrix::CRS<double>, double>>(true)));
  CHECK((lobpcg_harness<float, equation::none<matrix::CRS<float>, float>>(true)));
  CHECK((lobpcg_harness<double, equation::Jacobi<matrix::CRS<double>, double>>(true)));
  CHECK((lobpcg_harness<float, equation::Jacobi<matrix::CRS<float>, float>>(true)));
  CHECK((lobpcg_harness<double, equation::none<matrix::CRS<double>, double>>(false)));
  fused_and_callback_paths_agree<double>();
  fused_and_callback_paths_agree<float>();
  // operands on the host are refused, not computed on the CPU
  bool threw = false;
  try {
    vector<double> a(10, 1.0), c(10, 2.0);
    (void)blas::dot(a, c);
  } catch (const std::runtime_error &) {
    threw = true;
  }
  CHECK(threw);
  if (failures == 0) std::cout << "PASS" << std::endl;
  return failures == 0 ? 0 : 1;
}
