// monolish_blas.hpp -- drop-in monolish::blas for the device path: CRS SpMV
// (matvec, matvec_N, matvec_T) and the vector BLAS (dot, nrm2, axpy, xpay,
// scal, copy, axpyz) on vector<T> and view1D arguments.  Mirrors
// include/monolish/blas/monolish_matvec_blas.hpp and
// monolish_vector_blas.hpp of monolish 0.17.3; the arithmetic runs in the
// gfx950 kernels behind monolish_amd.h (operands must be device resident).
#pragma once

#include "monolish/common/monolish_common.hpp"

namespace monolish {
namespace blas {

namespace detail_blas {
template <class V> using elem_t = typename std::remove_cv<typename V::value_type>::type;

template <class V> void need(const V &v, const char *what) {
  detail::need_device(v.get_device_mem_stat(), what);
}
}  // namespace detail_blas

// matvec(A, x, y): y = A x  (matvec_blas.cpp:3849-3860, crs_matvec.hpp:8-114)
template <typename Float, class V1, class V2>
void matvec(const matrix::CRS<Float> &A, const V1 &x, V2 &y) {
  detail::FuncLog log_(detail::core_name<Float>("matvec"));
  assert(A.get_row() == y.size() && A.get_col() == x.size());
  detail_blas::need(x, "matvec");
  detail_blas::need(y, "matvec");
  detail::VecArg X(x), Y(y);
  detail::check(mgx_matvec(A.dev(), X.h, Y.h));
}
// matvec(a, A, x, b, y): y = a A x + b y (the device path's meaning)
template <typename Float, class V1, class V2>
void matvec(const Float &a, const matrix::CRS<Float> &A, const V1 &x, const Float &b, V2 &y) {
  detail::FuncLog log_(detail::core_name<Float>("matvec"));
  detail_blas::need(x, "matvec");
  detail_blas::need(y, "matvec");
  detail::VecArg X(x), Y(y);
  detail::check(mgx_matvec_ab((double)a, A.dev(), X.h, (double)b, Y.h));
}
template <typename Float, class V1, class V2>
void matvec_N(const matrix::CRS<Float> &A, const V1 &x, V2 &y) {
  matvec(A, x, y);
}
template <typename Float, class V1, class V2>
void matvec_N(const Float &a, const matrix::CRS<Float> &A, const V1 &x, const Float &b, V2 &y) {
  matvec(a, A, x, b, y);
}
// matvec_T(A, x, y): y = A^T x (crs_matvec.hpp:88-98)
template <typename Float, class V1, class V2>
void matvec_T(const matrix::CRS<Float> &A, const V1 &x, V2 &y) {
  detail::FuncLog log_(detail::core_name<Float>("matvec"));
  assert(A.get_row() == x.size() && A.get_col() == y.size());
  detail_blas::need(x, "matvec_T");
  detail_blas::need(y, "matvec_T");
  detail::VecArg X(x), Y(y);
  detail::check(mgx_matvec_t(A.dev(), X.h, Y.h));
}

// dot (src/blas/vector/dot.hpp:6-78): blocking, scalar to the host
template <class V1, class V2> detail_blas::elem_t<V1> dot(const V1 &x, const V2 &y) {
  detail::FuncLog log_(detail::core_name<detail_blas::elem_t<V1>>("dot"));
  assert(util::is_same_size(x, y));
  detail_blas::need(x, "dot");
  detail_blas::need(y, "dot");
  detail::VecArg X(x), Y(y);
  double out = 0;
  detail::check(mgx_dot(X.h, Y.h, &out));
  return (detail_blas::elem_t<V1>)out;
}
template <class V1, class V2, class Float> void dot(const V1 &x, const V2 &y, Float &ans) {
  ans = dot(x, y);
}

// nrm2 (src/blas/vector/nrm2.hpp:6-67): sqrt(dot(x, x))
template <class V> detail_blas::elem_t<V> nrm2(const V &x) {
  detail::FuncLog log_(detail::core_name<detail_blas::elem_t<V>>("nrm2"));
  detail_blas::need(x, "nrm2");
  detail::VecArg X(x);
  double out = 0;
  detail::check(mgx_nrm2(X.h, &out));
  return (detail_blas::elem_t<V>)out;
}
template <class V, class Float> void nrm2(const V &x, Float &ans) { ans = nrm2(x); }

// axpy (axpy.hpp:7-61): y = a x + y
template <class Float, class V1, class V2> void axpy(const Float alpha, const V1 &x, V2 &y) {
  detail::FuncLog log_(detail::core_name<detail_blas::elem_t<V1>>("axpy"));
  assert(util::is_same_size(x, y));
  detail_blas::need(x, "axpy");
  detail_blas::need(y, "axpy");
  detail::VecArg X(x), Y(y);
  detail::check(mgx_axpy((double)alpha, X.h, Y.h));
}
// xpay (xpay.hpp:7-66): y = x + a y
template <class Float, class V1, class V2> void xpay(const Float alpha, const V1 &x, V2 &y) {
  detail::FuncLog log_(detail::core_name<detail_blas::elem_t<V1>>("xpay"));
  assert(util::is_same_size(x, y));
  detail_blas::need(x, "xpay");
  detail_blas::need(y, "xpay");
  detail::VecArg X(x), Y(y);
  detail::check(mgx_xpay((double)alpha, X.h, Y.h));
}
// scal (scal.hpp:6-49): x = a x
template <class Float, class V> void scal(const Float alpha, V &x) {
  detail::FuncLog log_(detail::core_name<detail_blas::elem_t<V>>("scal"));
  detail_blas::need(x, "scal");
  detail::VecArg X(x);
  detail::check(mgx_scal((double)alpha, X.h));
}
// copy (copy.hpp:5-13): y = x
template <class V1, class V2> void copy(const V1 &x, V2 &y) {
  detail::FuncLog log_("copy_core");
  assert(util::is_same_size(x, y));
  detail_blas::need(x, "copy");
  detail_blas::need(y, "copy");
  detail::VecArg X(x), Y(y);
  detail::check(mgx_copy(X.h, Y.h));
}
// axpyz (axpyz.hpp:7-66): z = a x + y
template <class Float, class V1, class V2, class V3>
void axpyz(const Float alpha, const V1 &x, const V2 &y, V3 &z) {
  detail::FuncLog log_(detail::core_name<detail_blas::elem_t<V1>>("axpyz"));
  assert(util::is_same_size(x, y, z));
  detail_blas::need(x, "axpyz");
  detail_blas::need(y, "axpyz");
  detail_blas::need(z, "axpyz");
  detail::VecArg X(x), Y(y), Z(z);
  detail::check(mgx_axpyz((double)alpha, X.h, Y.h, Z.h));
}

}  // namespace blas
}  // namespace monolish
