// monolish_vml.hpp -- drop-in monolish::vml element-wise vector operations
// used by the Krylov loop: add, sub, mul, reciprocal (src/vml/vector/
// vector_vml.cpp: sub :614, mul :1198, reciprocal :4952; kernels
// src/internal/arithmetic/double_arithmetic.cpp:112-330).  Device resident
// operands only.
#pragma once

#include "monolish/common/monolish_common.hpp"

namespace monolish {
namespace vml {

namespace detail_vml {
template <class A, class B, class Y>
void vv(int (*fn)(mgx_vec *, mgx_vec *, mgx_vec *), const A &a, const B &b, Y &y,
        const char *what, const char *core) {
  detail::FuncLog log_(core);
  assert(util::is_same_size(a, b, y));
  detail::need_device(a.get_device_mem_stat() && b.get_device_mem_stat() &&
                          y.get_device_mem_stat(),
                      what);
  detail::VecArg AA(a), BB(b), YY(y);
  detail::check(fn(AA.h, BB.h, YY.h));
}
}  // namespace detail_vml

// y = a + b
template <class A, class B, class Y> void add(const A &a, const B &b, Y &y) {
  detail_vml::vv(mgx_vadd, a, b, y, "vml::add", "vvadd_core");
}
// y = a - b
template <class A, class B, class Y> void sub(const A &a, const B &b, Y &y) {
  detail_vml::vv(mgx_vsub, a, b, y, "vml::sub", "vvsub_core");
}
// y = a * b (element-wise)
template <class A, class B, class Y> void mul(const A &a, const B &b, Y &y) {
  detail_vml::vv(mgx_vmul, a, b, y, "vml::mul", "vvmul_core");
}
// y = 1 / a
template <class A, class Y> void reciprocal(const A &a, Y &y) {
  detail::FuncLog log_("vreciprocal_core");
  assert(util::is_same_size(a, y));
  detail::need_device(a.get_device_mem_stat() && y.get_device_mem_stat(), "vml::reciprocal");
  detail::VecArg AA(a), YY(y);
  detail::check(mgx_vreciprocal(AA.h, YY.h));
}

}  // namespace vml
}  // namespace monolish
