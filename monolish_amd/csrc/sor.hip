// sor.hip -- SOR sweeps on the device, bit-identical to the reference's
// serial loops (src/equation/kernel/sor_kernel.hpp:5-73), and the level
// schedules of every triangular sweep (SOR, IC / ILU).
//
// Each sweep runs level by level (mgx_sweep.hpp): rows of one level are
// independent, and each row adds its entries in storage order from the
// values the serial loop would have seen -- the same bits.
#include <algorithm>
#include <cstring>
#include <vector>

#include "mgx_device.hpp"
#include "mgx_sweep.hpp"

namespace mgx {

namespace {

// kind 0: forward, precond (sor_kernel_precond's first loop: every entry
//         left of the diagonal, z_i = (r_i - sum A_ij z_j) D_i)
// kind 1: forward, solver (sor_kernel_lower: entries in order while their
//         column is left of the row; the first other one ends the row)
// kind 2: backward, precond (second loop: entries right of the diagonal,
//         the sum kept in double, z_i -= sum * D_i)
template <class T, int KIND> struct SorRow {
  const int32_t *__restrict__ rp;
  const int32_t *__restrict__ ci;
  const T *__restrict__ val;
  const T *__restrict__ D;
  const T *__restrict__ r;
  T *z;
  __device__ __forceinline__ void operator()(int32_t i) const {
    const int32_t e0 = rp[i], e1 = rp[i + 1];
    if (KIND == 2) {
      double tmp = 0.0;
      for (int32_t j = e0; j < e1; j++) {
        const int32_t c = ci[j];
        if (i < c) {
          T p = val[j] * z[c];
          tmp = tmp + (double)p;
        }
      }
      z[i] = (T)((double)z[i] - tmp * (double)D[i]);
    } else {
      T tmp = r[i];
      for (int32_t j = e0; j < e1; j++) {
        const int32_t c = ci[j];
        if (i > c) {
          T p = val[j] * z[c];
          tmp = tmp - p;
        } else if (KIND == 1) {
          break;
        }
      }
      z[i] = tmp * D[i];
    }
  }
};

}  // namespace

// levels of one sweep kind from a host row structure
void build_levels(const std::vector<int32_t> &rp, const std::vector<int32_t> &ci, int64_t n,
                  int kind, SorLevels &out) {
  std::vector<int32_t> lev(n, 0);
  int32_t nlev = 0;
  if (kind == 2) {
    for (int64_t i = n - 1; i >= 0; i--) {
      int32_t l = 0;
      for (int32_t j = rp[i]; j < rp[i + 1]; j++)
        if (ci[j] > i) l = std::max(l, lev[ci[j]] + 1);
      lev[i] = l;
      nlev = std::max(nlev, l + 1);
    }
  } else {
    for (int64_t i = 0; i < n; i++) {
      int32_t l = 0;
      for (int32_t j = rp[i]; j < rp[i + 1]; j++) {
        if (ci[j] < i)
          l = std::max(l, lev[ci[j]] + 1);
        else if (kind == 1)
          break;
      }
      lev[i] = l;
      nlev = std::max(nlev, l + 1);
    }
  }
  out.ptr.assign(nlev + 1, 0);
  for (int64_t i = 0; i < n; i++) out.ptr[lev[i] + 1]++;
  for (int32_t l = 0; l < nlev; l++) out.ptr[l + 1] += out.ptr[l];
  std::vector<int32_t> pos(out.ptr.begin(), out.ptr.end() - 1), rows(n);
  for (int64_t i = 0; i < n; i++) rows[pos[lev[i]]++] = (int32_t)i;
  out.nlev = nlev;
  out.max_width = 0;
  for (int32_t l = 0; l < nlev; l++)
    out.max_width = std::max(out.max_width, out.ptr[l + 1] - out.ptr[l]);
  out.d_rows = (int32_t *)dev_alloc((size_t)std::max<int64_t>(n, 1) * 4);
  out.d_ptr = (int32_t *)dev_alloc((size_t)(nlev + 1) * 4);
  if (n > 0) MGX_HIP(hipMemcpy(out.d_rows, rows.data(), (size_t)n * 4, hipMemcpyHostToDevice));
  MGX_HIP(hipMemcpy(out.d_ptr, out.ptr.data(), (size_t)(nlev + 1) * 4, hipMemcpyHostToDevice));
}

const SorLevels &sor_levels(Csr *A, int kind) {
  if (!A->sor[kind].d_ptr) {
    const int64_t n = A->rows;
    std::vector<int32_t> rp(n + 1), ci;
    MGX_HIP(hipStreamSynchronize(A->ctx->stream));
    MGX_HIP(hipMemcpy(rp.data(), A->rp, (size_t)(n + 1) * 4, hipMemcpyDeviceToHost));
    ci.resize(std::max<int32_t>(rp[n], 0));
    if (!ci.empty())
      MGX_HIP(hipMemcpy(ci.data(), A->ci, ci.size() * 4, hipMemcpyDeviceToHost));
    build_levels(rp, ci, n, kind, A->sor[kind]);
  }
  return A->sor[kind];
}

namespace {

template <class T, int KIND>
void sweep_t(Csr *A, const T *val, const T *D, const T *r, T *z) {
  run_sweep(A->ctx, sor_levels(A, KIND), SorRow<T, KIND>{A->rp, A->ci, val, D, r, z});
}

template <class T> void sweep_kind(Csr *A, int kind, const T *val, const T *D, const T *r, T *z) {
  if (kind == 0)
    sweep_t<T, 0>(A, val, D, r, z);
  else if (kind == 1)
    sweep_t<T, 1>(A, val, D, r, z);
  else
    sweep_t<T, 2>(A, val, D, r, z);
}

}  // namespace

void sor_sweep(Csr *A, const void *val, mgx_dtype dt, int kind, const void *D, const void *r,
               void *z) {
  MGX_REQUIRE(A->rows == A->cols, "SOR: A.row != A.col");
  if (dt == MGX_F64)
    sweep_kind<double>(A, kind, static_cast<const double *>(val), static_cast<const double *>(D),
                       static_cast<const double *>(r), static_cast<double *>(z));
  else
    sweep_kind<float>(A, kind, static_cast<const float *>(val), static_cast<const float *>(D),
                      static_cast<const float *>(r), static_cast<float *>(z));
}

void levels_free(SorLevels &L) {
  dev_free(L.d_rows);
  dev_free(L.d_ptr);
  L = SorLevels();
}

void sor_levels_free(Csr *A) {
  for (SorLevels &L : A->sor) levels_free(L);
}

}  // namespace mgx
