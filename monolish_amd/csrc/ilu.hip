// ilu.hip -- IC(0) and ILU(0) preconditioners on the device.
//
// equation::IC / equation::ILU (src/equation/ic.cpp:12-80, ilu.cpp:12-85)
// factor a copy of A's values in place with cuSPARSE's csric02 /
// csrilu02 on the pattern of A (no fill) and apply the preconditioner as
// two csrsv2 triangular solves (kernel/ic_kernel.hpp, ilu_kernel.hpp).
// Here the factorisation is a level-scheduled row sweep (mgx_sweep.hpp),
// one thread per row, and the solves are triangular sweep plans
// (trisolve.hip); each row does the textbook arithmetic in a fixed order:
//   ILU(0), IKJ order (Saad, Iterative Methods for Sparse Linear Systems,
//   Alg. 10.4): for the row's entries left of the diagonal in increasing
//   column k: a_ik /= a_kk, then a_ij -= a_ik a_kj for the later entries j
//   of the row that row k also holds (sorted merge of the two rows)
//   IC(0): for k < i in increasing order, l_ik = (a_ik - sum_{m<k} l_im
//   l_km) / l_kk (sorted merge), then l_ii = sqrt(a_ii - sum_{m<i} l_im^2)
// A row reads only rows of earlier levels, which are final, so the factor
// and the solves are bit-identical to the serial loops that do the same
// operations in row order.
#include <algorithm>
#include <vector>

#include "mgx_device.hpp"
#include "mgx_sweep.hpp"

namespace mgx {

namespace {

template <class T> struct IluFactorRow {
  const int32_t *__restrict__ rp;
  const int32_t *__restrict__ ci;
  const int32_t *__restrict__ diag;
  T *f;
  int32_t *bad;  // smallest row with a zero pivot
  __device__ __forceinline__ void operator()(int32_t i) const {
    const int32_t e1 = rp[i + 1], di = diag[i];
    for (int32_t e = rp[i]; e < di; e++) {
      const int32_t k = ci[e];
      const int32_t dk = diag[k];
      const T lik = f[e] / f[dk];
      f[e] = lik;
      int32_t p = e + 1, q = dk + 1;
      const int32_t qe = rp[k + 1];
      int32_t cp = p < e1 ? ci[p] : 0, cq = q < qe ? ci[q] : 0;
      while (p < e1 && q < qe) {
        if (cp == cq) {
          const T prod = lik * f[q];
          f[p] = f[p] - prod;
          p++;
          q++;
          if (p < e1) cp = ci[p];
          if (q < qe) cq = ci[q];
        } else if (cp < cq) {
          p++;
          if (p < e1) cp = ci[p];
        } else {
          q++;
          if (q < qe) cq = ci[q];
        }
      }
    }
    const T piv = f[di];
    if (piv == (T)0 || piv != piv) atomicMin(bad, i);
  }
};

template <class T> struct IcFactorRow {
  const int32_t *__restrict__ rp;
  const int32_t *__restrict__ ci;
  const int32_t *__restrict__ diag;
  T *f;
  int32_t *bad;  // smallest row with a non-positive pivot
  __device__ __forceinline__ void operator()(int32_t i) const {
    const int32_t e0 = rp[i], di = diag[i];
    for (int32_t e = e0; e < di; e++) {
      const int32_t k = ci[e];
      const int32_t dk = diag[k];
      T sum = f[e];
      int32_t p = e0, q = rp[k];
      int32_t cp = p < e ? ci[p] : 0, cq = q < dk ? ci[q] : 0;
      while (p < e && q < dk) {
        if (cp == cq) {
          const T prod = f[p] * f[q];
          sum = sum - prod;
          p++;
          q++;
          if (p < e) cp = ci[p];
          if (q < dk) cq = ci[q];
        } else if (cp < cq) {
          p++;
          if (p < e) cp = ci[p];
        } else {
          q++;
          if (q < dk) cq = ci[q];
        }
      }
      f[e] = sum / f[dk];
    }
    T sum = f[di];
    for (int32_t e = e0; e < di; e++) {
      const T prod = f[e] * f[e];
      sum = sum - prod;
    }
    if (!(sum > (T)0)) atomicMin(bad, i);
    f[di] = sqrt(sum);
  }
};

__global__ void k_fill_i32(int32_t *p, int32_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

int32_t *upload_i32(const std::vector<int32_t> &h) {
  int32_t *d = (int32_t *)dev_alloc(std::max<size_t>(h.size(), 1) * 4);
  if (!h.empty()) MGX_HIP(hipMemcpy(d, h.data(), h.size() * 4, hipMemcpyHostToDevice));
  return d;
}

// diagonal positions and the checks csric02 / csrilu02 make on the pattern
// (sorted columns, every A(i,i) stored); for IC the plan of the L^T solve
// (rows of L^T: entries (j, i) of L with j > i, in increasing j)
void ensure_struct(Csr *A, int kind) {
  IluStruct &S = A->ilu;
  const int64_t n = A->rows;
  if (S.built && (kind != MGX_PRECOND_IC || A->tri[kTriLt].built)) return;
  std::vector<int32_t> rp(n + 1), ci;
  MGX_HIP(hipStreamSynchronize(A->ctx->stream));
  MGX_HIP(hipMemcpy(rp.data(), A->rp, (size_t)(n + 1) * 4, hipMemcpyDeviceToHost));
  ci.resize(std::max<int32_t>(rp[n], 0));
  if (!ci.empty()) MGX_HIP(hipMemcpy(ci.data(), A->ci, ci.size() * 4, hipMemcpyDeviceToHost));
  std::vector<int32_t> diag(n, -1);
  for (int64_t i = 0; i < n; i++) {
    for (int32_t e = rp[i]; e < rp[i + 1]; e++) {
      MGX_REQUIRE(e == rp[i] || ci[e] > ci[e - 1],
                  "IC/ILU: column indices of row %lld are not strictly increasing",
                  (long long)i);
      if (ci[e] == i) diag[i] = e;
    }
    MGX_REQUIRE(diag[i] >= 0, "IC/ILU: A(%lld,%lld) is missing", (long long)i, (long long)i);
  }
  if (!S.built) {
    S.d_diag = upload_i32(diag);
    S.built = true;
  }
  if (kind == MGX_PRECOND_IC && !A->tri[kTriLt].built) {
    std::vector<int32_t> lt_rp(n + 1, 0);
    for (int64_t j = 0; j < n; j++)
      for (int32_t e = rp[j]; e < diag[j]; e++) lt_rp[ci[e] + 1]++;
    for (int64_t i = 0; i < n; i++) lt_rp[i + 1] += lt_rp[i];
    std::vector<int32_t> pos(lt_rp.begin(), lt_rp.end() - 1), lt_e(lt_rp[n]), lt_c(lt_rp[n]);
    for (int64_t j = 0; j < n; j++)
      for (int32_t e = rp[j]; e < diag[j]; e++) {
        lt_c[pos[ci[e]]] = (int32_t)j;
        lt_e[pos[ci[e]]++] = e;
      }
    build_tri_plan(lt_rp, lt_c, &lt_e, n, kTriUpper, A->tri[kTriLt]);
  }
}

template <class T> void factor_t(Csr *A, int kind, T *f) {
  Ctx *c = A->ctx;
  ensure_struct(A, kind);
  const IluStruct &S = A->ilu;
  int32_t *bad = (int32_t *)ws_get(c, 4);
  k_fill_i32<<<1, 1, 0, c->stream>>>(bad, INT32_MAX);
  const SorLevels &lev = tri_plan(A, kTriLower).lev;
  if (kind == MGX_PRECOND_ILU)
    run_sweep(c, lev, IluFactorRow<T>{A->rp, A->ci, S.d_diag, f, bad});
  else
    run_sweep(c, lev, IcFactorRow<T>{A->rp, A->ci, S.d_diag, f, bad});
  int32_t h = INT32_MAX;
  MGX_HIP(hipMemcpyAsync(&h, bad, 4, hipMemcpyDeviceToHost, c->stream));
  MGX_HIP(hipStreamSynchronize(c->stream));
  MGX_REQUIRE(h == INT32_MAX, "%s: %s pivot at row %d", kind == MGX_PRECOND_IC ? "IC" : "ILU",
              kind == MGX_PRECOND_IC ? "non-positive" : "zero", (int)h);
}

}  // namespace

void ilu_factor(Csr *A, mgx_dtype dt, int kind, void *f) {
  MGX_REQUIRE(A->rows == A->cols, "IC/ILU: A.row != A.col");
  if (dt == MGX_F64)
    factor_t<double>(A, kind, static_cast<double *>(f));
  else
    factor_t<float>(A, kind, static_cast<float *>(f));
}

const int32_t *ilu_diag(Csr *A) { return A->ilu.d_diag; }

void ilu_free(Csr *A) {
  dev_free(A->ilu.d_diag);
  A->ilu = IluStruct();
}

}  // namespace mgx
