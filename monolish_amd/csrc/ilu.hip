// ilu.hip -- IC(0) and ILU(0) preconditioners on the device.
//
// equation::IC / equation::ILU (src/equation/ic.cpp:12-80, ilu.cpp:12-85)
// factor a copy of A's values in place with cuSPARSE's csric02 /
// csrilu02 on the pattern of A (no fill) and apply the preconditioner as
// two csrsv2 triangular solves (kernel/ic_kernel.hpp, ilu_kernel.hpp).
// Here both the factorisation and the solves are level-scheduled row
// sweeps (mgx_sweep.hpp), one thread per row, each row doing the textbook
// arithmetic in a fixed order:
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

// t_i = (r_i - sum_{j<i} l_ij t_j) [/ l_ii unless UNIT]
template <class T, bool UNIT> struct LowerRow {
  const int32_t *__restrict__ rp;
  const int32_t *__restrict__ ci;
  const int32_t *__restrict__ diag;
  const T *__restrict__ f;
  const T *__restrict__ r;
  T *t;
  __device__ __forceinline__ void operator()(int32_t i) const {
    const int32_t di = diag[i];
    T sum = r[i];
    for (int32_t e = rp[i]; e < di; e++) {
      const T prod = f[e] * t[ci[e]];
      sum = sum - prod;
    }
    t[i] = UNIT ? sum : sum / f[di];
  }
};

// ILU: z_i = (t_i - sum_{j>i} u_ij z_j) / u_ii
template <class T> struct UpperRow {
  const int32_t *__restrict__ rp;
  const int32_t *__restrict__ ci;
  const int32_t *__restrict__ diag;
  const T *__restrict__ f;
  const T *__restrict__ t;
  T *z;
  __device__ __forceinline__ void operator()(int32_t i) const {
    const int32_t di = diag[i];
    T sum = t[i];
    for (int32_t e = di + 1; e < rp[i + 1]; e++) {
      const T prod = f[e] * z[ci[e]];
      sum = sum - prod;
    }
    z[i] = sum / f[di];
  }
};

// IC: z_i = (t_i - sum_{j>i} l_ji z_j) / l_ii over the rows of L^T
template <class T> struct LtRow {
  const int32_t *__restrict__ lt_rp;
  const int32_t *__restrict__ lt_e;
  const int32_t *__restrict__ lt_c;
  const int32_t *__restrict__ diag;
  const T *__restrict__ f;
  const T *__restrict__ t;
  T *z;
  __device__ __forceinline__ void operator()(int32_t i) const {
    T sum = t[i];
    for (int32_t k = lt_rp[i]; k < lt_rp[i + 1]; k++) {
      const T prod = f[lt_e[k]] * z[lt_c[k]];
      sum = sum - prod;
    }
    z[i] = sum / f[diag[i]];
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

// diagonal positions, the checks csric02 / csrilu02 make on the pattern
// (sorted columns, every A(i,i) stored), the schedules of the forward and
// backward sweeps, and for IC the rows of L^T
void ensure_struct(Csr *A, int kind) {
  IluStruct &S = A->ilu;
  const int64_t n = A->rows;
  if (!S.built) {
    std::vector<int32_t> rp(n + 1), ci;
    MGX_HIP(hipStreamSynchronize(A->ctx->stream));
    MGX_HIP(hipMemcpy(rp.data(), A->rp, (size_t)(n + 1) * 4, hipMemcpyDeviceToHost));
    ci.resize(std::max<int32_t>(rp[n], 0));
    if (!ci.empty())
      MGX_HIP(hipMemcpy(ci.data(), A->ci, ci.size() * 4, hipMemcpyDeviceToHost));
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
    // rows of L^T: entries (j, i), j > i, in increasing j
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
    S.d_diag = upload_i32(diag);
    S.d_lt_rp = upload_i32(lt_rp);
    S.d_lt_e = upload_i32(lt_e);
    S.d_lt_c = upload_i32(lt_c);
    build_levels(lt_rp, lt_c, n, 2, S.lt);
    S.built = true;
  }
  sor_levels(A, 0);
  if (kind == MGX_PRECOND_ILU) sor_levels(A, 2);
}

template <class T> void factor_t(Csr *A, int kind, T *f) {
  Ctx *c = A->ctx;
  ensure_struct(A, kind);
  const IluStruct &S = A->ilu;
  int32_t *bad = (int32_t *)ws_get(c, 4);
  k_fill_i32<<<1, 1, 0, c->stream>>>(bad, INT32_MAX);
  if (kind == MGX_PRECOND_ILU)
    run_sweep(c, sor_levels(A, 0), IluFactorRow<T>{A->rp, A->ci, S.d_diag, f, bad});
  else
    run_sweep(c, sor_levels(A, 0), IcFactorRow<T>{A->rp, A->ci, S.d_diag, f, bad});
  int32_t h = INT32_MAX;
  MGX_HIP(hipMemcpyAsync(&h, bad, 4, hipMemcpyDeviceToHost, c->stream));
  MGX_HIP(hipStreamSynchronize(c->stream));
  MGX_REQUIRE(h == INT32_MAX, "%s: %s pivot at row %d", kind == MGX_PRECOND_IC ? "IC" : "ILU",
              kind == MGX_PRECOND_IC ? "non-positive" : "zero", (int)h);
}

template <class T>
void apply_t(Csr *A, int kind, const T *f, const T *r, T *t, T *z) {
  Ctx *c = A->ctx;
  ensure_struct(A, kind);
  const IluStruct &S = A->ilu;
  const SorLevels &lo = sor_levels(A, 0);
  if (kind == MGX_PRECOND_ILU) {
    run_sweep(c, lo, LowerRow<T, true>{A->rp, A->ci, S.d_diag, f, r, t});
    run_sweep(c, sor_levels(A, 2), UpperRow<T>{A->rp, A->ci, S.d_diag, f, t, z});
  } else {
    run_sweep(c, lo, LowerRow<T, false>{A->rp, A->ci, S.d_diag, f, r, t});
    run_sweep(c, S.lt, LtRow<T>{S.d_lt_rp, S.d_lt_e, S.d_lt_c, S.d_diag, f, t, z});
  }
}

}  // namespace

void ilu_factor(Csr *A, mgx_dtype dt, int kind, void *f) {
  MGX_REQUIRE(A->rows == A->cols, "IC/ILU: A.row != A.col");
  if (dt == MGX_F64)
    factor_t<double>(A, kind, static_cast<double *>(f));
  else
    factor_t<float>(A, kind, static_cast<float *>(f));
}

void ilu_apply(Csr *A, mgx_dtype dt, int kind, const void *f, const void *r, void *t, void *z) {
  if (dt == MGX_F64)
    apply_t<double>(A, kind, static_cast<const double *>(f), static_cast<const double *>(r),
                    static_cast<double *>(t), static_cast<double *>(z));
  else
    apply_t<float>(A, kind, static_cast<const float *>(f), static_cast<const float *>(r),
                   static_cast<float *>(t), static_cast<float *>(z));
}

void ilu_free(Csr *A) {
  IluStruct &S = A->ilu;
  dev_free(S.d_diag);
  dev_free(S.d_lt_rp);
  dev_free(S.d_lt_e);
  dev_free(S.d_lt_c);
  levels_free(S.lt);
  S = IluStruct();
}

}  // namespace mgx
