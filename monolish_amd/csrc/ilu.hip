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

// Wave-per-row factorisation for the chip-wide path (one launch per level):
// the 64 lanes of a wave hold the entries of row i (rows of at most 64
// entries); the rows k that row i reads (one per entry left of its
// diagonal, final since an earlier level) are staged in LDS in one
// coalesced pass; then the k are taken one at a time in column order, each
// from LDS.  ILU(0): the owner lane of a_ik forms l_ik = a_ik / a_kk, and
// every later entry p of row i that row k also holds (binary search of its
// column in row k's staged upper part) does a_ip -= l_ik a_kp.  IC(0): the
// lanes p < k find their column in row k's staged lower part and form the
// products l_ip l_kp; the owner lane of l_ik subtracts them in increasing
// column order, then divides by l_kk; the diagonal lane subtracts the l_ip^2
// in order and takes the square root.  Every value goes through the same
// operations, in the same order, as in the one-thread-per-row sweep, so the
// factor is bit-identical; the serial chain per row shrinks from one global
// load per merge step to one staging pass.
constexpr int kFacWaves = 4;       // rows (waves) per block
constexpr int kFacStage = 1024;    // staged entries of the rows read, per wave

__device__ __forceinline__ void fac_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// position of column c in the sorted scol[a, b), or -1
__device__ __forceinline__ int32_t fac_find(const int32_t *scol, int32_t a, int32_t b, int32_t c) {
  while (a < b) {
    const int32_t mid = (a + b) >> 1;
    const int32_t v = scol[mid];
    if (v == c) return mid;
    if (v < c)
      a = mid + 1;
    else
      b = mid;
  }
  return -1;
}

template <class T, bool IC>
__global__ __launch_bounds__(64 * kFacWaves) void k_factor_level_wave(
    const int32_t *__restrict__ rows, int32_t count, const int32_t *__restrict__ rp,
    const int32_t *__restrict__ ci, const int32_t *__restrict__ diag, T *f, int32_t *bad) {
  __shared__ int32_t scol[kFacWaves][kFacStage];
  __shared__ T sval[kFacWaves][kFacStage];
  __shared__ int32_t soff[kFacWaves][65];  // staged row j0 + a at [soff[a], soff[a + 1])
  __shared__ int32_t ssrc[kFacWaves][64];  // its first staged entry in ci / f
  __shared__ int32_t spiv[kFacWaves][64];  // the stage position of its diagonal
  __shared__ T sli[kFacWaves][64];         // IC: products, then row i's final values
  __shared__ int32_t sm[kFacWaves][64];    // IC: lane p found its column in row k
  const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
  const int32_t r = (int32_t)blockIdx.x * kFacWaves + w;
  if (r >= count) return;  // a whole wave; only wave-level syncs follow
  const int32_t i = rows[r];
  const int32_t e0 = rp[i], e1 = rp[i + 1], di = diag[i];
  const int32_t len = e1 - e0, nlow = di - e0;
  int32_t *col = scol[w];
  T *val = sval[w];
  const int32_t mycol = lane < len ? ci[e0 + lane] : INT32_MAX;
  T v = lane < len ? f[e0 + lane] : T(0);
  // lane j < nlow reads row k_j = the column of row i's entry j: ILU stages
  // its pivot and upper part [dk, end), IC its lower part and pivot [start, dk]
  const bool owns = lane < nlow;
  const int32_t ka = owns ? rp[mycol] : 0;
  const int32_t kd = owns ? diag[mycol] - ka : 0;
  const int32_t klen = owns ? rp[mycol + 1] - ka : 0;
  const int32_t pa = IC ? 0 : kd;
  const int32_t pn = owns ? (IC ? kd + 1 : klen - kd) : 0;
  int32_t incl = pn;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int32_t t = __shfl_up(incl, d, 64);
    if (lane >= d) incl += t;
  }
  const int32_t excl = incl - pn;
  bool rl = false;  // IC: the single-group right-looking form ran
  T dsum = T(0);    // IC: the diagonal's sum (lane nlow)
  for (int32_t j0 = 0; j0 < nlow;) {
    // the rows j0 .. j1-1 that fit the stage (a row of <= 64 entries always does)
    const int32_t base = __shfl(excl, j0, 64);
    const bool fits = lane >= j0 && lane < nlow && incl - base <= kFacStage;
    const int32_t j1 = j0 + (int32_t)__popcll(__ballot(fits));
    if (lane >= j0 && lane < j1) {
      soff[w][lane - j0] = excl - base;
      ssrc[w][lane - j0] = ka + pa;
      spiv[w][lane - j0] = excl - base + (IC ? kd : 0);
      if (lane == j1 - 1) soff[w][j1 - j0] = incl - base;
    }
    fac_wave_sync();
    const int32_t total = soff[w][j1 - j0];
    for (int32_t t = lane; t < total; t += 64) {
      int32_t a = 0, b = j1 - j0;  // staged row holding entry t
      while (b - a > 1) {
        const int32_t mid = (a + b) >> 1;
        if (soff[w][mid] <= t)
          a = mid;
        else
          b = mid;
      }
      const int32_t src = ssrc[w][a] + (t - soff[w][a]);
      col[t] = ci[src];
      val[t] = f[src];
    }
    fac_wave_sync();
    if (IC && j0 == 0 && j1 == nlow) {
      // Every row this row reads is staged: as each l_ip becomes final it is
      // broadcast, and each later entry j (and the diagonal) subtracts its
      // product at once.  Each entry still subtracts its products in
      // increasing column order, the same operations as the loop below,
      // without the owner lane's serial pass over all earlier entries.
      T acc = v;
      for (int32_t p = 0; p < nlow; p++) {
        const T pivot = val[spiv[w][p]];
        if (lane == p) v = acc / pivot;  // l_ip
        const T lip = __shfl(v, p, 64);
        const int32_t cp = __shfl(mycol, p, 64);
        if (lane > p && lane < nlow) {
          const int32_t q = fac_find(col, soff[w][lane], spiv[w][lane], cp);
          if (q >= 0) {
            const T prod = lip * val[q];
            acc = acc - prod;
          }
        } else if (lane == nlow) {
          const T prod = lip * lip;
          acc = acc - prod;
        }
      }
      dsum = acc;
      rl = true;
      fac_wave_sync();
      j0 = j1;
      continue;
    }
    for (int32_t j = j0; j < j1; j++) {
      const int32_t sa = soff[w][j - j0], sb = soff[w][j - j0 + 1], pv = spiv[w][j - j0];
      const T pivot = val[pv];
      if (!IC) {
        if (lane == j) v = v / pivot;  // l_ij = a_ij / a_jj
        const T lik = __shfl(v, j, 64);
        if (lane > j && lane < len) {
          const int32_t q = fac_find(col, pv + 1, sb, mycol);
          if (q >= 0) {
            const T prod = lik * val[q];
            v = v - prod;
          }
        }
      } else {
        T prod = T(0);
        int32_t hit = 0;
        if (lane < j) {
          const int32_t q = fac_find(col, sa, pv, mycol);
          if (q >= 0) {
            prod = v * val[q];
            hit = 1;
          }
        }
        sli[w][lane] = prod;
        sm[w][lane] = hit;
        fac_wave_sync();
        if (lane == j) {
          T sum = v;
          for (int32_t p = 0; p < j; p++)
            if (sm[w][p]) sum = sum - sli[w][p];
          v = sum / pivot;
        }
        fac_wave_sync();
      }
    }
    fac_wave_sync();  // the stage is read before the next group overwrites it
    j0 = j1;
  }
  if (!IC) {
    if (lane == nlow && (v == T(0) || v != v)) atomicMin(bad, i);
    if (lane < len) f[e0 + lane] = v;
  } else {
    if (!rl) {
      sli[w][lane] = v;
      fac_wave_sync();
      if (lane == nlow) {
        T sum = v;
        for (int32_t p = 0; p < nlow; p++) {
          const T prod = sli[w][p] * sli[w][p];
          sum = sum - prod;
        }
        dsum = sum;
      }
    }
    if (lane == nlow) {
      if (!(dsum > (T)0)) atomicMin(bad, i);
      v = sqrt(dsum);
    }
    if (lane <= nlow) f[e0 + lane] = v;
  }
}

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
    S.max_len = 0;
    for (int64_t i = 0; i < n; i++) S.max_len = std::max(S.max_len, rp[i + 1] - rp[i]);
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
  if (sweep_launches_per_level(lev) && S.max_len <= 64) {
    // chip-wide levels: a wave per row
    for (int32_t l = 0; l < lev.nlev; l++) {
      const int32_t cnt = lev.ptr[l + 1] - lev.ptr[l];
      const unsigned g = (unsigned)((cnt + kFacWaves - 1) / kFacWaves);
      if (kind == MGX_PRECOND_ILU)
        k_factor_level_wave<T, false><<<g, 64 * kFacWaves, 0, c->stream>>>(
            lev.d_rows + lev.ptr[l], cnt, A->rp, A->ci, S.d_diag, f, bad);
      else
        k_factor_level_wave<T, true><<<g, 64 * kFacWaves, 0, c->stream>>>(
            lev.d_rows + lev.ptr[l], cnt, A->rp, A->ci, S.d_diag, f, bad);
    }
    MGX_HIP(hipGetLastError());
  } else if (kind == MGX_PRECOND_ILU) {
    run_sweep(c, lev, IluFactorRow<T>{A->rp, A->ci, S.d_diag, f, bad});
  } else {
    run_sweep(c, lev, IcFactorRow<T>{A->rp, A->ci, S.d_diag, f, bad});
  }
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
