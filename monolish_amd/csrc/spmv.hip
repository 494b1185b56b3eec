// spmv.hip -- row-length-adaptive CRS SpMV for gfx950.
//
// The matrix is cut, once at upload, into a list of row blocks:
//  * stream blocks: up to 256 consecutive rows whose entries (at most the
//    plan's entry budget) are streamed by the whole block with coalesced loads;
//    each product val*x[col] goes to LDS, then one lane per row adds its
//    products in CRS order starting from 0.  That is the order and rounding of
//    monolish's CPU row loop (crs_matvec.hpp:100-108), so y matches it bit
//    for bit.
//  * long blocks: one row longer than the long-row threshold; the 256 lanes
//    stride over it and the partial sums meet in the canonical block tree
//    (deterministic, within the forward-error bound of the CPU loop).
// Descriptors are (first row, first entry) pairs; the sign bit of the row
// marks a long block.  The kernel is HBM bound (roofline: nnz*(sv+4) +
// (rows+1)*4 + (rows+cols)*sv bytes); x gathers are served from L2/MALL.
#include <algorithm>
#include <cstring>

#include "mgx_device.hpp"

namespace mgx {

namespace {

// U = entry budget / 256: each lane issues all of its val/col loads, then all
// of its x gathers, before it multiplies, so up to U loads per lane are in
// flight at once.
template <class T, bool AB, int U>
__global__ __launch_bounds__(kBlock) void k_spmv(
    const int2 *__restrict__ blk, int64_t nblocks,
    const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
    const T *__restrict__ val, const T *__restrict__ x, T *__restrict__ y,
    T alpha, T beta, int remap, const int32_t *live) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *prod = reinterpret_cast<T *>(smem_raw);
  __shared__ T sh[8];
  const int64_t b = xcd_block(remap);
  if (b >= nblocks) return;
  // inside a device-driven solve: nothing to do once the solve has stopped
  if (live && *live != kStatRunning) return;
  // descriptor: (first row | long-row flag, first entry)
  const int2 d0 = blk[b], d1 = blk[b + 1];
  const int32_t r0 = d0.x & 0x7fffffff, r1 = d1.x & 0x7fffffff;
  const int32_t e0 = d0.y, e1 = d1.y;
  if (d0.x < 0) {
    T s = 0;
    for (int32_t e = e0 + (int32_t)threadIdx.x; e < e1; e += kBlock) {
      T p = val[e] * x[ci[e]];
      s = s + p;
    }
    T r = block_tree(s, sh);
    if (threadIdx.x == 0) {
      if (AB) {
        T t = alpha * r;
        if (beta != T(0)) {
          T u = beta * y[r0];
          t = t + u;
        }
        y[r0] = t;
      } else {
        y[r0] = r;
      }
    }
    return;
  }
  // this lane's row bounds are read before the products, so their latency
  // overlaps the entry loads
  const int32_t r = r0 + (int32_t)threadIdx.x;
  int32_t ra = 0, rb = 0;
  if (r < r1) {
    ra = rp[r];
    rb = rp[r + 1];
  }
  stream_products<T, U>(val, ci, e0, e1, prod, [&](int32_t c) { return x[c]; });
  __syncthreads();
  if (r < r1) {
    const int32_t a = ra - e0, bnd = rb - e0;
    T s = 0;
    for (int32_t j = a; j < bnd; j++) s = s + prod[j];
    if (AB) {
      T t = alpha * s;
      if (beta != T(0)) {
        T u = beta * y[r];
        t = t + u;
      }
      y[r] = t;
    } else {
      y[r] = s;
    }
  }
}

// CRS::diag: d[i] = A(i,i), last stored match wins, 0 if absent.  A block
// owns 256 consecutive rows and scans their entries with consecutive lanes on
// consecutive entries (coalesced); the owning row of an entry comes from a
// binary search of the block's row_ptr slice in LDS, and the largest matching
// entry index per row (= the last stored match) is kept with an LDS atomicMax.
template <class T>
__global__ __launch_bounds__(kBlock) void k_diag(int64_t n, const int32_t *__restrict__ rp,
                                                 const int32_t *__restrict__ ci,
                                                 const T *__restrict__ val,
                                                 T *__restrict__ d) {
  __shared__ int32_t srp[kBlock + 1];
  __shared__ int32_t best[kBlock];
  const int64_t r0 = (int64_t)blockIdx.x * kBlock;
  const int nr = (int)(n - r0 < kBlock ? n - r0 : kBlock);
  const int t = threadIdx.x;
  // nr + 1 row pointers (up to 257) from 256 lanes
  if (t < nr) srp[t] = rp[r0 + t];
  if (t == 0) srp[nr] = rp[r0 + nr];
  best[t] = -1;
  __syncthreads();
  const int32_t e0 = srp[0], e1 = srp[nr];
  for (int32_t j = e0 + t; j < e1; j += kBlock) {
    // largest local row lr with srp[lr] <= j
    int lo = 0, hi = nr - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (srp[mid] <= j)
        lo = mid;
      else
        hi = mid - 1;
    }
    if (ci[j] == (int32_t)(r0 + lo)) atomicMax(&best[lo], j);
  }
  __syncthreads();
  if (t < nr) {
    const int32_t b = best[t];
    d[r0 + t] = b >= 0 ? val[b] : T(0);
  }
}

}  // namespace

void spmv_plan_build(Ctx *c, Csr *A, const int32_t *rp) {
  spmv_plan_build_with(A, rp, c->spmv_rows, c->spmv_nnz, c->spmv_long, c->xcd_remap);
}

void spmv_plan_build_with(Csr *A, const int32_t *rp, int rows_req, int nnz_req, int long_req,
                          int remap_req) {
  Ctx *c = A->ctx;
  const int64_t rows = A->rows;
  const int rows_max = std::min(std::max(rows_req, 1), kBlock);
  const int nnz_max = std::max(nnz_req, 64);
  const int long_t = std::min(std::max(long_req, 1), nnz_max);
  std::vector<int2> blk;
  blk.reserve((size_t)(rows / std::max(rows_max, 1)) + 16);
  int32_t mx = 0, mn = rows > 0 ? INT32_MAX : 0;
  int64_t nlong = 0, r = 0;
  while (r < rows) {
    int32_t len = rp[r + 1] - rp[r];
    mx = std::max(mx, len);
    mn = std::min(mn, len);
    if (len > long_t) {
      blk.push_back(make_int2((int32_t)r | (int32_t)0x80000000u, rp[r]));
      nlong++;
      r++;
      continue;
    }
    const int64_t start = r;
    int64_t acc = 0;
    while (r < rows) {
      len = rp[r + 1] - rp[r];
      if (len > long_t || r - start >= rows_max || acc + len > nnz_max) break;
      mx = std::max(mx, len);
      mn = std::min(mn, len);
      acc += len;
      r++;
    }
    blk.push_back(make_int2((int32_t)start, rp[start]));
  }
  const int64_t nb = (int64_t)blk.size();
  blk.push_back(make_int2((int32_t)rows, rp[rows]));
  A->plan.nblocks = nb;
  A->plan.nlong = nlong;
  A->plan.max_row = mx;
  A->plan.min_row = mn;
  A->plan.mean_row = rows > 0 ? (double)A->nnz / (double)rows : 0.0;
  A->plan.nnz_max = nnz_max;
  A->plan.rows_max = rows_max;
  A->plan.long_row = long_t;
  A->plan.remap = remap_req ? 1 : 0;
  A->plan.tuned_us = 0;
  if (A->plan.d_blocks) dev_free(A->plan.d_blocks);
  A->plan.d_blocks = (int2 *)dev_alloc(blk.size() * sizeof(int2));
  MGX_HIP(hipMemcpyAsync(A->plan.d_blocks, blk.data(), blk.size() * sizeof(int2),
                         hipMemcpyHostToDevice, c->stream));
  MGX_HIP(hipStreamSynchronize(c->stream));
}

template <class T, bool AB>
static void spmv_dispatch(Csr *A, const T *V, const T *X, T *Y, T alpha,
                          T beta, const int32_t *live) {
  Ctx *c = A->ctx;
  const int64_t nb = A->plan.nblocks;
  const int remap = A->plan.remap && nb >= 64;
  const unsigned grid = (unsigned)(remap ? ((nb + 7) / 8) * 8 : nb);
  const size_t lds = (size_t)A->plan.nnz_max * sizeof(T);
  // vector loads per lane: cover nnz_max entries plus the alignment slack
  constexpr int W = 16 / sizeof(T);
  const int u = (A->plan.nnz_max + W - 1 + kBlock * W - 1) / (kBlock * W);  // <= 17
  const int2 *B = A->plan.d_blocks;
#define MGX_SPMV_U(UU)                                                          \
  k_spmv<T, AB, UU><<<grid, kBlock, lds, c->stream>>>(B, nb, A->rp, A->ci, V, X, \
                                                      Y, alpha, beta, remap, live)
  if (u <= 1) MGX_SPMV_U(1);
  else if (u <= 2) MGX_SPMV_U(2);
  else if (u <= 3) MGX_SPMV_U(3);
  else if (u <= 4) MGX_SPMV_U(4);
  else if (u <= 5) MGX_SPMV_U(5);
  else if (u <= 6) MGX_SPMV_U(6);
  else if (u <= 8) MGX_SPMV_U(8);
  else if (u <= 12) MGX_SPMV_U(12);
  else MGX_SPMV_U(17);
#undef MGX_SPMV_U
  MGX_HIP(hipGetLastError());
}

void launch_spmv(Csr *A, const void *val, mgx_dtype dt, const void *x,
                 void *y, double alpha, double beta, bool ab, const int32_t *live) {
  if (A->plan.nblocks <= 0) return;
  if (dt == MGX_F64) {
    auto V = static_cast<const double *>(val);
    auto X = static_cast<const double *>(x);
    auto Y = static_cast<double *>(y);
    if (ab)
      spmv_dispatch<double, true>(A, V, X, Y, alpha, beta, live);
    else
      spmv_dispatch<double, false>(A, V, X, Y, 1.0, 0.0, live);
  } else {
    auto V = static_cast<const float *>(val);
    auto X = static_cast<const float *>(x);
    auto Y = static_cast<float *>(y);
    if (ab)
      spmv_dispatch<float, true>(A, V, X, Y, (float)alpha, (float)beta, live);
    else
      spmv_dispatch<float, false>(A, V, X, Y, 1.0f, 0.0f, live);
  }
}

void launch_diag_v(Csr *A, const void *val, mgx_dtype dt, void *d) {
  Ctx *c = A->ctx;
  const int64_t n = std::min(A->rows, A->cols);
  if (n <= 0) return;
  const int64_t g = (n + kBlock - 1) / kBlock;  // one block per 256 rows
  if (dt == MGX_F64)
    k_diag<double><<<(unsigned)g, kBlock, 0, c->stream>>>(
        n, A->rp, A->ci, static_cast<const double *>(val),
        static_cast<double *>(d));
  else
    k_diag<float><<<(unsigned)g, kBlock, 0, c->stream>>>(
        n, A->rp, A->ci, static_cast<const float *>(val),
        static_cast<float *>(d));
  MGX_HIP(hipGetLastError());
}

void launch_diag(Csr *A, void *d) { launch_diag_v(A, A->val, A->dtype, d); }

// Autotuner: times A's SpMV under candidate block plans (rows and entries
// per stream block, XCD order) and keeps the fastest.  The long-row
// threshold is kept, so every candidate computes the same bits.
void spmv_autotune(Csr *A, int reps, mgx_spmv_tuning *out) {
  Ctx *c = A->ctx;
  const int64_t rows = A->rows, cols = A->cols;
  std::vector<int32_t> rp(rows + 1);
  MGX_HIP(hipMemcpy(rp.data(), A->rp, (size_t)(rows + 1) * 4, hipMemcpyDeviceToHost));
  const size_t es = A->elem();
  void *x = dev_alloc((size_t)std::max<int64_t>(cols, 1) * es);
  void *y = dev_alloc((size_t)std::max<int64_t>(rows, 1) * es);
  MGX_HIP(hipMemsetAsync(x, 0, (size_t)std::max<int64_t>(cols, 1) * es, c->stream));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const int long_t = A->plan.long_row > 0 ? A->plan.long_row : c->spmv_long;
  struct Cand {
    int rows_max, nnz_max, remap;
    double us;
  };
  std::vector<Cand> cands;
  for (int rm : {64, 128, 256})
    for (int nm : {1024, 2048, 4096})
      for (int rr : {1, 0})
        if (nm >= long_t) cands.push_back(Cand{rm, nm, rr, 0.0});
  try {
    MGX_HIP(hipEventCreate(&e0));
    MGX_HIP(hipEventCreate(&e1));
    for (Cand &k : cands) {
      spmv_plan_build_with(A, rp.data(), k.rows_max, k.nnz_max, long_t, k.remap);
      for (int i = 0; i < 3; i++) launch_spmv(A, A->val, A->dtype, x, y, 1, 0, false);
      MGX_HIP(hipEventRecord(e0, c->stream));
      for (int i = 0; i < reps; i++) launch_spmv(A, A->val, A->dtype, x, y, 1, 0, false);
      MGX_HIP(hipEventRecord(e1, c->stream));
      MGX_HIP(hipEventSynchronize(e1));
      float ms = 0;
      MGX_HIP(hipEventElapsedTime(&ms, e0, e1));
      k.us = 1e3 * ms / reps;
    }
  } catch (...) {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dev_free(x);
    dev_free(y);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dev_free(x);
  dev_free(y);
  const Cand best = *std::min_element(cands.begin(), cands.end(),
                                      [](const Cand &a, const Cand &b) { return a.us < b.us; });
  spmv_plan_build_with(A, rp.data(), best.rows_max, best.nnz_max, long_t, best.remap);
  A->plan.tuned_us = best.us;
  if (out) {
    out->rows_per_block = best.rows_max;
    out->nnz_per_block = best.nnz_max;
    out->long_row = long_t;
    out->xcd_remap = best.remap;
    out->us = best.us;
    out->candidates = (int32_t)cands.size();
  }
}

}  // namespace mgx
