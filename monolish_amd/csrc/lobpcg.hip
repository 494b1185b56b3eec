// lobpcg.hip -- device kernels of the LOBPCG eigensolver
// (src/eigen/lobpcg_standard.cpp:10-300): the Gram matrices of the basis,
// the basis update, the per-pair scalings and residuals for all pairs in
// one launch each, and W = A w for all pairs in one pass over A.
//
// The basis {w, x, p} (k = 2m or 3m vectors of n entries, one vector per
// row, as the reference's Dense wxp) and its image {W, X, P} = A{w, x, p}
// are read once per iteration by k_gram, which forms the lower triangles of
// Sa = {W,X,P}{w,x,p}^T and Sb = {w,x,p}{w,x,p}^T (the reference's
// matmul(WXP, twxp) and matmul(wxp, twxp)) from per-tile partial sums in a
// fixed order.  k_lobpcg_combine forms the next x, p, X, P of all m pairs
// from the previous basis in one pass (the reference's chain of axpy calls,
// lobpcg_standard.cpp:220-238, in its order).
#include "mgx_device.hpp"

namespace mgx {

namespace {

constexpr int kGramTile = 128;  // entries of each vector per block
constexpr int kMaxPairs = 10;   // eigenpairs per solve (the entry point's limit)

// per-row factors passed by value (at most 10 pairs)
template <class T> struct RowScal {
  T v[10];
};

// lower triangle of Sa (pairs 0..npair-1) and Sb (npair..2 npair-1), pair
// (i, j), i >= j, numbered i (i + 1) / 2 + j
template <class T>
__global__ __launch_bounds__(kBlock) void k_gram(int64_t n, int k, const T *__restrict__ wxp,
                                                 const T *__restrict__ WXP, double *part) {
  extern __shared__ unsigned char smem[];
  // k rows of wxp, then k rows of WXP; rows padded by one value so the
  // lanes of a wave, which read different rows at the same entry, fall in
  // different LDS banks
  constexpr int kLd = kGramTile + 1;
  T *sv = reinterpret_cast<T *>(smem);
  const int64_t e0 = (int64_t)blockIdx.x * kGramTile;
  const int len = (int)(n - e0 < kGramTile ? n - e0 : kGramTile);
  for (int q = threadIdx.x; q < 2 * k * kGramTile; q += kBlock) {
    const int r = q / kGramTile, e = q % kGramTile;
    const T *src = r < k ? wxp + (int64_t)r * n : WXP + (int64_t)(r - k) * n;
    sv[r * kLd + e] = e < len ? src[e0 + e] : (T)0;
  }
  __syncthreads();
  const int npair = k * (k + 1) / 2;
  for (int o = threadIdx.x; o < 2 * npair; o += kBlock) {
    const int pr = o < npair ? o : o - npair;
    int i = 0;
    while ((i + 1) * (i + 2) / 2 <= pr) i++;
    const int j = pr - i * (i + 1) / 2;
    const T *a = o < npair ? sv + (k + i) * kLd : sv + i * kLd;
    const T *b = sv + j * kLd;
    double s = 0.0;
    for (int e = 0; e < len; e++) s += (double)(a[e] * b[e]);
    part[(int64_t)blockIdx.x * 2 * npair + o] = s;
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_gram_reduce(int64_t nblk, int nout,
                                                        const double *__restrict__ part,
                                                        double *out) {
  const int o = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (o >= nout) return;
  double s = 0.0;
  int64_t b = 0;
  for (; b + 8 <= nblk; b += 8) {  // eight loads in flight, added in order
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = part[(b + u) * nout + o];
#pragma unroll
    for (int u = 0; u < 8; u++) s += v[u];
  }
  for (; b < nblk; b++) s += part[b * nout + o];
  out[o] = s;
}

// coef[i * k + j] (k = the row stride, 3m): coefficient j (of w_j,
// x_{j-m}, p_{j-2m}) of pair i; has_p: the basis holds p
template <class T>
__global__ __launch_bounds__(kBlock) void k_lobpcg_combine(
    int64_t n, int m, int k, bool has_p, const T *__restrict__ coef, const T *__restrict__ wxp_p,
    const T *__restrict__ WXP_p, T *wxp, T *WXP) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= n) return;
  // basis vector j outermost: each previous value is loaded once, and every
  // output still adds its terms in the order j = 0, 1, ...
  T p[kMaxPairs], x[kMaxPairs], P[kMaxPairs], X[kMaxPairs];
#pragma unroll
  for (int i = 0; i < kMaxPairs; i++) p[i] = x[i] = P[i] = X[i] = 0;
  for (int j = 0; j < m; j++) {
    const T wj = wxp_p[(int64_t)j * n + e], Wj = WXP_p[(int64_t)j * n + e];
    const T xj = wxp_p[(int64_t)(m + j) * n + e], Xj = WXP_p[(int64_t)(m + j) * n + e];
    T pj = 0, Pj = 0;
    if (has_p) {
      pj = wxp_p[(int64_t)(2 * m + j) * n + e];
      Pj = WXP_p[(int64_t)(2 * m + j) * n + e];
    }
#pragma unroll
    for (int i = 0; i < kMaxPairs; i++) {
      if (i < m) {
        const T *b = coef + (int64_t)i * k;
        const T bw = b[j], bx = b[m + j];
        p[i] = p[i] + bw * wj;
        x[i] = x[i] + bw * wj;
        P[i] = P[i] + bw * Wj;
        X[i] = X[i] + bw * Wj;
        if (has_p) {
          const T bp = b[2 * m + j];
          p[i] = p[i] + bp * pj;
          x[i] = x[i] + bp * pj;
          P[i] = P[i] + bp * Pj;
          X[i] = X[i] + bp * Pj;
        }
        x[i] = x[i] + bx * xj;
        X[i] = X[i] + bx * Xj;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < kMaxPairs; i++) {
    if (i < m) {
      wxp[(int64_t)(m + i) * n + e] = x[i];
      wxp[(int64_t)(2 * m + i) * n + e] = p[i];
      WXP[(int64_t)(m + i) * n + e] = X[i];
      WXP[(int64_t)(2 * m + i) * n + e] = P[i];
    }
  }
}

// rows r0 .. r0+m-1 of A and of B scaled by s[r] (blas::scal of each)
template <class T>
__global__ __launch_bounds__(kBlock) void k_rows_scale2(int64_t n, int m, RowScal<T> s, T *A,
                                                        T *B) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= n) return;
  for (int r = 0; r < m; r++) {
    A[(int64_t)r * n + e] = s.v[r] * A[(int64_t)r * n + e];
    if (B) B[(int64_t)r * n + e] = s.v[r] * B[(int64_t)r * n + e];
  }
}

// pair r: x *= s, X *= s, w = -lam x + X, then w = M o w (Jacobi) when M
template <class T>
__global__ __launch_bounds__(kBlock) void k_lobpcg_xw(int64_t n, int m, RowScal<T> s,
                                                      RowScal<T> nlam, T *x, T *X, T *w,
                                                      const T *__restrict__ M) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= n) return;
  for (int r = 0; r < m; r++) {
    const int64_t o = (int64_t)r * n + e;
    const T xv = s.v[r] * x[o], Xv = s.v[r] * X[o];
    x[o] = xv;
    X[o] = Xv;
    const T t = nlam.v[r] * xv + Xv;
    w[o] = M ? M[e] * t : t;
  }
}

// Y_r = A X_r for the m rows of X (row-major, stride n): one thread per
// matrix row reads the row's entries once and keeps m sums, each added in
// storage order (the plan-based SpMV's result for rows of at most 256
// entries)
template <class T>
__global__ __launch_bounds__(kBlock) void k_spmm_rows(int64_t n, int m,
                                                      const int32_t *__restrict__ rp,
                                                      const int32_t *__restrict__ ci,
                                                      const T *__restrict__ val,
                                                      const T *__restrict__ X, T *Y) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  T s[kMaxPairs];
#pragma unroll
  for (int r = 0; r < kMaxPairs; r++) s[r] = 0;
  for (int32_t e = rp[i]; e < rp[i + 1]; e++) {
    const T a = val[e];
    const int64_t c = ci[e];
#pragma unroll
    for (int r = 0; r < kMaxPairs; r++)
      if (r < m) {
        const T p = a * X[(int64_t)r * n + c];
        s[r] = s[r] + p;
      }
  }
#pragma unroll
  for (int r = 0; r < kMaxPairs; r++)
    if (r < m) Y[(int64_t)r * n + i] = s[r];
}

}  // namespace

bool lobpcg_spmm(Csr *A, mgx_dtype dt, int m, const void *X, void *Y) {
  if (A->plan.max_row > 256 || A->rows != A->cols) return false;
  Ctx *c = A->ctx;
  const int64_t n = A->rows;
  const unsigned g = (unsigned)((n + kBlock - 1) / kBlock);
  if (n == 0) return true;
  if (dt == MGX_F64)
    k_spmm_rows<double><<<g, kBlock, 0, c->stream>>>(n, m, A->rp, A->ci, (const double *)A->val,
                                                     (const double *)X, (double *)Y);
  else
    k_spmm_rows<float><<<g, kBlock, 0, c->stream>>>(n, m, A->rp, A->ci, (const float *)A->val,
                                                    (const float *)X, (float *)Y);
  MGX_HIP(hipGetLastError());
  return true;
}

void lobpcg_scale_rows(Ctx *c, mgx_dtype dt, int64_t n, int m, const double *s, void *A,
                       void *B) {
  const unsigned g = (unsigned)((n + kBlock - 1) / kBlock);
  if (dt == MGX_F64) {
    RowScal<double> v;
    for (int r = 0; r < m; r++) v.v[r] = s[r];
    k_rows_scale2<double><<<g, kBlock, 0, c->stream>>>(n, m, v, (double *)A, (double *)B);
  } else {
    RowScal<float> v;
    for (int r = 0; r < m; r++) v.v[r] = (float)s[r];
    k_rows_scale2<float><<<g, kBlock, 0, c->stream>>>(n, m, v, (float *)A, (float *)B);
  }
  MGX_HIP(hipGetLastError());
}

void lobpcg_xw(Ctx *c, mgx_dtype dt, int64_t n, int m, const double *s, const double *nlam,
               void *x, void *X, void *w, const void *M) {
  const unsigned g = (unsigned)((n + kBlock - 1) / kBlock);
  if (dt == MGX_F64) {
    RowScal<double> a, b;
    for (int r = 0; r < m; r++) a.v[r] = s[r], b.v[r] = nlam[r];
    k_lobpcg_xw<double><<<g, kBlock, 0, c->stream>>>(n, m, a, b, (double *)x, (double *)X,
                                                     (double *)w, (const double *)M);
  } else {
    RowScal<float> a, b;
    for (int r = 0; r < m; r++) a.v[r] = (float)s[r], b.v[r] = (float)nlam[r];
    k_lobpcg_xw<float><<<g, kBlock, 0, c->stream>>>(n, m, a, b, (float *)x, (float *)X,
                                                    (float *)w, (const float *)M);
  }
  MGX_HIP(hipGetLastError());
}

void lobpcg_gram(Ctx *c, mgx_dtype dt, int64_t n, int k, const void *wxp, const void *WXP,
                 double *part, double *out) {
  const int64_t nblk = (n + kGramTile - 1) / kGramTile;
  const size_t es = dt == MGX_F64 ? 8 : 4;
  const size_t lds = (size_t)2 * k * (kGramTile + 1) * es;
  const int nout = k * (k + 1);
  if (dt == MGX_F64)
    k_gram<double><<<(unsigned)nblk, kBlock, lds, c->stream>>>(n, k, (const double *)wxp,
                                                               (const double *)WXP, part);
  else
    k_gram<float><<<(unsigned)nblk, kBlock, lds, c->stream>>>(n, k, (const float *)wxp,
                                                              (const float *)WXP, part);
  k_gram_reduce<double><<<(unsigned)((nout + kBlock - 1) / kBlock), kBlock, 0, c->stream>>>(
      nblk, nout, part, out);
  MGX_HIP(hipGetLastError());
}

void lobpcg_combine(Ctx *c, mgx_dtype dt, int64_t n, int m, int stride, bool has_p,
                    const void *coef, const void *wxp_p, const void *WXP_p, void *wxp,
                    void *WXP) {
  const unsigned g = (unsigned)((n + kBlock - 1) / kBlock);
  if (dt == MGX_F64)
    k_lobpcg_combine<double><<<g, kBlock, 0, c->stream>>>(
        n, m, stride, has_p, (const double *)coef, (const double *)wxp_p,
        (const double *)WXP_p, (double *)wxp, (double *)WXP);
  else
    k_lobpcg_combine<float><<<g, kBlock, 0, c->stream>>>(
        n, m, stride, has_p, (const float *)coef, (const float *)wxp_p, (const float *)WXP_p,
        (float *)wxp, (float *)WXP);
  MGX_HIP(hipGetLastError());
}

}  // namespace mgx
