// blas1.hip -- vector BLAS / VML kernels for gfx950.
//
// Streaming kernels are HBM bound: grid-stride loops, 256 threads, every
// multiply and add rounded separately (built with -ffp-contract=off) so the
// results equal monolish's CPU loops bit for bit.  dot/nrm2 use the canonical
// tile order (mgx_device.hpp) so they are deterministic and reproducible on
// the host.
#include "mgx_device.hpp"

namespace mgx {

namespace {

inline int grid_for(Ctx *c, int64_t n, int per_thread = 4) {
  int64_t g = (n + (int64_t)kBlock * per_thread - 1) / ((int64_t)kBlock * per_thread);
  int64_t cap = (int64_t)c->num_cus * 16;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (int)g;
}

template <class T> __global__ void k_fill(T *y, int64_t n, T v) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    y[i] = v;
}

template <class T>
__global__ void k_axpy(int64_t n, T a, const T *__restrict__ x,
                       T *__restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    T t = a * x[i];
    y[i] = t + y[i];
  }
}

template <class T>
__global__ void k_xpay(int64_t n, T a, const T *__restrict__ x,
                       T *__restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    T t = a * y[i];
    y[i] = x[i] + t;
  }
}

// z may alias x or y (element-wise, each element read before written)
template <class T>
__global__ void k_axpyz(int64_t n, T a, const T *x, const T *y, T *z) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    T t = a * x[i];
    z[i] = t + y[i];
  }
}

template <class T> __global__ void k_scal(int64_t n, T a, T *x) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    x[i] = a * x[i];
}

template <class T, int OP>
__global__ void k_vv(int64_t n, const T *a, const T *b, T *y) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    T u = a[i], v = b[i];
    y[i] = OP == 0 ? u + v : (OP == 1 ? u - v : u * v);
  }
}

template <class T>
__global__ void k_reciprocal(int64_t n, const T *a, T *y) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    y[i] = T(1) / a[i];
}

template <class S, class D>
__global__ void k_convert(int64_t n, const S *x, D *y) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    y[i] = (D)x[i];
}

// canonical dot: one partial per 512-element tile, last block sums them
template <class T>
__global__ __launch_bounds__(kBlock) void k_dot(int64_t n,
                                                const T *__restrict__ x,
                                                const T *__restrict__ y,
                                                T *__restrict__ P,
                                                int64_t ntiles,
                                                uint32_t *ticket,
                                                double *out) {
  __shared__ T sh[8];
  __shared__ int flag;
  for (int64_t k = blockIdx.x; k < ntiles; k += gridDim.x) {
    const int64_t i0 = k * kTile + 2 * threadIdx.x;
    T p0 = 0, p1 = 0;
    if (i0 + 1 < n) {
      p0 = x[i0] * y[i0];
      p1 = x[i0 + 1] * y[i0 + 1];
    } else if (i0 < n) {
      p0 = x[i0] * y[i0];
    }
    T a = p0 + p1;
    T s = block_tree(a, sh);
    if (threadIdx.x == 0) P[k] = s;
    __syncthreads();
  }
  if (last_block_arrives(ticket, gridDim.x, &flag)) {
    T r = reduce_partials(P, ntiles, sh);
    if (threadIdx.x == 0) *out = (double)r;
  }
}

template <class T> T *tp(void *p) { return static_cast<T *>(p); }
template <class T> const T *tp(const void *p) {
  return static_cast<const T *>(p);
}

}  // namespace

void launch_fill(Ctx *c, mgx_dtype dt, void *y, int64_t n, double v) {
  if (n <= 0) return;
  if (dt == MGX_F64)
    k_fill<double><<<grid_for(c, n), kBlock, 0, c->stream>>>(tp<double>(y), n, v);
  else
    k_fill<float><<<grid_for(c, n), kBlock, 0, c->stream>>>(tp<float>(y), n, (float)v);
  MGX_HIP(hipGetLastError());
}

void launch_axpy(Ctx *c, mgx_dtype dt, int64_t n, double a, const void *x,
                 void *y) {
  if (n <= 0) return;
  if (dt == MGX_F64)
    k_axpy<double><<<grid_for(c, n), kBlock, 0, c->stream>>>(n, a, tp<double>(x), tp<double>(y));
  else
    k_axpy<float><<<grid_for(c, n), kBlock, 0, c->stream>>>(n, (float)a, tp<float>(x), tp<float>(y));
  MGX_HIP(hipGetLastError());
}

void launch_xpay(Ctx *c, mgx_dtype dt, int64_t n, double a, const void *x,
                 void *y) {
  if (n <= 0) return;
  if (dt == MGX_F64)
    k_xpay<double><<<grid_for(c, n), kBlock, 0, c->stream>>>(n, a, tp<double>(x), tp<double>(y));
  else
    k_xpay<float><<<grid_for(c, n), kBlock, 0, c->stream>>>(n, (float)a, tp<float>(x), tp<float>(y));
  MGX_HIP(hipGetLastError());
}

void launch_axpyz(Ctx *c, mgx_dtype dt, int64_t n, double a, const void *x,
                  const void *y, void *z) {
  if (n <= 0) return;
  if (dt == MGX_F64)
    k_axpyz<double><<<grid_for(c, n), kBlock, 0, c->stream>>>(n, a, tp<double>(x), tp<double>(y), tp<double>(z));
  else
    k_axpyz<float><<<grid_for(c, n), kBlock, 0, c->stream>>>(n, (float)a, tp<float>(x), tp<float>(y), tp<float>(z));
  MGX_HIP(hipGetLastError());
}

void launch_scal(Ctx *c, mgx_dtype dt, int64_t n, double a, void *x) {
  if (n <= 0) return;
  if (dt == MGX_F64)
    k_scal<double><<<grid_for(c, n), kBlock, 0, c->stream>>>(n, a, tp<double>(x));
  else
    k_scal<float><<<grid_for(c, n), kBlock, 0, c->stream>>>(n, (float)a, tp<float>(x));
  MGX_HIP(hipGetLastError());
}

void launch_copy(Ctx *c, mgx_dtype dt, int64_t n, const void *x, void *y) {
  if (n <= 0 || x == y) return;
  size_t bytes = (size_t)n * (dt == MGX_F64 ? 8 : 4);
  MGX_HIP(hipMemcpyAsync(y, x, bytes, hipMemcpyDeviceToDevice, c->stream));
}

void launch_vv(Ctx *c, mgx_dtype dt, int op, int64_t n, const void *a,
               const void *b, void *y) {
  if (n <= 0) return;
  const int g = grid_for(c, n);
  if (dt == MGX_F64) {
    auto A = tp<double>(a);
    auto B = tp<double>(b);
    auto Y = tp<double>(y);
    if (op == 0) k_vv<double, 0><<<g, kBlock, 0, c->stream>>>(n, A, B, Y);
    else if (op == 1) k_vv<double, 1><<<g, kBlock, 0, c->stream>>>(n, A, B, Y);
    else k_vv<double, 2><<<g, kBlock, 0, c->stream>>>(n, A, B, Y);
  } else {
    auto A = tp<float>(a);
    auto B = tp<float>(b);
    auto Y = tp<float>(y);
    if (op == 0) k_vv<float, 0><<<g, kBlock, 0, c->stream>>>(n, A, B, Y);
    else if (op == 1) k_vv<float, 1><<<g, kBlock, 0, c->stream>>>(n, A, B, Y);
    else k_vv<float, 2><<<g, kBlock, 0, c->stream>>>(n, A, B, Y);
  }
  MGX_HIP(hipGetLastError());
}

void launch_reciprocal(Ctx *c, mgx_dtype dt, int64_t n, const void *a,
                       void *y) {
  if (n <= 0) return;
  if (dt == MGX_F64)
    k_reciprocal<double><<<grid_for(c, n), kBlock, 0, c->stream>>>(n, tp<double>(a), tp<double>(y));
  else
    k_reciprocal<float><<<grid_for(c, n), kBlock, 0, c->stream>>>(n, tp<float>(a), tp<float>(y));
  MGX_HIP(hipGetLastError());
}

void launch_convert(Ctx *c, mgx_dtype from, mgx_dtype to, int64_t n,
                    const void *x, void *y) {
  if (n <= 0) return;
  if (from == to) {
    launch_copy(c, from, n, x, y);
    return;
  }
  if (from == MGX_F64)
    k_convert<double, float><<<grid_for(c, n), kBlock, 0, c->stream>>>(n, tp<double>(x), tp<float>(y));
  else
    k_convert<float, double><<<grid_for(c, n), kBlock, 0, c->stream>>>(n, tp<float>(x), tp<double>(y));
  MGX_HIP(hipGetLastError());
}

void launch_dot(Ctx *c, mgx_dtype dt, int64_t n, const void *x, const void *y,
                double *d_out, uint32_t *d_ticket) {
  if (n <= 0) {
    launch_fill(c, MGX_F64, d_out, 1, 0.0);
    return;
  }
  const int64_t nt = ntiles_of(n);
  ensure_partials(c, (size_t)nt * 8);
  int64_t g = nt;
  const int64_t cap = (int64_t)c->num_cus * 8;
  if (g > cap) g = cap;
  if (dt == MGX_F64)
    k_dot<double><<<(int)g, kBlock, 0, c->stream>>>(n, tp<double>(x), tp<double>(y),
        tp<double>(c->d_partials), nt, d_ticket, d_out);
  else
    k_dot<float><<<(int)g, kBlock, 0, c->stream>>>(n, tp<float>(x), tp<float>(y),
        tp<float>(c->d_partials), nt, d_ticket, d_out);
  MGX_HIP(hipGetLastError());
}

double dot_host(Ctx *c, mgx_dtype dt, int64_t n, const void *x,
                const void *y) {
  double *slot = &c->d_state->rr;
  launch_dot(c, dt, n, x, y, slot, &c->d_state->ticket[0]);
  MGX_HIP(hipMemcpyAsync(c->h_pinned, slot, sizeof(double),
                         hipMemcpyDeviceToHost, c->stream));
  MGX_HIP(hipStreamSynchronize(c->stream));
  return c->h_pinned[0];
}

}  // namespace mgx
