// blas1.hip -- vector BLAS / VML kernels for gfx950.
//
// Element-wise kernels are HBM bound.  Each lane moves 16 bytes per operand
// (2 doubles or 4 floats) with non-temporal loads and stores when all operands
// are 16-byte aligned (a scalar variant handles views at odd offsets), one
// pass, no grid-stride loop.  Every multiply and add rounds separately (built
// with -ffp-contract=off), so results equal monolish's CPU loops bit for bit.
//
// dot/nrm2 follow the canonical tile order (mgx_device.hpp): one block per
// 512-element tile, whose partial is the block tree of the tile; the blocks
// then meet in arrive_reduce, which runs stage 2 (lane sums in order, final
// block tree) spread over the last-arriving blocks.
#include <mutex>

#include "mgx_device.hpp"

namespace mgx {

namespace {

enum EwOp { EW_FILL, EW_AXPY, EW_XPAY, EW_SCAL, EW_ADD, EW_SUB, EW_MUL, EW_RECIP };

template <int OP, class T> __device__ __forceinline__ T ew(T a, T x, T y) {
  if (OP == EW_FILL) return a;
  if (OP == EW_AXPY) {  // y = a x + y
    T t = a * x;
    return t + y;
  }
  if (OP == EW_XPAY) {  // y = x + a y
    T t = a * y;
    return x + t;
  }
  if (OP == EW_SCAL) return a * x;
  if (OP == EW_ADD) return x + y;
  if (OP == EW_SUB) return x - y;
  if (OP == EW_MUL) return x * y;
  return T(1) / x;  // EW_RECIP
}

template <int OP> constexpr bool reads_x() { return OP != EW_FILL; }
template <int OP> constexpr bool reads_y() {
  return OP == EW_AXPY || OP == EW_XPAY || OP == EW_ADD || OP == EW_SUB || OP == EW_MUL;
}

// z[i] = op(a, x[i], y[i]); W elements per lane (W > 1 requires 16-byte
// aligned x, y, z); z may alias x or y
template <class T, int OP, int W>
__global__ __launch_bounds__(kBlock) void k_ew(int64_t n, T a, const T *x,
                                               const T *y, T *z) {
  const int64_t i = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * W;
  if (W > 1 && i + W <= n) {
    typedef T V __attribute__((ext_vector_type(W)));
    V xv = {}, yv = {}, r;
    if (reads_x<OP>()) xv = __builtin_nontemporal_load(reinterpret_cast<const V *>(x + i));
    if (reads_y<OP>()) yv = __builtin_nontemporal_load(reinterpret_cast<const V *>(y + i));
#pragma unroll
    for (int j = 0; j < W; j++) r[j] = ew<OP>(a, (T)xv[j], (T)yv[j]);
    __builtin_nontemporal_store(r, reinterpret_cast<V *>(z + i));
  } else {
    for (int64_t j = i; j < n && j < i + W; j++) {
      T xs = reads_x<OP>() ? x[j] : T(0);
      T ys = reads_y<OP>() ? y[j] : T(0);
      z[j] = ew<OP>(a, xs, ys);
    }
  }
}

template <class S, class D>
__global__ void k_convert(int64_t n, const S *x, D *y) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    y[i] = (D)x[i];
}

// canonical dot: one block per 512-element tile (lane t holds elements 2t
// and 2t+1), the tile partial from the block tree; the blocks then meet in
// arrive_reduce, which adds the tile partials in the canonical lane order
template <class T, bool VEC, bool SQ = false>
__global__ __launch_bounds__(kBlock) void k_dot(int64_t n, const T *x, const T *y, T *P,
                                                T *S, uint32_t *tick, double *out) {
  __shared__ T sh[8];
  __shared__ T tp[kBlock];
  __shared__ T res[2];
  __shared__ int flag;
  const int64_t tile = blockIdx.x;
  const int64_t i0 = tile * kTile + 2 * threadIdx.x;
  T p0 = 0, p1 = 0;
  if (VEC && i0 + 1 < n) {
    typedef T V2 __attribute__((ext_vector_type(2)));
    const V2 xv = __builtin_nontemporal_load(reinterpret_cast<const V2 *>(x + i0));
    const V2 yv = __builtin_nontemporal_load(reinterpret_cast<const V2 *>(y + i0));
    p0 = xv[0] * yv[0];
    p1 = xv[1] * yv[1];
  } else {
    if (i0 < n) p0 = x[i0] * y[i0];
    if (i0 + 1 < n) p1 = x[i0 + 1] * y[i0 + 1];
  }
  const T s = block_tree(p0 + p1, sh);
  if (threadIdx.x == 0) publish_partial(P, tile, s);
  if (arrive_reduce<T, 1, 1>(tick, P, 0, S, ntiles_of(n), tile, tp, sh, res, &flag))
    if (threadIdx.x == 0) *out = SQ ? (double)sqrt(res[0]) : (double)res[0];
}

// m canonical dots <x_r, x_r> of the rows of a row-major block (row r at
// X + r ld), one per blockIdx.y, each exactly as k_dot computes it
template <class T, bool VEC>
__global__ __launch_bounds__(kBlock) void k_mdot(int64_t n, const T *X, int64_t ld, T *P,
                                                 int64_t pst, uint32_t *tick, double *out) {
  __shared__ T sh[8];
  __shared__ T tp[kBlock];
  __shared__ T res[2];
  __shared__ int flag;
  const int64_t r = blockIdx.y;
  const T *x = X + r * ld;
  T *Pr = P + r * pst;
  const int64_t nt = ntiles_of(n);
  const int64_t tile = blockIdx.x;
  const int64_t i0 = tile * kTile + 2 * threadIdx.x;
  T p0 = 0, p1 = 0;
  if (VEC && i0 + 1 < n) {
    typedef T V2 __attribute__((ext_vector_type(2)));
    const V2 xv = __builtin_nontemporal_load(reinterpret_cast<const V2 *>(x + i0));
    p0 = xv[0] * xv[0];
    p1 = xv[1] * xv[1];
  } else {
    if (i0 < n) p0 = x[i0] * x[i0];
    if (i0 + 1 < n) p1 = x[i0 + 1] * x[i0 + 1];
  }
  const T s = block_tree(p0 + p1, sh);
  if (threadIdx.x == 0) publish_partial(Pr, tile, s);
  if (arrive_reduce<T, 1, 1>(tick + r * kTickWords, Pr, 0, Pr + nt, nt, tile, tp, sh, res,
                             &flag))
    if (threadIdx.x == 0) out[r] = (double)res[0];
}

template <class T> T *tp(void *p) { return static_cast<T *>(p); }
template <class T> const T *tp(const void *p) {
  return static_cast<const T *>(p);
}

inline bool aligned16(const void *p) {
  return p == nullptr || (reinterpret_cast<uintptr_t>(p) & 15) == 0;
}

template <class T, int OP>
void ew_launch_t(Ctx *c, int64_t n, T a, const void *x, const void *y, void *z) {
  constexpr int W = 16 / sizeof(T);
  const bool vec = aligned16(x) && aligned16(y) && aligned16(z);
  if (vec) {
    const int64_t g = (n + (int64_t)kBlock * W - 1) / ((int64_t)kBlock * W);
    k_ew<T, OP, W><<<(unsigned)g, kBlock, 0, c->stream>>>(n, a, tp<T>(x), tp<T>(y), tp<T>(z));
  } else {
    const int64_t g = (n + kBlock - 1) / kBlock;
    k_ew<T, OP, 1><<<(unsigned)g, kBlock, 0, c->stream>>>(n, a, tp<T>(x), tp<T>(y), tp<T>(z));
  }
  MGX_HIP(hipGetLastError());
}

template <int OP>
void ew_launch(Ctx *c, mgx_dtype dt, int64_t n, double a, const void *x, const void *y,
               void *z) {
  if (n <= 0) return;
  if (dt == MGX_F64)
    ew_launch_t<double, OP>(c, n, a, x, y, z);
  else
    ew_launch_t<float, OP>(c, n, (float)a, x, y, z);
}

}  // namespace

void launch_fill(Ctx *c, mgx_dtype dt, void *y, int64_t n, double v) {
  ew_launch<EW_FILL>(c, dt, n, v, nullptr, nullptr, y);
}

void launch_axpy(Ctx *c, mgx_dtype dt, int64_t n, double a, const void *x,
                 void *y) {
  ew_launch<EW_AXPY>(c, dt, n, a, x, y, y);
}

void launch_xpay(Ctx *c, mgx_dtype dt, int64_t n, double a, const void *x,
                 void *y) {
  ew_launch<EW_XPAY>(c, dt, n, a, x, y, y);
}

void launch_axpyz(Ctx *c, mgx_dtype dt, int64_t n, double a, const void *x,
                  const void *y, void *z) {
  ew_launch<EW_AXPY>(c, dt, n, a, x, y, z);
}

void launch_scal(Ctx *c, mgx_dtype dt, int64_t n, double a, void *x) {
  ew_launch<EW_SCAL>(c, dt, n, a, x, nullptr, x);
}

void launch_copy(Ctx *c, mgx_dtype dt, int64_t n, const void *x, void *y) {
  if (n <= 0 || x == y) return;
  size_t bytes = (size_t)n * (dt == MGX_F64 ? 8 : 4);
  MGX_HIP(hipMemcpyAsync(y, x, bytes, hipMemcpyDeviceToDevice, c->stream));
}

void launch_vv(Ctx *c, mgx_dtype dt, int op, int64_t n, const void *a,
               const void *b, void *y) {
  if (op == 0)
    ew_launch<EW_ADD>(c, dt, n, 0.0, a, b, y);
  else if (op == 1)
    ew_launch<EW_SUB>(c, dt, n, 0.0, a, b, y);
  else
    ew_launch<EW_MUL>(c, dt, n, 0.0, a, b, y);
}

void launch_reciprocal(Ctx *c, mgx_dtype dt, int64_t n, const void *a,
                       void *y) {
  ew_launch<EW_RECIP>(c, dt, n, 0.0, a, nullptr, y);
}

void launch_convert(Ctx *c, mgx_dtype from, mgx_dtype to, int64_t n,
                    const void *x, void *y) {
  if (n <= 0) return;
  if (from == to) {
    launch_copy(c, from, n, x, y);
    return;
  }
  int64_t g = std::min<int64_t>((n + kBlock * 4 - 1) / (kBlock * 4), (int64_t)c->num_cus * 16);
  if (from == MGX_F64)
    k_convert<double, float><<<(unsigned)g, kBlock, 0, c->stream>>>(n, tp<double>(x), tp<float>(y));
  else
    k_convert<float, double><<<(unsigned)g, kBlock, 0, c->stream>>>(n, tp<float>(x), tp<double>(y));
  MGX_HIP(hipGetLastError());
}

// the canonical dot into *d_out with the tile partials and lane sums in
// `partials` ((ntiles + 2 * kRedGroups) values) and zeroed counters d_ticket
template <class T, bool SQ>
static void dot_kernel(Ctx *c, int64_t n, const void *x, const void *y, void *partials,
                       double *d_out, uint32_t *d_ticket) {
  const int64_t nt = ntiles_of(n);
  T *P = tp<T>(partials), *S = P + nt;
  if (aligned16(x) && aligned16(y))
    k_dot<T, true, SQ><<<(unsigned)nt, kBlock, 0, c->stream>>>(n, tp<T>(x), tp<T>(y), P, S,
                                                               d_ticket, d_out);
  else
    k_dot<T, false, SQ><<<(unsigned)nt, kBlock, 0, c->stream>>>(n, tp<T>(x), tp<T>(y), P, S,
                                                                d_ticket, d_out);
}

// the canonical dot (sq: its square root, in the vectors' precision) into
// *d_out with the tile partials and lane sums in `partials`
// ((ntiles + 2 * kRedGroups) values) and zeroed counters d_ticket
static void launch_dot_into(Ctx *c, mgx_dtype dt, int64_t n, const void *x, const void *y,
                            void *partials, double *d_out, uint32_t *d_ticket,
                            bool sq = false) {
  if (n <= 0) {
    launch_fill(c, MGX_F64, d_out, 1, 0.0);
    return;
  }
  if (dt == MGX_F64) {
    if (sq)
      dot_kernel<double, true>(c, n, x, y, partials, d_out, d_ticket);
    else
      dot_kernel<double, false>(c, n, x, y, partials, d_out, d_ticket);
  } else {
    if (sq)
      dot_kernel<float, true>(c, n, x, y, partials, d_out, d_ticket);
    else
      dot_kernel<float, false>(c, n, x, y, partials, d_out, d_ticket);
  }
  MGX_HIP(hipGetLastError());
}

void launch_dot(Ctx *c, mgx_dtype dt, int64_t n, const void *x, const void *y,
                double *d_out, uint32_t *d_ticket, bool sq) {
  if (n > 0) {
    // tile partials, then the 256 lane sums
    const size_t es = dt == MGX_F64 ? 8 : 4;
    ensure_partials(c, (size_t)(ntiles_of(n) + 2 * kRedGroups) * es);
  }
  launch_dot_into(c, dt, n, x, y, c->d_partials, d_out, d_ticket, sq);
}

void launch_mnrm2sq(Ctx *c, mgx_dtype dt, int64_t n, int m, const void *X, int64_t ld,
                    void *partials, uint32_t *ticks, double *d_out) {
  if (n <= 0) {
    launch_fill(c, MGX_F64, d_out, m, 0.0);
    return;
  }
  const int64_t nt = ntiles_of(n), pst = nt + 2 * kRedGroups;
  const bool vec = aligned16(X) && (ld % 2 == 0);
  const dim3 g((unsigned)nt, (unsigned)m);
  if (dt == MGX_F64) {
    if (vec)
      k_mdot<double, true><<<g, kBlock, 0, c->stream>>>(n, tp<double>(X), ld,
                                                        tp<double>(partials), pst, ticks, d_out);
    else
      k_mdot<double, false><<<g, kBlock, 0, c->stream>>>(n, tp<double>(X), ld,
                                                         tp<double>(partials), pst, ticks, d_out);
  } else {
    if (vec && ld % 4 == 0)
      k_mdot<float, true><<<g, kBlock, 0, c->stream>>>(n, tp<float>(X), ld, tp<float>(partials),
                                                       pst, ticks, d_out);
    else
      k_mdot<float, false><<<g, kBlock, 0, c->stream>>>(n, tp<float>(X), ld,
                                                        tp<float>(partials), pst, ticks, d_out);
  }
  MGX_HIP(hipGetLastError());
}

// ---- allocation-free dot ------------------------------------------------
// For callers that cannot allocate device memory (code under HIP graph
// capture, serving loops that forbid hipMalloc): the dot's arrival counters,
// tile partials and lane sums live in one file-scope device array.  Every
// call on a device uses that device's copy and overwrites what the last
// call left there, so calls that overlap on one device (two streams, two
// threads) race.
constexpr int64_t kNoAllocMaxN = int64_t(1) << 26;  // 67,108,864 elements
constexpr size_t kNoAllocTickBytes = 2048;          // >= kTickWords counters, 256-aligned
constexpr size_t kNoAllocBytes =
    kNoAllocTickBytes + (size_t)(kNoAllocMaxN / kTile + 2 * kRedGroups) * sizeof(double);
static_assert(kTickWords * sizeof(uint32_t) <= kNoAllocTickBytes, "counter area");
static_assert(kNoAllocMaxN % kTile == 0, "whole tiles");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocDevices = 64;
static void *g_scratch_ptr[kNoAllocDevices] = {nullptr};
static std::mutex g_scratch_mu;

// zeroes n counters (a kernel rather than a memset node, so the call can be
// captured into a graph)
namespace {
__global__ void k_zero_u32(uint32_t *p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}
}  // namespace

// Looks up this device's address of `scratch` (the array's name is only a
// host-side stand-in).  Called by mgx_init with `device` current.
void noalloc_scratch_init(int device) {
  MGX_REQUIRE(device >= 0 && device < kNoAllocDevices, "static scratch: device %d out of range",
              device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void *ptr = nullptr;
  MGX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  MGX_REQUIRE(ptr != nullptr, "static scratch: no device address on device %d", device);
  g_scratch_ptr[device] = ptr;
}

static void *noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return device >= 0 && device < kNoAllocDevices ? g_scratch_ptr[device] : nullptr;
}

double dot_host(Ctx *c, mgx_dtype dt, int64_t n, const void *x,
                const void *y) {
  double *slot = &c->d_state->rr;
  launch_dot(c, dt, n, x, y, slot, c->d_tick);
  MGX_HIP(hipMemcpyAsync(c->h_pinned, slot, sizeof(double),
                         hipMemcpyDeviceToHost, c->stream));
  MGX_HIP(hipStreamSynchronize(c->stream));
  return c->h_pinned[0];
}

}  // namespace mgx

using namespace mgx;

extern "C" int mgx_dot_device_noalloc(mgx_vec *x, mgx_vec *y, double *d_out) {
  return guard([&] {
    MGX_REQUIRE(x && y && d_out, "mgx_dot_device_noalloc: NULL argument");
    MGX_REQUIRE(x->ctx == y->ctx, "mgx_dot_device_noalloc: vectors of different contexts");
    MGX_REQUIRE(x->dtype == y->dtype, "mgx_dot_device_noalloc: dtype mismatch");
    MGX_REQUIRE(x->n == y->n, "mgx_dot_device_noalloc: size mismatch (%lld vs %lld)",
                (long long)x->n, (long long)y->n);
    MGX_REQUIRE(x->n <= kNoAllocMaxN,
                "mgx_dot_device_noalloc: %lld elements exceed the static scratch's limit of "
                "%lld",
                (long long)x->n, (long long)kNoAllocMaxN);
    Ctx *c = x->ctx;
    activate(c);
    char *base = static_cast<char *>(noalloc_scratch(c->device));
    MGX_REQUIRE(base, "mgx_dot_device_noalloc: static scratch not set up on device %d",
                c->device);
    uint32_t *ticks = reinterpret_cast<uint32_t *>(base);
    if (x->n > 0) {
      k_zero_u32<<<1, 256, 0, c->stream>>>(ticks, kTickWords);
      MGX_HIP(hipGetLastError());
    }
    launch_dot_into(c, x->dtype, x->n, x->d, y->d, base + kNoAllocTickBytes, d_out, ticks);
    return MGX_OK;
  });
}
