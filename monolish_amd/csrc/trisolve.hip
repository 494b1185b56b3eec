// trisolve.hip -- level-scheduled triangular sweeps: the SOR sweeps
// (src/equation/kernel/sor_kernel.hpp:5-73) and the IC / ILU triangular
// solves (csrsv2 in kernel/ic_kernel.hpp, ilu_kernel.hpp).
//
// Every such sweep has one form: out_i = fin(init_i - sum_d coef_d *
// out_{col_d}), the sum over the row's dependency entries in the row's
// storage order.  A plan, built once per matrix structure, lists the rows
// in level order (level of a row = 1 + the largest level of the rows it
// reads) with their dependency entries packed behind each other, so a row
// needs one 16-byte record and contiguous entry loads; the coefficients and
// the per-row scale are gathered into the same order once per solve.  Rows
// of a level are independent and each row does the serial loop's
// arithmetic in its order, so the result is bit-identical to the serial
// sweep.  Within a row the entries are loaded eight at a time before their
// products are subtracted in order, so a row costs a few memory round
// trips, not two per entry.
#include <algorithm>
#include <vector>

#include "mgx_device.hpp"
#include "mgx_sweep.hpp"

namespace mgx {

namespace {

constexpr int kTriChunk = 8;
// level-order positions whose values the single-workgroup sweeps keep in
// LDS (a read of an older position goes to out)
constexpr int kTriRing = 4096;

// FIN 0: out = acc * sc (SOR forward); 1: SOR backward, acc in double from
// 0, out = (T)((double)out - acc * (double)sc); 2: out = acc / sc; 3: out = acc
template <class T, int FIN>
__device__ __forceinline__ void tri_row(const int4 rc, const int32_t *__restrict__ dcol,
                                        const T *__restrict__ dcoef, const T scv,
                                        const T *__restrict__ in, T *out) {
  const int32_t i = rc.x, d0 = rc.y, d1 = rc.z;
  using Acc = typename std::conditional<FIN == 1, double, T>::type;
  Acc acc = (Acc)0;
  if (FIN != 1) acc = (Acc)in[i];
  for (int32_t d = d0; d < d1; d += kTriChunk) {
    int32_t c[kTriChunk];
    T a[kTriChunk], zv[kTriChunk];
#pragma unroll
    for (int u = 0; u < kTriChunk; u++) {
      c[u] = d + u < d1 ? dcol[d + u] : i;
      a[u] = d + u < d1 ? dcoef[d + u] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < kTriChunk; u++) zv[u] = d + u < d1 ? out[c[u]] : (T)0;
#pragma unroll
    for (int u = 0; u < kTriChunk; u++) {
      if (d + u < d1) {
        const T p = a[u] * zv[u];
        if (FIN == 1)
          acc = acc + (double)p;
        else
          acc = acc - p;
      }
    }
  }
  if (FIN == 0)
    out[i] = acc * scv;
  else if (FIN == 1)
    out[i] = (T)((double)out[i] - acc * (double)scv);
  else if (FIN == 2)
    out[i] = acc / scv;
  else
    out[i] = acc;
}

template <class T, int FIN>
__global__ __launch_bounds__(kSweepWgThreads) void k_tri_wg(
    const int32_t *__restrict__ lptr, int32_t nlev, const int4 *__restrict__ rec,
    const int32_t *__restrict__ dcol, const T *__restrict__ dcoef, const T *__restrict__ sc,
    const T *__restrict__ in, T *out) {
  for (int32_t L = 0; L < nlev; L++) {
    const int32_t k1 = lptr[L + 1];
    for (int32_t k = lptr[L] + (int32_t)threadIdx.x; k < k1; k += kSweepWgThreads)
      tri_row<T, FIN>(rec[k], dcol, dcoef, sc[k], in, out);
    __syncthreads();  // the level's rows are written before the next reads them
  }
}

// The pipelined form (levels at most kSweepWgThreads rows wide, one row per
// thread and level): while level L runs, the dependency entries of level
// L+1 and the row records of level L+2 are already in flight, and the
// values of the last kTriRing level-order positions live in LDS (dpos >= 0
// names the slot; older values are read from out), so a level costs LDS
// reads, a store and the barrier.  Within a wave that has rows the prefetch
// loads are unconditional (clamped to entry 0 for missing rows), waves
// without rows load nothing, and the level loop is unrolled six times so
// that no register of a load in flight is ever copied: together these keep
// the compiler's wait counts exact, which is what lets the loads overlap.
constexpr int kTriPre = 4;  // dependency entries held in registers per row

template <class T> struct TriHead {
  int32_t i, d0, d1;  // d1 < 0: no row
  T s, init;
};
template <class T> struct TriDeps {
  int32_t c[kTriPre], p[kTriPre];
  T a[kTriPre];
};

template <class T, int FIN>
__device__ __forceinline__ void tri_head(TriHead<T> &h, bool valid, int32_t k,
                                         const int4 *__restrict__ rec, const T *__restrict__ sc,
                                         const T *__restrict__ in, const T *out) {
  const int32_t kk = valid ? k : 0;
  const int4 r = rec[kk];
  h.i = r.x;
  h.d0 = r.y;
  h.d1 = valid ? r.z : -1;
  h.s = sc[kk];
  h.init = FIN == 1 ? out[r.x] : in[r.x];
}

template <class T>
__device__ __forceinline__ void tri_deps(const TriHead<T> &h, TriDeps<T> &d,
                                         const int32_t *__restrict__ dcol,
                                         const int32_t *__restrict__ dpos,
                                         const T *__restrict__ dcoef) {
#pragma unroll
  for (int u = 0; u < kTriPre; u++) {
    const int32_t e = h.d0 + u < h.d1 ? h.d0 + u : 0;
    d.c[u] = dcol[e];
    d.p[u] = dpos[e];
    d.a[u] = dcoef[e];
  }
}

// the values read by row h: the first kTriPre from the prefetched entries,
// then the product sum over all of them in the row's order
template <class T, int FIN>
__device__ __forceinline__ void tri_finish(const TriHead<T> &h, const TriDeps<T> &d, int32_t k,
                                           const T *zv, T *ring,
                                           const int32_t *__restrict__ dcol,
                                           const int32_t *__restrict__ dpos,
                                           const T *__restrict__ dcoef, T *out) {
  using Acc = typename std::conditional<FIN == 1, double, T>::type;
  Acc acc = FIN == 1 ? (Acc)0 : (Acc)h.init;
#pragma unroll
  for (int u = 0; u < kTriPre; u++) {
    if (h.d0 + u < h.d1) {
      const T p = d.a[u] * zv[u];
      if (FIN == 1)
        acc = acc + (double)p;
      else
        acc = acc - p;
    }
  }
  for (int32_t e = h.d0 + kTriPre; e < h.d1; e += kTriChunk) {
    int32_t c[kTriChunk], q[kTriChunk];
    T a[kTriChunk], z[kTriChunk];
#pragma unroll
    for (int u = 0; u < kTriChunk; u++) {
      const bool v = e + u < h.d1;
      c[u] = v ? dcol[e + u] : 0;
      q[u] = v ? dpos[e + u] : 0;
      a[u] = v ? dcoef[e + u] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < kTriChunk; u++)
      z[u] = e + u < h.d1 ? (q[u] >= 0 ? ring[q[u] & (kTriRing - 1)] : out[c[u]]) : (T)0;
#pragma unroll
    for (int u = 0; u < kTriChunk; u++) {
      if (e + u < h.d1) {
        const T p = a[u] * z[u];
        if (FIN == 1)
          acc = acc + (double)p;
        else
          acc = acc - p;
      }
    }
  }
  T v;
  if (FIN == 0)
    v = acc * h.s;
  else if (FIN == 1)
    v = (T)((double)h.init - acc * (double)h.s);
  else if (FIN == 2)
    v = acc / h.s;
  else
    v = acc;
  out[h.i] = v;
  ring[k & (kTriRing - 1)] = v;
}

template <class T, int FIN>
__global__ __launch_bounds__(kSweepWgThreads) void k_tri_pipe(
    const int32_t *__restrict__ lptr, int32_t nlev, const int4 *__restrict__ rec,
    const int32_t *__restrict__ dcol, const int32_t *__restrict__ dpos,
    const T *__restrict__ dcoef, const T *__restrict__ sc, const T *__restrict__ in, T *out) {
  __shared__ T ring[kTriRing];
  const int32_t t = (int32_t)threadIdx.x;
  // level L covers positions [lptr[L], lptr[L+1]); kb[] holds the bounds of
  // levels L, L+1 and L+2 as the loop advances
  int32_t k0 = lptr[0], k1 = nlev > 0 ? lptr[1] : 0, k2 = nlev > 1 ? lptr[2] : k1;
  TriHead<T> h0, h1, h2;
  TriDeps<T> e0, e1;
  tri_head<T, FIN>(h0, k0 + t < k1, k0 + t, rec, sc, in, out);
  tri_deps(h0, e0, dcol, dpos, dcoef);
  tri_head<T, FIN>(h1, nlev > 1 && k1 + t < k2, k1 + t, rec, sc, in, out);
  h2.d1 = -1;
  // one level: cur (its entries cd) runs, the entries of nxt and the record
  // of the level after it (into nn) are loaded
  const int32_t w0 = t & ~63;  // first thread of this wave: waves without rows load nothing
  auto step = [&](int32_t L, TriHead<T> &cur, TriDeps<T> &cd, TriHead<T> &nxt, TriDeps<T> &nd,
                  TriHead<T> &nn) {
    const int32_t k3 = L + 3 <= nlev ? lptr[L + 3] : k2;
    T zv[kTriPre];
#pragma unroll
    for (int u = 0; u < kTriPre; u++)
      zv[u] = cur.d0 + u < cur.d1
                  ? (cd.p[u] >= 0 ? ring[cd.p[u] & (kTriRing - 1)] : out[cd.c[u]])
                  : (T)0;
    // the values (and every load of the step before) have arrived before the
    // prefetch loads go out, so nothing below waits for those
    __builtin_amdgcn_s_waitcnt(0);
    if (w0 < k2 - k1) tri_deps(nxt, nd, dcol, dpos, dcoef);
    if (L + 2 < nlev && w0 < k3 - k2)
      tri_head<T, FIN>(nn, k2 + t < k3, k2 + t, rec, sc, in, out);
    else
      nn.d1 = -1;
    if (cur.d1 >= 0) tri_finish<T, FIN>(cur, cd, k0 + t, zv, ring, dcol, dpos, dcoef, out);
    __syncthreads();  // the level's values (LDS and out) are written before the next reads them
    k0 = k1;
    k1 = k2;
    k2 = k3;
  };
  for (int32_t L = 0; L < nlev; L += 6) {
    step(L, h0, e0, h1, e1, h2);
    if (L + 1 >= nlev) break;
    step(L + 1, h1, e1, h2, e0, h0);
    if (L + 2 >= nlev) break;
    step(L + 2, h2, e0, h0, e1, h1);
    if (L + 3 >= nlev) break;
    step(L + 3, h0, e1, h1, e0, h2);
    if (L + 4 >= nlev) break;
    step(L + 4, h1, e0, h2, e1, h0);
    if (L + 5 >= nlev) break;
    step(L + 5, h2, e1, h0, e0, h1);
  }
}

// The staged form: levels are grouped into blocks of at most kStPos rows
// and kStDeps dependency entries (whole levels, contiguous in the plan).
// While the rows of block b run out of LDS, the records of block b+1 are
// loaded into registers by the whole workgroup in one coalesced pass and
// written to the other LDS buffer afterwards, so a level costs only LDS
// traffic, its stores and the barrier; global memory latency is paid once
// per block.  The initial values (in, or out for the SOR backward sweep)
// are gathered into level order by k_tri_gather before the sweep.
constexpr int kStPos = 1024, kStDeps = 1536;

template <class T>
__global__ __launch_bounds__(kBlock) void k_tri_gather(int64_t n, const int4 *__restrict__ rec,
                                                       const T *__restrict__ src, T *dst) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k < n) dst[k] = src[rec[k].x];
}

template <class T, int FIN>
__global__ __launch_bounds__(kSweepWgThreads) void k_tri_staged(
    const int4 *__restrict__ blk, int32_t nblk, const int32_t *__restrict__ lptr,
    const int4 *__restrict__ rec, const T *__restrict__ sc, const T *__restrict__ ginit,
    const int32_t *__restrict__ dpos, const int32_t *__restrict__ dcol,
    const T *__restrict__ dcoef, T *out) {
  __shared__ int4 s_rec[2][kStPos];
  __shared__ T s_sc[2][kStPos], s_init[2][kStPos];
  __shared__ int32_t s_dpos[2][kStDeps], s_dcol[2][kStDeps];
  __shared__ T s_coef[2][kStDeps];
  __shared__ int32_t s_lptr[2][kStPos];
  __shared__ T ring[kTriRing];
  constexpr int kDU = (kStDeps + kSweepWgThreads - 1) / kSweepWgThreads;
  const int32_t t = (int32_t)threadIdx.x;
  // registers of the block being loaded (block descriptors: x = first
  // level, y = first position, z = first dependency entry)
  int4 g_rec;
  T g_sc, g_init, g_coef[kDU];
  int32_t g_dpos[kDU], g_dcol[kDU], g_lptr;
  auto load = [&](const int4 B0, const int4 B1) {
    const int32_t np = B1.y - B0.y, nd = B1.z - B0.z, nl = B1.x - B0.x;
    const int32_t kp = B0.y + (t < np ? t : 0);
    g_rec = rec[kp];
    g_sc = sc[kp];
    g_init = ginit[kp];
#pragma unroll
    for (int u = 0; u < kDU; u++) {
      const int32_t dd = t + u * kSweepWgThreads;
      const int32_t idx = dd < nd ? B0.z + dd : 0;
      g_dpos[u] = dpos[idx];
      g_dcol[u] = dcol[idx];
      g_coef[u] = dcoef[idx];
    }
    g_lptr = lptr[B0.x + (t < nl ? t : 0)];
  };
  auto store = [&](int buf, const int4 B0, const int4 B1) {
    const int32_t np = B1.y - B0.y, nd = B1.z - B0.z, nl = B1.x - B0.x;
    if (t < np) {
      s_rec[buf][t] = g_rec;
      s_sc[buf][t] = g_sc;
      s_init[buf][t] = g_init;
    }
#pragma unroll
    for (int u = 0; u < kDU; u++) {
      const int32_t dd = t + u * kSweepWgThreads;
      if (dd < nd && dd < kStDeps) {
        s_dpos[buf][dd] = g_dpos[u];
        s_dcol[buf][dd] = g_dcol[u];
        s_coef[buf][dd] = g_coef[u];
      }
    }
    if (t < nl) s_lptr[buf][t] = g_lptr;
  };
  int4 B0 = blk[0], B1 = blk[1];
  load(B0, B1);
  store(0, B0, B1);
  __syncthreads();
  for (int32_t b = 0; b < nblk; b++) {
    const int buf = b & 1;
    const int4 N0 = B1, N1 = b + 2 <= nblk ? blk[b + 2] : B1;
    if (b + 1 < nblk) load(N0, N1);  // in flight while this block runs
    const int32_t nl = B1.x - B0.x;
    for (int32_t l = 0; l < nl; l++) {
      const int32_t ka = s_lptr[buf][l], kb = l + 1 < nl ? s_lptr[buf][l + 1] : B1.y;
      if (ka + t < kb) {
        const int32_t kk = ka + t - B0.y;
        const int4 rc = s_rec[buf][kk];
        const int32_t d0 = rc.y - B0.z, d1 = rc.z - B0.z;
        using Acc = typename std::conditional<FIN == 1, double, T>::type;
        const T init = s_init[buf][kk];
        Acc acc = FIN == 1 ? (Acc)0 : (Acc)init;
        for (int32_t d = d0; d < d1; d++) {
          const int32_t q = s_dpos[buf][d];
          const T z = q >= 0 ? ring[q & (kTriRing - 1)] : out[s_dcol[buf][d]];
          const T p = s_coef[buf][d] * z;
          if (FIN == 1)
            acc = acc + (double)p;
          else
            acc = acc - p;
        }
        const T s = s_sc[buf][kk];
        T v;
        if (FIN == 0)
          v = acc * s;
        else if (FIN == 1)
          v = (T)((double)init - acc * (double)s);
        else if (FIN == 2)
          v = acc / s;
        else
          v = acc;
        out[rc.x] = v;
        ring[(ka + t) & (kTriRing - 1)] = v;
      }
      __syncthreads();  // the level's values are written before the next reads them
    }
    if (b + 1 < nblk) store(buf ^ 1, N0, N1);
    __syncthreads();
    B0 = N0;
    B1 = N1;
  }
}

// The dependency-driven form, for plans too large or too wide for one
// workgroup: persistent waves take 64 consecutive level-order positions at a
// time from a ticket counter and each lane computes its row as soon as the
// rows it reads are written, with no barrier between levels.  Readiness
// travels with the value: before the sweep every out_i holds kEmpty, a
// signalling NaN that no arithmetic produces, and a row publishes its value
// with one device-coherent atomic exchange, so a reader that finds
// something else has the final value.  A lane loads its entries eight at a
// time (plain loads) and consumes them strictly in the row's order; at the
// first one still kEmpty it polls that one entry with a device-coherent
// atomic read (a plain load may keep seeing a cached copy) until it is
// written, then goes on, so the arithmetic is the serial sweep's.  Rows only read rows of earlier
// levels, which sit at earlier positions, so their tickets were taken by
// waves that are already running: the sweep cannot deadlock.  A lane that
// has consumed none of its row's entries for kFlowSpinTicks of the 100 MHz
// wall clock gives up: it publishes NaN and raises the plan's stall flag
// (ticket[1]), so the kernel always ends and the solver returns
// MGX_ERR_TIMEOUT.  At most 1024 waves x 64 rows are in flight, so a valid
// plan's longest wait is 65536 dependency hops (measured 4-7 us each), far
// inside the 2 s limit.
constexpr int kFlowThreads = 256;
constexpr int kFlowBlocks = 256;
constexpr uint64_t kFlowSpinTicks = 200000000;  // 2 s without progress
// plans whose levels hold kFlowMinWidth rows on average take the
// dependency-driven form when they have at least kFlowMinRows rows or a
// level wider than one workgroup's block; deep plans (a long chain beside
// wide levels) stay in one workgroup, which sweeps any width level by level
constexpr int64_t kFlowMinRows = 1 << 17, kFlowMinWidth = 256;

template <class T> struct FlowBits;
template <> struct FlowBits<double> {
  using U = uint64_t;
  static constexpr U kEmpty = 0x7ff0deadbeef0001ull, kNan = 0x7ff8000000000000ull;
};
template <> struct FlowBits<float> {
  using U = uint32_t;
  static constexpr U kEmpty = 0x7f80dea1u, kNan = 0x7fc00000u;
};

// ginit[k] = the initial value of row rec[k].x (out for the SOR backward
// sweep, in otherwise), then out of that row = kEmpty; in may alias out
template <class T, int FIN>
__global__ __launch_bounds__(kBlock) void k_flow_init(int64_t n, const int4 *__restrict__ rec,
                                                      const T *in, T *out, T *ginit,
                                                      uint32_t *ticket) {
  using U = typename FlowBits<T>::U;
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k == 0) *ticket = 0u;
  if (k < n) {
    const int32_t i = rec[k].x;
    ginit[k] = FIN == 1 ? out[i] : in[i];
    reinterpret_cast<U *>(out)[i] = FlowBits<T>::kEmpty;
  }
}

template <class T, int FIN>
__global__ __launch_bounds__(kFlowThreads) void k_tri_flow(
    int32_t n, const int4 *__restrict__ rec, const int32_t *__restrict__ dcol,
    const T *__restrict__ dcoef, const T *__restrict__ sc, const T *__restrict__ ginit, T *out,
    uint32_t *ticket) {
  using U = typename FlowBits<T>::U;
  using Acc = typename std::conditional<FIN == 1, double, T>::type;
  U *ob = reinterpret_cast<U *>(out);
  const int lane = (int)(threadIdx.x & 63);
  const uint32_t nchunk = (uint32_t)((n + 63) / 64);
  for (;;) {
    uint32_t t = 0;
    if (lane == 0)
      t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    t = (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
    if (t >= nchunk) break;
    const int32_t k = (int32_t)(t * 64u) + lane;
    const bool has = k < n;
    const int4 rc = has ? rec[k] : make_int4(0, 0, 0, 0);
    int32_t d = rc.y;
    const int32_t d1 = has ? rc.z : rc.y;
    const T init = has ? ginit[k] : (T)0;
    Acc acc = FIN == 1 ? (Acc)0 : (Acc)init;
    uint64_t t0 = wall_clock64();  // the lane's last progress
    int32_t wc = -1;  // >= 0: column of entry d, which was still kEmpty
    T wa = (T)0;
    // a lane publishes its row as soon as its last entry is consumed: a row
    // later in the same 64 may read it
    bool pending = has;
    while (__any(pending)) {
      if (pending && d < d1) {
        if (wc >= 0) {
          const U z = __hip_atomic_fetch_or(ob + wc, (U)0, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
          if (z != FlowBits<T>::kEmpty) {
            const T p = wa * __builtin_bit_cast(T, z);
            if (FIN == 1)
              acc = acc + (double)p;
            else
              acc = acc - p;
            d++;
            wc = -1;
            t0 = wall_clock64();
          }
        } else {
          int32_t c[kTriChunk];
          T a[kTriChunk];
          U zb[kTriChunk];
#pragma unroll
          for (int u = 0; u < kTriChunk; u++) {
            const int32_t e = d + u < d1 ? d + u : d;
            c[u] = dcol[e];
            a[u] = dcoef[e];
          }
#pragma unroll
          // plain loads: out held kEmpty when this kernel started (k_flow_init,
          // an earlier kernel of the stream) and changes once, to the final
          // value, so a cached copy is either that value or a stale kEmpty,
          // which the device-coherent poll below resolves
          for (int u = 0; u < kTriChunk; u++) zb[u] = d + u < d1 ? ob[c[u]] : (U)0;
          bool stop = false;
          int32_t taken = 0;
#pragma unroll
          for (int u = 0; u < kTriChunk; u++) {
            if (!stop && d + u < d1) {
              if (zb[u] == FlowBits<T>::kEmpty) {
                stop = true;
                wc = c[u];
                wa = a[u];
              } else {
                const T p = a[u] * __builtin_bit_cast(T, zb[u]);
                if (FIN == 1)
                  acc = acc + (double)p;
                else
                  acc = acc - p;
                taken++;
              }
            }
          }
          d += taken;
          if (taken > 0) t0 = wall_clock64();
        }
        if (d < d1 && wc >= 0 && wall_clock64() - t0 > kFlowSpinTicks) {
          acc = (Acc)__builtin_bit_cast(T, FlowBits<T>::kNan);
          d = d1;
          __hip_atomic_store(ticket + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      if (pending && d >= d1) {
        const T s = sc[k];
        T v;
        if (FIN == 0)
          v = acc * s;
        else if (FIN == 1)
          v = (T)((double)init - acc * (double)s);
        else if (FIN == 2)
          v = acc / s;
        else
          v = acc;
        U vb = __builtin_bit_cast(U, v);
        if (vb == FlowBits<T>::kEmpty) vb = FlowBits<T>::kNan;
        __hip_atomic_exchange(ob + rc.x, vb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        pending = false;
      }
      if (__any(pending && wc >= 0)) __builtin_amdgcn_s_sleep(2);
    }
  }
}

// coef[d] = vals[dent[d]]
template <class T>
__global__ __launch_bounds__(kBlock) void k_tri_pack(int64_t nd, const int32_t *__restrict__ dent,
                                                     const T *__restrict__ vals, T *coef) {
  const int64_t d = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (d < nd) coef[d] = vals[dent[d]];
}

// sc[k] = vec[pos ? pos[row_k] : row_k]
template <class T>
__global__ __launch_bounds__(kBlock) void k_tri_scale(int64_t n, const int4 *__restrict__ rec,
                                                      const T *__restrict__ vec,
                                                      const int32_t *__restrict__ pos, T *sc) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k < n) {
    const int32_t i = rec[k].x;
    sc[k] = vec[pos ? pos[i] : i];
  }
}

template <class T, int FIN>
void tri_run_t(Ctx *c, const TriPlan &P, const T *coef, const T *sc, const T *in, T *out,
               T *ginit) {
  const SorLevels &L = P.lev;
  if (L.nlev == 0) return;
  if (P.flow) {
    const int64_t n = L.ptr[L.nlev];
    k_flow_init<T, FIN><<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, c->stream>>>(
        n, P.d_rec, in, out, ginit, P.d_ticket);
    const int64_t want = (n + 63) / 64 / (kFlowThreads / 64) + 1;
    k_tri_flow<T, FIN><<<(unsigned)std::min<int64_t>(want, kFlowBlocks), kFlowThreads, 0,
                         c->stream>>>((int32_t)n, P.d_rec, P.d_dcol, coef, sc, ginit, out,
                                      P.d_ticket);
  } else if (P.nblk > 0) {
    const int64_t n = L.ptr[L.nlev];
    k_tri_gather<T><<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, c->stream>>>(
        n, P.d_rec, FIN == 1 ? out : in, ginit);
    k_tri_staged<T, FIN><<<1, kSweepWgThreads, 0, c->stream>>>(
        P.d_blk, P.nblk, L.d_ptr, P.d_rec, sc, ginit, P.d_dpos, P.d_dcol, coef, out);
  } else if (L.max_width <= kSweepWgThreads) {
    k_tri_pipe<T, FIN><<<1, kSweepWgThreads, 0, c->stream>>>(
        L.d_ptr, L.nlev, P.d_rec, P.d_dcol, P.d_dpos, coef, sc, in, out);
  } else {
    k_tri_wg<T, FIN><<<1, kSweepWgThreads, 0, c->stream>>>(L.d_ptr, L.nlev, P.d_rec, P.d_dcol,
                                                           coef, sc, in, out);
  }
  MGX_HIP(hipGetLastError());
}

template <class T> void tri_run_fin(Ctx *c, const TriOp &op, const void *in, void *out) {
  const T *coef = static_cast<const T *>(op.coef), *sc = static_cast<const T *>(op.sc);
  const T *x = static_cast<const T *>(in);
  T *y = static_cast<T *>(out), *g = static_cast<T *>(op.ginit);
  switch (op.fin) {
    case kTriMul: tri_run_t<T, 0>(c, *op.plan, coef, sc, x, y, g); break;
    case kTriSorBack: tri_run_t<T, 1>(c, *op.plan, coef, sc, x, y, g); break;
    case kTriDiv: tri_run_t<T, 2>(c, *op.plan, coef, sc, x, y, g); break;
    default: tri_run_t<T, 3>(c, *op.plan, coef, sc, x, y, g); break;
  }
}

bool dep_of(int filter, int64_t i, int32_t c) { return filter == 2 ? c > i : c < i; }

}  // namespace

void build_tri_plan(const std::vector<int32_t> &ptr, const std::vector<int32_t> &col,
                    const std::vector<int32_t> *ent, int64_t n, int filter, TriPlan &P) {
  // dependency entries of each row, in the row's order (filter 1 stops at
  // the first entry that is not left of the diagonal)
  std::vector<int32_t> cnt(n, 0), lev(n, 0);
  int32_t nlev = 0;
  auto for_deps = [&](int64_t i, auto &&f) {
    for (int32_t j = ptr[i]; j < ptr[i + 1]; j++) {
      if (dep_of(filter, i, col[j]))
        f(j);
      else if (filter == 1)
        break;
    }
  };
  if (filter == 2) {
    for (int64_t i = n - 1; i >= 0; i--) {
      int32_t l = 0;
      for_deps(i, [&](int32_t j) { l = std::max(l, lev[col[j]] + 1), cnt[i]++; });
      lev[i] = l;
      nlev = std::max(nlev, l + 1);
    }
  } else {
    for (int64_t i = 0; i < n; i++) {
      int32_t l = 0;
      for_deps(i, [&](int32_t j) { l = std::max(l, lev[col[j]] + 1), cnt[i]++; });
      lev[i] = l;
      nlev = std::max(nlev, l + 1);
    }
  }
  SorLevels &L = P.lev;
  L.ptr.assign(nlev + 1, 0);
  for (int64_t i = 0; i < n; i++) L.ptr[lev[i] + 1]++;
  for (int32_t l = 0; l < nlev; l++) L.ptr[l + 1] += L.ptr[l];
  std::vector<int32_t> pos(L.ptr.begin(), L.ptr.end() - 1), rows(n);
  for (int64_t i = 0; i < n; i++) rows[pos[lev[i]]++] = (int32_t)i;
  L.nlev = nlev;
  L.max_width = 0;
  for (int32_t l = 0; l < nlev; l++) L.max_width = std::max(L.max_width, L.ptr[l + 1] - L.ptr[l]);
  int64_t nd = 0;
  for (int64_t i = 0; i < n; i++) nd += cnt[i];
  std::vector<int4> rec(n);
  std::vector<int32_t> dcol(std::max<int64_t>(nd, 1)), dent(std::max<int64_t>(nd, 1)),
      dpos(std::max<int64_t>(nd, 1)), where(n);
  for (int64_t k = 0; k < n; k++) where[rows[k]] = (int32_t)k;
  int64_t d = 0;
  for (int64_t k = 0; k < n; k++) {
    const int32_t i = rows[k];
    // a value read by this row is still in the LDS ring when no position of
    // this row's level or an earlier one has overwritten its slot
    const int64_t lo = (int64_t)L.ptr[lev[i] + 1] - kTriRing;
    rec[k] = make_int4(i, (int32_t)d, (int32_t)(d + cnt[i]), 0);
    for_deps(i, [&](int32_t j) {
      dcol[d] = col[j];
      dent[d] = ent ? (*ent)[j] : j;
      dpos[d] = where[col[j]] >= lo ? where[col[j]] : -1;
      d++;
    });
  }
  P.ndeps = nd;
  L.d_rows = (int32_t *)dev_alloc((size_t)std::max<int64_t>(n, 1) * 4);
  L.d_ptr = (int32_t *)dev_alloc((size_t)(nlev + 1) * 4);
  P.d_rec = (int4 *)dev_alloc((size_t)std::max<int64_t>(n, 1) * sizeof(int4));
  P.d_dcol = (int32_t *)dev_alloc(dcol.size() * 4);
  P.d_dent = (int32_t *)dev_alloc(dent.size() * 4);
  P.d_dpos = (int32_t *)dev_alloc(dpos.size() * 4);
  MGX_HIP(hipMemcpy(P.d_dpos, dpos.data(), dpos.size() * 4, hipMemcpyHostToDevice));
  // large plans with wide levels sweep dependency-driven across the chip;
  // the others in one workgroup, staged in blocks of whole levels when no
  // level alone exceeds a block
  P.flow = (L.max_width > kSweepWgRows || n >= kFlowMinRows) &&
           n >= kFlowMinWidth * (int64_t)nlev;
  if (P.flow) {
    P.d_ticket = (uint32_t *)dev_alloc(256);
    MGX_HIP(hipMemset(P.d_ticket, 0, 256));  // [0] ticket, [1] stall flag
  }
  std::vector<int4> blk;
  bool fits = !P.flow;
  for (int32_t l = 0; l < nlev && fits; l++) {
    const int32_t p0 = L.ptr[l], p1 = L.ptr[l + 1];
    const int32_t e0 = rec[p0].y, e1 = rec[p1 - 1].z;
    if (p1 - p0 > kStPos || e1 - e0 > kStDeps) fits = false;
    if (blk.empty() || p1 - blk.back().y > kStPos || e1 - blk.back().z > kStDeps)
      blk.push_back(make_int4(l, p0, e0, 0));
  }
  P.nblk = 0;
  if (fits && nlev > 0) {
    blk.push_back(make_int4(nlev, (int32_t)n, (int32_t)nd, 0));
    P.nblk = (int32_t)blk.size() - 1;
    P.d_blk = (int4 *)dev_alloc(blk.size() * sizeof(int4));
    MGX_HIP(hipMemcpy(P.d_blk, blk.data(), blk.size() * sizeof(int4), hipMemcpyHostToDevice));
  }
  if (n > 0) {
    MGX_HIP(hipMemcpy(L.d_rows, rows.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    MGX_HIP(hipMemcpy(P.d_rec, rec.data(), (size_t)n * sizeof(int4), hipMemcpyHostToDevice));
  }
  MGX_HIP(hipMemcpy(L.d_ptr, L.ptr.data(), (size_t)(nlev + 1) * 4, hipMemcpyHostToDevice));
  MGX_HIP(hipMemcpy(P.d_dcol, dcol.data(), dcol.size() * 4, hipMemcpyHostToDevice));
  MGX_HIP(hipMemcpy(P.d_dent, dent.data(), dent.size() * 4, hipMemcpyHostToDevice));
  P.built = true;
}

const TriPlan &tri_plan(Csr *A, int which) {
  TriPlan &P = A->tri[which];
  if (!P.built) {
    MGX_REQUIRE(which <= kTriUpper, "internal: the L^T plan is built by the IC set-up");
    const int64_t n = A->rows;
    std::vector<int32_t> rp(n + 1), ci;
    MGX_HIP(hipStreamSynchronize(A->ctx->stream));
    MGX_HIP(hipMemcpy(rp.data(), A->rp, (size_t)(n + 1) * 4, hipMemcpyDeviceToHost));
    ci.resize(std::max<int32_t>(rp[n], 0));
    if (!ci.empty())
      MGX_HIP(hipMemcpy(ci.data(), A->ci, ci.size() * 4, hipMemcpyDeviceToHost));
    build_tri_plan(rp, ci, nullptr, n, which, P);
  }
  return P;
}

TriOp tri_prepare(Csr *A, const TriPlan &P, mgx_dtype dt, int fin, const void *vals,
                  const void *scale, const int32_t *scale_pos) {
  Ctx *c = A->ctx;
  const size_t es = dt == MGX_F64 ? 8 : 4;
  const int64_t n = A->rows;
  TriOp op;
  op.plan = &P;
  op.fin = fin;
  op.coef = ws_get(c, (size_t)std::max<int64_t>(P.ndeps, 1) * es);
  op.sc = ws_get(c, (size_t)std::max<int64_t>(n, 1) * es);
  op.ginit = P.nblk > 0 || P.flow ? ws_get(c, (size_t)std::max<int64_t>(n, 1) * es) : nullptr;
  const unsigned gd = (unsigned)((P.ndeps + kBlock - 1) / kBlock);
  const unsigned gn = (unsigned)((n + kBlock - 1) / kBlock);
  if (dt == MGX_F64) {
    if (P.ndeps > 0)
      k_tri_pack<double><<<gd, kBlock, 0, c->stream>>>(P.ndeps, P.d_dent, (const double *)vals,
                                                       (double *)op.coef);
    if (n > 0)
      k_tri_scale<double><<<gn, kBlock, 0, c->stream>>>(n, P.d_rec, (const double *)scale,
                                                        scale_pos, (double *)op.sc);
  } else {
    if (P.ndeps > 0)
      k_tri_pack<float><<<gd, kBlock, 0, c->stream>>>(P.ndeps, P.d_dent, (const float *)vals,
                                                      (float *)op.coef);
    if (n > 0)
      k_tri_scale<float><<<gn, kBlock, 0, c->stream>>>(n, P.d_rec, (const float *)scale,
                                                       scale_pos, (float *)op.sc);
  }
  MGX_HIP(hipGetLastError());
  return op;
}

void tri_run(Csr *A, const TriOp &op, mgx_dtype dt, const void *in, void *out) {
  if (dt == MGX_F64)
    tri_run_fin<double>(A->ctx, op, in, out);
  else
    tri_run_fin<float>(A->ctx, op, in, out);
}

void tri_check_stall(Csr *A) {
  for (TriPlan &P : A->tri) {
    if (!P.flow || !P.d_ticket) continue;
    uint32_t flag = 0;
    MGX_HIP(hipMemcpy(&flag, P.d_ticket + 1, 4, hipMemcpyDeviceToHost));
    if (flag) {
      MGX_HIP(hipMemset(P.d_ticket + 1, 0, 4));
      set_error("triangular sweep: a row waited more than 2 s for the rows it reads; the "
                "sweep gave up (result NaN)");
      throw Error{MGX_ERR_TIMEOUT};
    }
  }
}

void levels_free(SorLevels &L) {
  dev_free(L.d_rows);
  dev_free(L.d_ptr);
  L = SorLevels();
}

void tri_free(Csr *A) {
  for (TriPlan &P : A->tri) {
    levels_free(P.lev);
    dev_free(P.d_rec);
    dev_free(P.d_dcol);
    dev_free(P.d_dent);
    dev_free(P.d_dpos);
    dev_free(P.d_blk);
    dev_free(P.d_ticket);
    P = TriPlan();
  }
}

}  // namespace mgx
