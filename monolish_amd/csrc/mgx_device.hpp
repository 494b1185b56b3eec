// mgx_device.hpp -- device-side building blocks shared by the kernels.
//
// Canonical reduction order (identical on host checker and device): a block
// of 256 lanes holds one value per lane; within each 64-lane wavefront the
// tree v[l] += v[l+off] for off = 32..1 leaves the wave sum in lane 0; the
// block result is (W0 + W1) + (W2 + W3).  A vector dot over n elements uses
// tiles of 512 elements, lane t of tile k holding x[2t]*y[2t] + x[2t+1]*y[2t+1]
// of that tile; the tile partials P_k are then summed with lane t adding
// P_t, P_{t+256}, ... in order, followed by the same block tree.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mgx_internal.hpp"

namespace mgx {

template <class T> __device__ __forceinline__ T wave_tree(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    T o = __shfl_down(v, off, 64);
    v = v + o;
  }
  return v;
}

// block of 256 threads; returns the block sum in thread 0 (other threads get
// an unspecified value).  sh must hold 4 elements; callers that reuse sh
// must __syncthreads() before writing it again.
template <class T> __device__ __forceinline__ T block_tree(T v, T *sh) {
  v = wave_tree(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  T r = 0;
  if (threadIdx.x == 0) {
    T lo = sh[0] + sh[1];
    T hi = sh[2] + sh[3];
    r = lo + hi;
  }
  return r;
}

// two independent block sums at once (shares the barrier)
template <class T>
__device__ __forceinline__ void block_tree2(T &a, T &b, T *sh) {
  a = wave_tree(a);
  b = wave_tree(b);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    sh[wv] = a;
    sh[4 + wv] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    T lo = sh[0] + sh[1], hi = sh[2] + sh[3];
    a = lo + hi;
    T lo2 = sh[4] + sh[5], hi2 = sh[6] + sh[7];
    b = lo2 + hi2;
  }
}

// Cross-block partial sums.  Partials are published with device-coherent
// (agent-scope) stores and read back with device-coherent loads, so no L2
// write-back/invalidate fence is needed: the writer waits until its store
// has completed (s_waitcnt) before it takes a ticket, and the block that
// takes the last ticket reads every partial coherently.
template <class T>
__device__ __forceinline__ void publish_partial(T *P, int64_t k, T v) {
  __hip_atomic_store(&P[k], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T> __device__ __forceinline__ T load_partial(const T *P, int64_t k) {
  return __hip_atomic_load(const_cast<T *>(&P[k]), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// stage 2 of the canonical dot: lane t sums P[t], P[t+256], ... in order
template <class T>
__device__ __forceinline__ T reduce_partials(const T *P, int64_t ntiles,
                                             T *sh) {
  T s = 0;
  int64_t k = threadIdx.x;
  // issue 8 coherent loads before adding them in order (latency bound)
  for (; k + 7 * kBlock < ntiles; k += 8 * kBlock) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = load_partial(P, k + u * kBlock);
#pragma unroll
    for (int u = 0; u < 8; u++) s = s + v[u];
  }
  for (; k < ntiles; k += kBlock) s = s + load_partial(P, k);
  return block_tree(s, sh);
}

// stage 2 over quarter (wave) partials W: tile k's partial is
// (W[4k] + W[4k+1]) + (W[4k+2] + W[4k+3]), the same value block_tree gives
// for the tile; then the canonical lane sums and block tree
template <class T>
__device__ __forceinline__ T reduce_quarters(const T *W, int64_t ntiles, T *sh) {
  T s = 0;
  for (int64_t k = threadIdx.x; k < ntiles; k += kBlock) {
    T w0 = load_partial(W, 4 * k), w1 = load_partial(W, 4 * k + 1);
    T w2 = load_partial(W, 4 * k + 2), w3 = load_partial(W, 4 * k + 3);
    T lo = w0 + w1;
    T hi = w2 + w3;
    T p = lo + hi;
    s = s + p;
  }
  return block_tree(s, sh);
}

// "last block done": every block calls this after publishing its partials
// (publish_partial); returns true in all threads of the block that arrives
// last, which may then read every block's partials with load_partial.
// tick[0] counts finished groups, tick[1 + g] the arrivals of group g; block
// blk (0 <= blk < nblocks) belongs to group blk % G with G = min(nblocks,
// kTicketGroups).  The last arrival of a group takes a group ticket; the
// block that takes the last group ticket is the last block overall.  Every
// counter it finishes is reset to zero for the next use.
__device__ __forceinline__ bool last_block_arrives(uint32_t *tick, uint32_t nblocks,
                                                   uint32_t blk, int *sh_flag) {
  __builtin_amdgcn_s_waitcnt(0);  // this lane's published stores completed
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t G = nblocks < kTicketGroups ? nblocks : kTicketGroups;
    const uint32_t g = blk % G;
    const uint32_t members = (nblocks - g + G - 1) / G;
    int last = 0;
    uint32_t t = __hip_atomic_fetch_add(&tick[1 + g], 1u, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
    if (t == members - 1) {
      __hip_atomic_store(&tick[1 + g], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      uint32_t t2 = __hip_atomic_fetch_add(&tick[0], 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
      if (t2 == G - 1) {
        __hip_atomic_store(&tick[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = 1;
      }
    }
    *sh_flag = last;
  }
  __syncthreads();
  return *sh_flag != 0;
}

// Stage 2 of the canonical reduction, spread over the arriving blocks.
// NV independent sums; sum v has one partial per unit at P + v * pstride.
// Units are tiles (Q = 1) or 128-row quarters (Q = 4, tile k's partial is
// (W[4k] + W[4k+1]) + (W[4k+2] + W[4k+3])).  Every block of the launch calls
// this once, after publishing its unit's partials with publish_partial.
// Stage-2 lane g of the canonical order adds the tile partials g, g + 256,
// ... in order; here the blocks of those tiles form group g, and the last of
// them to arrive computes lane g's sum and publishes it in S[v * 256 + g].
// Groups arrive at supergroups (g mod 16), those at the top counter; the
// block that completes the top counter runs the final block tree over the
// 256 lane sums.  It gets true and res[0..NV) (LDS, visible to all its
// threads); every other block gets false.  tp is LDS scratch of NV * 256
// values; tick holds zeroed counters and is zero again after the call.
template <class T, int NV, int Q>
__device__ __forceinline__ bool arrive_reduce(uint32_t *tick, const T *P, int64_t pstride,
                                              T *S, int64_t ntiles, int64_t unit, T *tp,
                                              T *sh, T *res, int *flag) {
  __builtin_amdgcn_s_waitcnt(0);  // this lane's published stores completed
  __syncthreads();
  if (ntiles <= kRedGroups) {
    // Small reductions (each stage-2 lane holds at most one tile): the units
    // count straight into the supergroups and those into the top counter;
    // the last block forms every lane sum itself (0 + the tile partial),
    // two fewer dependent memory round trips than the general path and the
    // same bits.
    const int64_t U = Q * ntiles;
    const int64_t SG = U < kRedSuper ? U : kRedSuper;
    if (threadIdx.x == 0) {
      const int64_t sg = unit % SG;
      const uint32_t members = (uint32_t)((U - sg + SG - 1) / SG);
      int last = 0;
      const uint32_t t = __hip_atomic_fetch_add(&tick[1 + sg], 1u, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
      if (t == members - 1) {
        __hip_atomic_store(&tick[1 + sg], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t t2 = __hip_atomic_fetch_add(&tick[0], 1u, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
        if (t2 == (uint32_t)SG - 1) {
          __hip_atomic_store(&tick[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          last = 1;
        }
      }
      *flag = last;
    }
    __syncthreads();
    if (!*flag) return false;
    T lane[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
      lane[v] = T(0);
      if ((int64_t)threadIdx.x < ntiles) {
        const T *Pv = P + v * pstride;
        const int64_t k = threadIdx.x;
        T p;
        if (Q == 4) {
          const T w0 = load_partial(Pv, 4 * k), w1 = load_partial(Pv, 4 * k + 1);
          const T w2 = load_partial(Pv, 4 * k + 2), w3 = load_partial(Pv, 4 * k + 3);
          const T lo = w0 + w1;
          const T hi = w2 + w3;
          p = lo + hi;
        } else {
          p = load_partial(Pv, k);
        }
        lane[v] = lane[v] + p;
      }
    }
    if (NV == 1) {
      const T r = block_tree(lane[0], sh);
      if (threadIdx.x == 0) res[0] = r;
    } else {
      T s0 = lane[0], s1 = lane[NV - 1];
      block_tree2(s0, s1, sh);
      if (threadIdx.x == 0) {
        res[0] = s0;
        res[1] = s1;
      }
    }
    __syncthreads();
    return true;
  }
  const int64_t G = ntiles < kRedGroups ? ntiles : kRedGroups;
  const int64_t g = (Q == 4 ? (unit >> 2) : unit) % G;
  const int64_t m = (ntiles - g + G - 1) / G;  // tiles of group g
  uint32_t *gtick = tick + 1 + kRedSuper;
  if (threadIdx.x == 0) {
    const uint32_t members = (uint32_t)(Q * m);
    int last = 0;
    const uint32_t t = __hip_atomic_fetch_add(&gtick[g], 1u, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
    if (t == members - 1) {
      __hip_atomic_store(&gtick[g], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      last = 1;
    }
    *flag = last;
  }
  __syncthreads();
  if (!*flag) return false;
  // lane g's chain: tile partials loaded in parallel, added in order
  T acc[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) acc[v] = 0;
  for (int64_t base = 0; base < m; base += kBlock) {
    const int64_t l = base + threadIdx.x;
    if (l < m) {
      const int64_t k = g + l * G;
#pragma unroll
      for (int v = 0; v < NV; v++) {
        const T *Pv = P + v * pstride;
        T p;
        if (Q == 4) {
          const T w0 = load_partial(Pv, 4 * k), w1 = load_partial(Pv, 4 * k + 1);
          const T w2 = load_partial(Pv, 4 * k + 2), w3 = load_partial(Pv, 4 * k + 3);
          const T lo = w0 + w1;
          const T hi = w2 + w3;
          p = lo + hi;
        } else {
          p = load_partial(Pv, k);
        }
        tp[v * kBlock + threadIdx.x] = p;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const int64_t cnt = m - base < kBlock ? m - base : kBlock;
      for (int64_t j = 0; j < cnt; j++) {
#pragma unroll
        for (int v = 0; v < NV; v++) acc[v] = acc[v] + tp[v * kBlock + j];
      }
    }
    __syncthreads();
  }
  const int64_t SG = G < kRedSuper ? G : kRedSuper;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int v = 0; v < NV; v++) publish_partial(S, v * kRedGroups + g, acc[v]);
    __builtin_amdgcn_s_waitcnt(0);
    const int64_t sg = g % SG;
    const uint32_t smembers = (uint32_t)((G - sg + SG - 1) / SG);
    int last = 0;
    const uint32_t t = __hip_atomic_fetch_add(&tick[1 + sg], 1u, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
    if (t == smembers - 1) {
      __hip_atomic_store(&tick[1 + sg], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t t2 = __hip_atomic_fetch_add(&tick[0], 1u, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
      if (t2 == (uint32_t)SG - 1) {
        __hip_atomic_store(&tick[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = 1;
      }
    }
    *flag = last;
  }
  __syncthreads();
  if (!*flag) return false;
  // final block tree over the lane sums (lanes without tiles hold 0)
  T s0 = threadIdx.x < G ? load_partial(S, threadIdx.x) : T(0);
  if (NV == 1) {
    const T r = block_tree(s0, sh);
    if (threadIdx.x == 0) res[0] = r;
  } else {
    T s1 = threadIdx.x < G ? load_partial(S, kRedGroups + threadIdx.x) : T(0);
    block_tree2(s0, s1, sh);
    if (threadIdx.x == 0) {
      res[0] = s0;
      res[1] = s1;
    }
  }
  __syncthreads();
  return true;
}

// Products val[e] * gv(ci[e]) for e in [e0, e1) into prod[e - e0] by the 256
// lanes of a block.  val / col_ind are read with 16-byte non-temporal loads
// from e0 rounded down to W = 16 / sizeof(T) entries (the arrays are padded
// by W entries; entries outside [e0, e1) are loaded but never used); each
// lane issues its U vector loads before any gather, then all gathers before
// any product.  Requires e1 - e0 + W - 1 <= U * 256 * W.
template <class T, int U, class G>
__device__ __forceinline__ void stream_products(const T *__restrict__ val,
                                                const int32_t *__restrict__ ci, int32_t e0,
                                                int32_t e1, T *prod, G gv) {
  constexpr int W = 16 / sizeof(T);
  typedef T VT __attribute__((ext_vector_type(W)));
  typedef int32_t VI __attribute__((ext_vector_type(W)));
  const int32_t ea = e0 & ~(W - 1);
  VT v[U];
  VI c[U];
#pragma unroll
  for (int k = 0; k < U; k++) {
    const int32_t base = ea + ((int32_t)threadIdx.x + k * kBlock) * W;
    if (base < e1) {
      v[k] = __builtin_nontemporal_load(reinterpret_cast<const VT *>(val + base));
      c[k] = __builtin_nontemporal_load(reinterpret_cast<const VI *>(ci + base));
    }
  }
  T xv[U][W];
#pragma unroll
  for (int k = 0; k < U; k++) {
    const int32_t base = ea + ((int32_t)threadIdx.x + k * kBlock) * W;
#pragma unroll
    for (int j = 0; j < W; j++)
      if (base + j >= e0 && base + j < e1) xv[k][j] = gv(c[k][j]);
  }
#pragma unroll
  for (int k = 0; k < U; k++) {
    const int32_t base = ea + ((int32_t)threadIdx.x + k * kBlock) * W;
#pragma unroll
    for (int j = 0; j < W; j++)
      if (base + j >= e0 && base + j < e1) prod[base + j - e0] = (T)v[k][j] * xv[k][j];
  }
}

// vector loads per lane that stream_products needs for `entries` entries
template <class T> constexpr int stream_u(int entries) {
  return (entries + (int)(16 / sizeof(T)) - 1 + kBlock * (int)(16 / sizeof(T)) - 1) /
         (kBlock * (int)(16 / sizeof(T)));
}

// workgroups are dispatched round-robin over the 8 XCDs (workgroup i runs on
// XCD i % 8).  remap 1 gives each XCD one contiguous range of logical blocks
// (grid a multiple of 8); remap c > 1 gives each XCD runs of c consecutive
// logical blocks, the runs dealt to the XCDs in turn (grid a multiple of
// 8 c), so the rows a run shares with its neighbour runs are fetched into
// fewer L2s while all XCDs still stream the matrix front to back together.
// remap_grid() sizes the grid.
__device__ __forceinline__ int64_t xcd_block(int remap) {
  const int64_t g = gridDim.x, b = blockIdx.x;
  if (!remap) return b;
  if (remap > 1) {
    const int64_t c = remap, k = b >> 3;
    return ((k / c) * 8 + (b & 7)) * c + k % c;
  }
  const int64_t per = g >> 3;
  return (b & 7) * per + (b >> 3);
}

// the remap actually used for nb logical blocks (none below 64 blocks) and
// the grid it needs; blocks past nb return at once
struct RemapGrid {
  int remap;
  unsigned grid;
};
inline RemapGrid remap_grid(int remap, int64_t nb) {
  const int r = nb >= 64 ? remap : 0;
  const int64_t q = r > 1 ? 8 * (int64_t)r : 8;
  return {r, (unsigned)(r ? (nb + q - 1) / q * q : nb)};
}

template <class T> struct vec2;
template <> struct vec2<double> { using type = double2; };
template <> struct vec2<float> { using type = float2; };

}  // namespace mgx
