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

// "last block done": every block calls this after publishing its partials
// (publish_partial); returns true in all threads of the block that arrives
// last, which may then read every block's partials with load_partial.  The
// counter is reset to zero by the last block for the next use.
__device__ __forceinline__ bool last_block_arrives(uint32_t *ticket,
                                                   uint32_t nblocks,
                                                   int *sh_flag) {
  __builtin_amdgcn_s_waitcnt(0);  // this lane's published stores completed
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
    int last = (t == nblocks - 1);
    if (last)
      __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    *sh_flag = last;
  }
  __syncthreads();
  return *sh_flag != 0;
}

// workgroups are dispatched round-robin over the 8 XCDs (workgroup i runs on
// XCD i % 8).  Remapping gives each XCD a contiguous range of logical blocks
// so neighbouring rows share that XCD's L2.  grid must be a multiple of 8.
__device__ __forceinline__ int64_t xcd_block(int remap) {
  const int64_t g = gridDim.x, b = blockIdx.x;
  if (!remap) return b;
  const int64_t per = g >> 3;
  return (b & 7) * per + (b >> 3);
}

template <class T> struct vec2;
template <> struct vec2<double> { using type = double2; };
template <> struct vec2<float> { using type = float2; };

}  // namespace mgx
