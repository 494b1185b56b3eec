// batch.hip -- fused Krylov iterations over many independent CRS systems in
// one launch (CG and BiCGSTAB, Jacobi or no preconditioner).
//
// Every system has a descriptor (matrix arrays, work vectors, device solve
// state) and is cut into 512-row tiles; a launch processes a list of
// (system, tile) pairs.  Reductions are per system, in the canonical tile
// order over that system's own tiles, with the last block of each system
// doing its stage 2.  Nothing is shared across systems, so a system's iterates
// are bit-identical whether it runs alone or among others, and identical to
// the reference-shaped loop (src/equation/cg.cpp:46-80,
// src/equation/bicgstab.cpp:53-129).  Systems whose solve has stopped turn
// their tiles into no-ops, so the host can queue several iterations and can
// add or retire systems between launches (continuous batching).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "mgx_batch.hpp"
#include "mgx_device.hpp"
#include "mgx_sell.hpp"

namespace mgx {

namespace {

template <class T> __device__ __forceinline__ T dsqrt(T v);
template <> __device__ __forceinline__ double dsqrt(double v) { return __builtin_sqrt(v); }
template <> __device__ __forceinline__ float dsqrt(float v) { return __builtin_sqrtf(v); }

template <class T> __device__ __forceinline__ T tload(const T *p) { return *p; }

struct Loc {
  const void *d;  // SysDesc<T>*
  int64_t tile, R0, R1;
};

template <class T>
__device__ __forceinline__ const SysDesc<T> &locate(const SysDesc<T> *sys,
                                                    const TileRef *tiles,
                                                    int64_t b, int64_t &tile) {
  if (tiles) {
    const TileRef tr = tiles[b];
    tile = tr.tile;
    return sys[tr.slot];
  }
  tile = b;
  return sys[0];
}

// the row-sum part shared by every fused SpMV: rows of one tile, chunk by
// chunk; gv(c) gives the gathered vector value at column c.  Results go to
// qloc (tile-local) and out[row].
template <class T, class G>
__device__ __forceinline__ void tile_spmv(const SysDesc<T> &d, int64_t tile,
                                          int64_t R0, T *prod, T *qloc,
                                          T *out, T *sh, G gv) {
  const int t = threadIdx.x;
  for (int32_t ck = d.tile_chunk[tile]; ck < d.tile_chunk[tile + 1]; ck++) {
    const int32_t d0 = d.chunks[ck], d1 = d.chunks[ck + 1];
    const int32_t ra = d0 & 0x7fffffff, rb = d1 & 0x7fffffff;
    if (d0 < 0) {
      const int32_t e0 = d.rp[ra], e1 = d.rp[ra + 1];
      T s = 0;
      for (int32_t e = e0 + t; e < e1; e += kBlock) {
        T pr = d.val[e] * gv(d.ci[e]);
        s = s + pr;
      }
      T res = block_tree(s, sh);
      if (t == 0) {
        qloc[ra - R0] = res;
        out[ra] = res;
      }
      __syncthreads();
      continue;
    }
    const int32_t e0 = d.rp[ra], e1 = d.rp[rb];
    // a chunk holds at most kFusedNnz entries (fused_plan_build)
    stream_products<T, stream_u<T>(kFusedNnz)>(d.val, d.ci, e0, e1, prod, gv);
    __syncthreads();
    const int32_t row = ra + t;
    if (row < rb) {
      const int32_t lo = d.rp[row] - e0, hi = d.rp[row + 1] - e0;
      T s = 0;
      for (int32_t j = lo; j < hi; j++) s = s + prod[j];
      qloc[row - R0] = s;
      out[row] = s;
    }
    __syncthreads();
  }
}

// the same row sums for a system with a sliced copy (SysDesc::sell): lanes
// 0..127 walk one row each of the 128-row quarter at R0 (sell_row: the
// products from 0 in CRS order, as tile_spmv adds them), no LDS staging
template <class T, class G>
__device__ __forceinline__ void tile_sell(const SysDesc<T> &d, int64_t R0, int64_t R1, T *qloc,
                                          T *out, G gv) {
  constexpr int W = 16 / (int)sizeof(T);
  const int t = threadIdx.x;
  if (t < kQuarter && R0 + t < R1) {
    const int64_t r = R0 + t;
    const int32_t len = sell_len<T>(d.slen, d.rp, r);
    const T s = len ? sell_row<T, W>(r, len, d.soff, d.sci, d.sval, gv) : T(0);
    qloc[t] = s;
    out[r] = s;
  }
  __syncthreads();
}

// Solve-state fields are read and written with device-coherent (agent-scope)
// accesses: the persistent kernels read them in every block after a grid
// barrier, and the other kernels share the same helpers.
template <class V> __device__ __forceinline__ V sld(const V *p) {
  return __hip_atomic_load(const_cast<V *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class V> __device__ __forceinline__ void sst(V *p, V v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// end of a CG iteration (cg.cpp:64-84): beta = <z,r>_new / <z,r>_old,
// residual nrm2(r), history, then the stop tests in the reference's order;
// called by one thread of the block that completed the reductions; S: the
// scalars' type (T, or double for the mixed-precision solver's FP32 systems)
template <class T, class S = T>
__device__ __forceinline__ void cg_finish(const SysDesc<T> &d, SolveState *st, S zr_new, S rr,
                                          int32_t *stopped) {
  const S zr_old = (S)sld(&st->zr);
  const S beta = zr_new / zr_old;
  const S resid = dsqrt(rr);
  const int64_t it = sld(&st->iter);
  if (d.rhist && it < d.rhist_len) d.rhist[it] = (double)resid;
  sst(&st->beta, (double)beta);
  sst(&st->zr, (double)zr_new);
  sst(&st->rr, (double)rr);
  sst(&st->resid, (double)resid);
  sst(&st->first, 0);
  sst(&st->iter, it + 1);
  int32_t s = kStatRunning;
  if ((double)resid < sld(&st->tol) && sld(&st->miniter) <= it + 1)
    s = MGX_OK;
  else if (resid != resid)
    s = MGX_SOLVER_RESIDUAL_NAN;
  else if (it + 1 >= sld(&st->maxiter))
    s = MGX_SOLVER_MAXITER;
  sst(&st->status, s);
  if (s != kStatRunning && stopped) atomicAdd(stopped, 1);
}

// Kernel-finished reductions (SysDesc::lf, at most kRedGroups tiles): a
// kernel publishes its tile partials and ends; every block of the next
// kernel of that system reads all of them and finishes the reduction in the
// canonical order (stage-2 lane t holds 0 + P_t, then the block tree, as
// arrive_reduce's short path), so no block waits for the others and no
// arrival counter is touched.  The block of the system's first tile writes
// the solve state; fields that the same kernel's blocks read are never
// written by it (iteration numbers alternate between ia and ib, zr between
// zrs[0] and zrs[1]).

// lane t's stage-2 value of one reduction over nt tile partials at P (Q = 4:
// quarter partials, the tile partial rebuilt as (W0 + W1) + (W2 + W3)).  The
// partials were written by an earlier kernel of the stream, so plain
// (cached) loads see them; only partials of the same kernel need the
// device-coherent loads of load_partial.
template <class T, int Q>
__device__ __forceinline__ T lf_lane(const T *P, int64_t nt) {
  T lane = T(0);
  const int64_t k = threadIdx.x;
  if (k < nt) {
    T p;
    if (Q == 4) {
      const T w0 = P[4 * k], w1 = P[4 * k + 1];
      const T w2 = P[4 * k + 2], w3 = P[4 * k + 3];
      const T lo = w0 + w1;
      const T hi = w2 + w3;
      p = lo + hi;
    } else {
      p = P[k];
    }
    lane = lane + p;
  }
  return lane;
}

// offset of the <p, q> quarter partials of an lf system in its P (the
// update kernel's <z,r> and <r,r> tile partials use [0, 2 ntiles))
template <class T> __device__ __forceinline__ int64_t lf_pq_off(const SysDesc<T> &d) {
  return 4 * d.ntiles;
}

// Start of CG iteration `it` (= the iterations completed) of an lf system,
// run by every block of the kernel that opens the iteration: finishes
// <z,r> and <r,r> of the last update, then beta, the residual and the stop
// tests of cg.cpp:64-84 (as cg_finish).  Returns false (in every thread)
// when the system stops; else beta (pz: p = z, the first iteration) for all
// threads.  `writer` (the block of tile 0) records the state.
template <class T>
__device__ __forceinline__ bool lf_open(const SysDesc<T> &d, SolveState *st, bool writer,
                                        int32_t *stopped, T *sh, T *bc, int *flag, T &beta,
                                        bool &pz, int64_t &it) {
  it = sld(&st->ia);
  const int32_t fresh = sld(&st->fresh);
  if (fresh != 0) {
    pz = fresh == 1;
    beta = pz ? T(0) : (T)sld(&st->beta);
    if (writer && threadIdx.x == 0) {
      sst(&st->zrs[it & 1], sld(&st->zr));
      sst(&st->ib, it);
    }
    return true;
  }
  pz = false;
  T a = lf_lane<T, 1>(d.P, d.ntiles), c = lf_lane<T, 1>(d.P + d.ntiles, d.ntiles);
  block_tree2(a, c, sh);
  if (threadIdx.x == 0) {
    const T zr_new = a, rr = c;
    const T zr_old = (T)sld(&st->zrs[(it - 1) & 1]);
    const T b = zr_new / zr_old;
    const T resid = dsqrt(rr);
    int32_t s = kStatRunning;
    if ((double)resid < sld(&st->tol) && sld(&st->miniter) <= it)
      s = MGX_OK;
    else if (resid != resid)
      s = MGX_SOLVER_RESIDUAL_NAN;
    else if (it >= sld(&st->maxiter))
      s = MGX_SOLVER_MAXITER;
    bc[0] = b;
    *flag = s;
    if (writer) {
      if (d.rhist && it - 1 < d.rhist_len) d.rhist[it - 1] = (double)resid;
      sst(&st->beta, (double)b);
      sst(&st->zr, (double)zr_new);
      sst(&st->rr, (double)rr);
      sst(&st->resid, (double)resid);
      sst(&st->iter, it);
      sst(&st->zrs[it & 1], (double)zr_new);
      sst(&st->ib, it);
      if (s != kStatRunning) {
        sst(&st->status, s);
        if (stopped) atomicAdd(stopped, 1);
      }
    }
  }
  __syncthreads();
  beta = bc[0];
  return *flag == kStatRunning;
}

// ---------------------------------------------------------------- CG ----
// p = z + beta p formed at each gather; q = A p; <p, q>; alpha = <z,r>/<p,q>
template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_spmv(const SysDesc<T> *sys,
                                                    const TileRef *tiles,
                                                    int64_t nlaunch, int remap,
                                                    int32_t *stopped) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *prod = reinterpret_cast<T *>(smem_raw);
  __shared__ T qloc[kQuarter];
  __shared__ T sh[8];
  __shared__ T res[2];
  __shared__ int flag;
  __shared__ T bc[2];
  const int64_t b = xcd_block(remap);  // consecutive quarters share an XCD's L2
  if (b >= nlaunch) return;
  int64_t tile;
  const SysDesc<T> &d = locate(sys, tiles, b >> 2, tile);
  SolveState *st = d.st;
  if (st->status != kStatRunning) return;
  // split-form systems are not listed: batch_launch_split runs them
  const bool once = d.once != 0;  // p already formed in pa by k_cg_p
  bool first;
  T beta;
  int64_t iter;
  if (d.lf && !once) {
    // this kernel opens the iteration (finishes the last update's reductions)
    if (!lf_open(d, st, tile == 0 && (b & 3) == 0, stopped, sh, bc, &flag, beta, first, iter))
      return;
  } else if (d.lf) {
    iter = sld(&st->ib);  // k_cg_p opened it
    first = false;
    beta = T(0);
  } else {
    iter = st->iter;
    first = st->first != 0;
    beta = (T)st->beta;
  }
  const bool odd = (iter & 1) != 0;
  const T *p_old = once ? d.pa : (odd ? d.pb : d.pa);
  T *p_new = odd ? d.pa : d.pb;
  const T *z = d.z;
  auto pv = [&](int64_t c) -> T {
    if (once) return p_old[c];
    if (first) return z[c];
    T bp = beta * p_old[c];
    return z[c] + bp;
  };
  // one block per 128-row quarter of the tile; wave 0 forms the quarter's
  // share of the tile's <p, q> tree
  const int64_t qi = tile * 4 + (b & 3);
  const int64_t R0 = d.row0 + qi * kQuarter;
  const int64_t R1 = min(d.row0 + d.n, R0 + kQuarter);
  if (d.sell)
    tile_sell(d, R0, R1, qloc, d.q, pv);
  else
    tile_spmv(d, qi, R0, prod, qloc, d.q, sh, pv);
  const int t = threadIdx.x;
  T acc0 = 0, acc1 = 0;
  if (t < 64) {
    const int64_t i0 = R0 + 2 * t;
    if (i0 < R1) {
      T p = pv(i0);
      if (!once) p_new[i0] = p;
      acc0 = p * qloc[2 * t];
    }
    if (i0 + 1 < R1) {
      T p = pv(i0 + 1);
      if (!once) p_new[i0 + 1] = p;
      acc1 = p * qloc[2 * t + 1];
    }
  }
  T s = wave_tree(acc0 + acc1);
  if (d.lf) {
    // k_cg_update finishes <p, q>
    if (t == 0) publish_partial(d.P, lf_pq_off(d) + qi, s);
    return;
  }
  if (t == 0) publish_partial(d.P, qi, s);
  if (arrive_reduce<T, 1, 4>(d.tick, d.P, 0, d.S, d.ntiles, qi, prod, sh, res, &flag)) {
    if (t == 0) {
      const T pq = res[0];
      T zr = (T)st->zr;
      T alpha = zr / pq;
      st->pq = (double)pq;
      st->alpha = (double)alpha;
    }
  }
}

// x += alpha p; r -= alpha q; z = M r; <z,r>, <r,r>; beta, residual, stop
template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_update(const SysDesc<T> *sys,
                                                      const TileRef *tiles,
                                                      int64_t nlaunch,
                                                      int32_t *stopped) {
  __shared__ T sh[8];
  __shared__ T tp[2 * kBlock];
  __shared__ T res[2];
  __shared__ int flag;
  __shared__ T bc[2];
  const int64_t b = blockIdx.x;
  if (b >= nlaunch) return;
  int64_t tile;
  const SysDesc<T> &d = locate(sys, tiles, b, tile);
  SolveState *st = d.st;
  if (st->status != kStatRunning) return;
  int64_t iter;
  T alpha;
  if (d.lf) {
    // finish <p, q> of the SpMV kernel; alpha = <z,r> / <p,q> (cg.cpp:55-58)
    iter = sld(&st->ib);
    const T lane = lf_lane<T, 4>(d.P + lf_pq_off(d), d.ntiles);
    const T pq = block_tree(lane, sh);
    if (threadIdx.x == 0) {
      const T zr = (T)sld(&st->zrs[iter & 1]);
      const T a = zr / pq;
      bc[0] = a;
      if (tile == 0) {
        sst(&st->pq, (double)pq);
        sst(&st->alpha, (double)a);
      }
    }
    __syncthreads();
    alpha = bc[0];
  } else {
    iter = st->iter;
    alpha = (T)st->alpha;
  }
  const bool odd = (iter & 1) != 0;
  const T *p_new = odd || d.once ? d.pa : d.pb;
  const T nalpha = -alpha;
  const int t = threadIdx.x;
  const int64_t R0 = d.row0 + tile * kTile;
  const int64_t R1 = min(d.row0 + d.n, R0 + kTile);
  const int64_t i0 = R0 + 2 * t;
  T zr2[2] = {0, 0}, rr2[2] = {0, 0};
  typedef T V2 __attribute__((ext_vector_type(2)));
  if (d.vec && i0 + 1 < R1) {
    // i0 is even: one 2-element vector access per array and lane
    const V2 pv = *reinterpret_cast<const V2 *>(p_new + i0);
    const V2 qv = *reinterpret_cast<const V2 *>(d.q + i0);
    V2 xv = *reinterpret_cast<const V2 *>(d.x + i0);
    V2 rv = *reinterpret_cast<const V2 *>(d.r + i0);
    V2 mv = {};
    if (d.M) mv = *reinterpret_cast<const V2 *>(d.M + i0);
    V2 zv;
#pragma unroll
    for (int e = 0; e < 2; e++) {
      T t1 = alpha * pv[e];
      xv[e] = t1 + xv[e];
      T t2 = nalpha * qv[e];
      T r_ = t2 + rv[e];
      rv[e] = r_;
      T z_ = d.M ? mv[e] * r_ : r_;
      zv[e] = z_;
      zr2[e] = z_ * r_;
      rr2[e] = r_ * r_;
    }
    *reinterpret_cast<V2 *>(d.x + i0) = xv;
    *reinterpret_cast<V2 *>(d.r + i0) = rv;
    *reinterpret_cast<V2 *>(d.z + i0) = zv;
  } else {
#pragma unroll
    for (int e = 0; e < 2; e++) {
      const int64_t i = i0 + e;
      if (i < R1) {
        T pv = p_new[i], qv = d.q[i];
        T t1 = alpha * pv;
        d.x[i] = t1 + d.x[i];
        T t2 = nalpha * qv;
        T rv = t2 + d.r[i];
        d.r[i] = rv;
        T zv = d.M ? d.M[i] * rv : rv;
        d.z[i] = zv;
        zr2[e] = zv * rv;
        rr2[e] = rv * rv;
      }
    }
  }
  T zr_a = zr2[0] + zr2[1];
  T rr_a = rr2[0] + rr2[1];
  block_tree2(zr_a, rr_a, sh);
  if (t == 0) {
    publish_partial(d.P, tile, zr_a);
    publish_partial(d.P, d.ntiles + tile, rr_a);
  }
  if (d.lf) {
    // the next iteration's opening kernel finishes <z,r> and <r,r>
    if (t == 0 && tile == 0) {
      sst(&st->ia, iter + 1);
      sst(&st->fresh, 0);
    }
    return;
  }
  if (arrive_reduce<T, 2, 1>(d.tick, d.P, d.ntiles, d.S, d.ntiles, tile, tp, sh, res,
                             &flag)) {
    if (t == 0) cg_finish<T>(d, st, res[0], res[1], stopped);
  }
}

// split form: stage 2 of the canonical reductions by one block per system,
// after the kernel that published the tile partials.  stage 0: <p, q> ->
// alpha (cg.cpp:55-58); stage 1: <z, r> and <r, r> -> cg_finish
template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_finish(const SysDesc<T> *sys,
                                                      const int32_t *slots, int nslots,
                                                      int stage, int32_t *stopped) {
  __shared__ T sh[8];
  if ((int)blockIdx.x >= nslots) return;
  const SysDesc<T> &d = sys[slots[blockIdx.x]];
  SolveState *st = d.st;
  if (st->status != kStatRunning) return;
  if (d.acc64) {
    // FP64 partials and scalars (mixed-precision inner CG)
    __shared__ double shd[8];
    const double *Pd = reinterpret_cast<const double *>(d.P);
    if (stage == 0) {
      const double pq = reduce_partials(Pd, d.ntiles, shd);
      if (threadIdx.x == 0) {
        st->pq = pq;
        st->alpha = st->zr / pq;
      }
    } else {
      const double zr_new = reduce_partials(Pd, d.ntiles, shd);
      __syncthreads();
      const double rr = reduce_partials(Pd + d.ntiles, d.ntiles, shd);
      if (threadIdx.x == 0) cg_finish<T, double>(d, st, zr_new, rr, stopped);
    }
    return;
  }
  if (stage == 0) {
    const T pq = reduce_partials(d.P, d.ntiles, sh);
    if (threadIdx.x == 0) {
      const T zr = (T)st->zr;
      st->pq = (double)pq;
      st->alpha = (double)(zr / pq);
    }
  } else {
    const T zr_new = reduce_partials(d.P, d.ntiles, sh);
    __syncthreads();
    const T rr = reduce_partials(d.P + d.ntiles, d.ntiles, sh);
    if (threadIdx.x == 0) cg_finish<T>(d, st, zr_new, rr, stopped);
  }
}

// ------------------------------------------------- persistent CG (small) --
// One cooperative launch runs up to max_iters CG iterations of one system.
// Every block owns the 128-row quarters qi = b, b + G, ... in both stages.
// Cross-block data (p, z, partials, solve state) moves through
// device-coherent accesses; the blocks meet at a grid barrier after each
// stage, where the last block to arrive runs the canonical stage 2 and
// writes the scalars before it releases the others.  Waits are bounded:
// a block that waits too long sets gs->err and leaves, and so do the others.
constexpr uint32_t kSpinLimit = 1u << 22;  // polls of about 0.5-1 us each

__device__ __forceinline__ bool grid_arrive(GridSync *gs, uint32_t G, uint32_t ngroups,
                                            uint32_t blk, uint64_t gen, int *flag) {
  __builtin_amdgcn_s_waitcnt(0);  // this lane's coherent stores completed
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t g = blk % ngroups;
    const uint64_t members = (G - g + ngroups - 1) / ngroups;
    int last = 0;
    const uint64_t t = __hip_atomic_fetch_add(&gs->group[g], (uint64_t)1, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
    if (t == members * gen - 1) {
      const uint64_t t2 = __hip_atomic_fetch_add(&gs->top, (uint64_t)1, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
      if (t2 == (uint64_t)ngroups * gen - 1) last = 1;
    }
    *flag = last;
  }
  __syncthreads();
  return *flag != 0;
}

__device__ __forceinline__ void grid_release(GridSync *gs, uint64_t gen) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) sst(&gs->release, gen);
}

// false when the wait timed out or another block reported an error
__device__ __forceinline__ bool grid_wait(GridSync *gs, uint64_t gen, int *flag) {
  if (threadIdx.x == 0) {
    int ok = 1;
    uint32_t spins = 0;
    while (sld(&gs->release) < gen) {
      __builtin_amdgcn_s_sleep(2);
      if ((++spins & 1023u) == 0 && sld(&gs->err) != 0) {
        ok = 0;
        break;
      }
      if (spins >= kSpinLimit) {
        sst(&gs->err, 1u);
        ok = 0;
        break;
      }
    }
    *flag = ok;
  }
  __syncthreads();
  return *flag != 0;
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_persist(const SysDesc<T> *sys, GridSync *gs,
                                                       uint32_t ngroups, int32_t *stopped,
                                                       int max_iters) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *prod = reinterpret_cast<T *>(smem_raw);
  __shared__ T qloc[kQuarter];
  __shared__ T sh[8];
  __shared__ int flag;
  const SysDesc<T> &d = sys[0];
  SolveState *st = d.st;
  const uint32_t G = gridDim.x, blk = blockIdx.x;
  const int t = threadIdx.x;
  const int64_t nq = 4 * d.ntiles;
  uint64_t gen = 0;
  for (int k = 0; k < max_iters; k++) {
    if (sld(&st->status) != kStatRunning) break;  // the same value in every block
    const int64_t iter = sld(&st->iter);
    const bool first = sld(&st->first) != 0;
    const T beta = (T)sld(&st->beta);
    const bool odd = (iter & 1) != 0;
    const T *p_old = odd ? d.pb : d.pa;
    T *p_new = odd ? d.pa : d.pb;
    const T *z = d.z;
    auto pv = [&](int64_t c) -> T {
      const T zc = sld(&z[c]);
      if (first) return zc;
      T bp = beta * sld(&p_old[c]);
      return zc + bp;
    };
    // stage 1: q = A p over the owned quarters, p formed at each gather,
    // quarter partials of <p, q>
    for (int64_t qi = blk; qi < nq; qi += G) {
      const int64_t R0 = qi * kQuarter;
      const int64_t R1 = min(d.n, R0 + kQuarter);
      tile_spmv(d, qi, R0, prod, qloc, d.q, sh, pv);
      T acc0 = 0, acc1 = 0;
      if (t < 64) {
        const int64_t i0 = R0 + 2 * t;
        if (i0 < R1) {
          const T p = pv(i0);
          sst(&p_new[i0], p);
          acc0 = p * qloc[2 * t];
        }
        if (i0 + 1 < R1) {
          const T p = pv(i0 + 1);
          sst(&p_new[i0 + 1], p);
          acc1 = p * qloc[2 * t + 1];
        }
      }
      const T s = wave_tree(acc0 + acc1);
      if (t == 0) publish_partial(d.P, qi, s);
      __syncthreads();  // qloc and prod are reused by the next quarter
    }
    if (grid_arrive(gs, G, ngroups, blk, ++gen, &flag)) {
      const T pq = reduce_quarters(d.P, d.ntiles, sh);
      if (t == 0) {
        const T zr = (T)sld(&st->zr);
        sst(&st->pq, (double)pq);
        sst(&st->alpha, (double)(zr / pq));
      }
      grid_release(gs, gen);
    } else if (!grid_wait(gs, gen, &flag)) {
      return;
    }
    // stage 2: x += alpha p, r -= alpha q, z = M r over the owned quarters,
    // quarter partials of <z, r> and <r, r>
    const T alpha = (T)sld(&st->alpha);
    const T nalpha = -alpha;
    for (int64_t qi = blk; qi < nq; qi += G) {
      const int64_t R0 = qi * kQuarter;
      const int64_t R1 = min(d.n, R0 + kQuarter);
      T zr2[2] = {0, 0}, rr2[2] = {0, 0};
      if (t < 64) {
#pragma unroll
        for (int e = 0; e < 2; e++) {
          const int64_t i = R0 + 2 * t + e;
          if (i < R1) {
            const T pvv = sld(&p_new[i]), qv = d.q[i];
            T t1 = alpha * pvv;
            d.x[i] = t1 + d.x[i];
            T t2 = nalpha * qv;
            T rv = t2 + d.r[i];
            d.r[i] = rv;
            T zv = d.M ? d.M[i] * rv : rv;
            sst(&d.z[i], zv);
            zr2[e] = zv * rv;
            rr2[e] = rv * rv;
          }
        }
      }
      const T zr_w = wave_tree(zr2[0] + zr2[1]);
      const T rr_w = wave_tree(rr2[0] + rr2[1]);
      if (t == 0) {
        publish_partial(d.P, qi, zr_w);
        publish_partial(d.P, nq + qi, rr_w);
      }
    }
    if (grid_arrive(gs, G, ngroups, blk, ++gen, &flag)) {
      const T zr_new = reduce_quarters(d.P, d.ntiles, sh);
      __syncthreads();
      const T rr = reduce_quarters(d.P + nq, d.ntiles, sh);
      if (t == 0) cg_finish<T>(d, st, zr_new, rr, stopped);
      grid_release(gs, gen);
    } else if (!grid_wait(gs, gen, &flag)) {
      return;
    }
  }
}

// gather-once two-kernel form, stage 1: p = z in the first iteration, else
// p = z + beta p (xpay, cg.cpp:76), in pa; k_cg_spmv then forms q = A p
template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_p(const SysDesc<T> *sys, const TileRef *tiles,
                                                 int64_t nlaunch, int32_t *stopped) {
  const int64_t b = blockIdx.x;
  if (b >= nlaunch) return;
  int64_t tile;
  const SysDesc<T> &d = locate(sys, tiles, b, tile);
  if (!d.once) return;
  SolveState *st = d.st;
  if (st->status != kStatRunning) return;
  bool first;
  T beta;
  if (d.lf) {
    // this kernel opens the iteration (finishes the last update's reductions)
    __shared__ T sh[8];
    __shared__ T bc[2];
    __shared__ int flag;
    int64_t iter;
    if (!lf_open(d, st, tile == 0, stopped, sh, bc, &flag, beta, first, iter)) return;
  } else {
    first = st->first != 0;
    beta = (T)st->beta;
  }
  const int64_t R1 = min(d.n, (tile + 1) * kTile);
  const int64_t i0 = tile * kTile + 2 * threadIdx.x;
  typedef T V2 __attribute__((ext_vector_type(2)));
  if (i0 + 1 < R1) {
    // i0 is even and the work buffers are 256-byte aligned
    const V2 zv = *reinterpret_cast<const V2 *>(d.z + i0);
    V2 pv = zv;
    if (!first) {
      pv = *reinterpret_cast<const V2 *>(d.pa + i0);
      const T b0 = beta * pv[0], b1 = beta * pv[1];
      pv[0] = zv[0] + b0;
      pv[1] = zv[1] + b1;
    }
    *reinterpret_cast<V2 *>(d.pa + i0) = pv;
  } else if (i0 < R1) {
    T pv = d.z[i0];
    if (!first) {
      const T bp = beta * d.pa[i0];
      pv = d.z[i0] + bp;
    }
    d.pa[i0] = pv;
  }
}

// ------------------------------------------- split form, one system ----
// The split form's streaming stages launched for one system at a time, its
// descriptor passed by value (a block starts without loading it through the
// tile list) and kSplitTiles consecutive tiles per block, every load of
// those tiles issued before any arithmetic.  Each tile's partials are the
// ones the tile-list kernels form (lane t: elements 2t and 2t+1, the sum of
// the two, then the block tree), so the bits do not change.  Work vectors
// start at row 0 of the system (row0 is 0 for every engine system).

// p = z + beta p (cg.cpp:75-78), 16-byte accesses
template <class T>
__global__ __launch_bounds__(kBlock) void k_split_p(const SysDesc<T> d) {
  const SolveState *st = d.st;
  if (st->status != kStatRunning) return;
  const bool first = st->first != 0;
  const T beta = (T)st->beta;
  constexpr int V = 16 / (int)sizeof(T);
  typedef T VT __attribute__((ext_vector_type(V)));
  constexpr int NV = kSplitTiles * kTile / V;  // vectors per block
  constexpr int PER = (NV + kBlock - 1) / kBlock;
  const int64_t e0 = (int64_t)blockIdx.x * kSplitTiles * kTile;
  VT zv[PER], pv[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int v = k * kBlock + (int)threadIdx.x;
    const int64_t j = e0 + (int64_t)v * V;
    if (v < NV && j + V <= d.n) {
      zv[k] = *reinterpret_cast<const VT *>(d.z + j);
      if (!first) pv[k] = *reinterpret_cast<const VT *>(d.pa + j);
    }
  }
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int v = k * kBlock + (int)threadIdx.x;
    const int64_t j = e0 + (int64_t)v * V;
    if (v >= NV || j >= d.n) continue;
    if (j + V <= d.n) {
      VT o = zv[k];
      if (!first) {
#pragma unroll
        for (int e = 0; e < V; e++) {
          const T be = beta * pv[k][e];
          o[e] = zv[k][e] + be;
        }
      }
      *reinterpret_cast<VT *>(d.pa + j) = o;
    } else {
      for (int64_t i = j; i < d.n; i++) {
        T pi = d.z[i];
        if (!first) {
          const T bp = beta * d.pa[i];
          pi = d.z[i] + bp;
        }
        d.pa[i] = pi;
      }
    }
  }
}

// the per-wave sums of kSplitTiles tiles' lane values (a: slots 0..3, b:
// slots 4..7 of each tile), then the tiles' block sums published by threads
// 0..kSplitTiles-1 in block_tree2's order
template <class T, int NR, class R = T>
__device__ __forceinline__ void split_publish(const SysDesc<T> &d, const R (*lanev)[kSplitTiles],
                                              R *sh) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kSplitTiles; k++)
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const T w = wave_tree(lanev[r][k]);
      if (lane == 0) sh[k * 8 + 4 * r + wv] = w;
    }
  __syncthreads();
  const int64_t tile = (int64_t)blockIdx.x * kSplitTiles + threadIdx.x;
  if ((int)threadIdx.x < kSplitTiles && tile < d.ntiles) {
    const R *w = sh + threadIdx.x * 8;
    R *P = reinterpret_cast<R *>(d.P);
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const R lo = w[4 * r] + w[4 * r + 1], hi = w[4 * r + 2] + w[4 * r + 3];
      publish_partial(P, r * d.ntiles + tile, lo + hi);
    }
  }
}

// x += alpha p, r -= alpha q, z = M r; tile partials of <z, r> and <r, r>
// (cg.cpp:60-73), as k_cg_update for a split system; R: the partials' type
template <class T, class R>
__global__ __launch_bounds__(kBlock) void k_split_update(const SysDesc<T> d) {
  __shared__ R sh[kSplitTiles * 8];
  const SolveState *st = d.st;
  if (st->status != kStatRunning) return;
  const T alpha = (T)st->alpha;
  const T nalpha = -alpha;
  const int t = threadIdx.x;
  typedef T V2 __attribute__((ext_vector_type(2)));
  V2 pv[kSplitTiles], qv[kSplitTiles], xv[kSplitTiles], rv[kSplitTiles], mv[kSplitTiles];
  bool full[kSplitTiles];
#pragma unroll
  for (int k = 0; k < kSplitTiles; k++) {
    const int64_t tile = (int64_t)blockIdx.x * kSplitTiles + k;
    const int64_t i0 = tile * kTile + 2 * t;
    full[k] = d.vec && tile < d.ntiles && i0 + 1 < d.n;
    if (full[k]) {
      pv[k] = *reinterpret_cast<const V2 *>(d.pa + i0);
      qv[k] = *reinterpret_cast<const V2 *>(d.q + i0);
      xv[k] = *reinterpret_cast<const V2 *>(d.x + i0);
      rv[k] = *reinterpret_cast<const V2 *>(d.r + i0);
      if (d.M) mv[k] = *reinterpret_cast<const V2 *>(d.M + i0);
    }
  }
  R lanev[2][kSplitTiles];
#pragma unroll
  for (int k = 0; k < kSplitTiles; k++) {
    const int64_t tile = (int64_t)blockIdx.x * kSplitTiles + k;
    const int64_t i0 = tile * kTile + 2 * t;
    R zr2[2] = {0, 0}, rr2[2] = {0, 0};
    if (full[k]) {
      V2 zv;
#pragma unroll
      for (int e = 0; e < 2; e++) {
        T t1 = alpha * pv[k][e];
        xv[k][e] = t1 + xv[k][e];
        T t2 = nalpha * qv[k][e];
        T r_ = t2 + rv[k][e];
        rv[k][e] = r_;
        T z_ = d.M ? mv[k][e] * r_ : r_;
        zv[e] = z_;
        zr2[e] = (R)z_ * (R)r_;
        rr2[e] = (R)r_ * (R)r_;
      }
      *reinterpret_cast<V2 *>(d.x + i0) = xv[k];
      *reinterpret_cast<V2 *>(d.r + i0) = rv[k];
      *reinterpret_cast<V2 *>(d.z + i0) = zv;
    } else if (tile < d.ntiles) {
      const int64_t R1 = min(d.n, (tile + 1) * kTile);
#pragma unroll
      for (int e = 0; e < 2; e++) {
        const int64_t i = i0 + e;
        if (i < R1) {
          T pi = d.pa[i], qi = d.q[i];
          T t1 = alpha * pi;
          d.x[i] = t1 + d.x[i];
          T t2 = nalpha * qi;
          T ri = t2 + d.r[i];
          d.r[i] = ri;
          T zi = d.M ? d.M[i] * ri : ri;
          d.z[i] = zi;
          zr2[e] = (R)zi * (R)ri;
          rr2[e] = (R)ri * (R)ri;
        }
      }
    }
    lanev[0][k] = zr2[0] + zr2[1];
    lanev[1][k] = rr2[0] + rr2[1];
  }
  split_publish<T, 2, R>(d, lanev, sh);
}

// tile partials of <p, q> when the SpMV did not form them
template <class T>
__global__ __launch_bounds__(kBlock) void k_split_pq(const SysDesc<T> d) {
  __shared__ T sh[kSplitTiles * 8];
  const SolveState *st = d.st;
  if (st->status != kStatRunning) return;
  const int t = threadIdx.x;
  T lanev[1][kSplitTiles];
#pragma unroll
  for (int k = 0; k < kSplitTiles; k++) {
    const int64_t tile = (int64_t)blockIdx.x * kSplitTiles + k;
    const int64_t R1 = min(d.n, (tile + 1) * kTile);
    const int64_t i0 = tile * kTile + 2 * t;
    T p0 = 0, p1 = 0;
    if (tile < d.ntiles) {
      if (i0 + 1 < R1) {
        p0 = d.pa[i0] * d.q[i0];
        p1 = d.pa[i0 + 1] * d.q[i0 + 1];
      } else if (i0 < R1) {
        p0 = d.pa[i0] * d.q[i0];
      }
    }
    lanev[0][k] = p0 + p1;
  }
  split_publish<T, 1>(d, lanev, sh);
}

// ---------------------------------------------------------- BiCGSTAB ----
// K1: p = r + beta (p - omega v) and phat = M p at each gather;
//     v = A phat; <v, r0>; alpha = rho / <v, r0>          (bicgstab.cpp:64-80)
template <class T>
__global__ __launch_bounds__(kBlock) void k_bcg_spmv1(const SysDesc<T> *sys,
                                                      const TileRef *tiles,
                                                      int64_t nlaunch, int remap) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *prod = reinterpret_cast<T *>(smem_raw);
  __shared__ T qloc[kQuarter];
  __shared__ T sh[8];
  __shared__ T res[2];
  __shared__ int flag;
  const int64_t b = xcd_block(remap);  // consecutive quarters share an XCD's L2
  if (b >= nlaunch) return;
  int64_t tile;
  const SysDesc<T> &d = locate(sys, tiles, b >> 2, tile);
  SolveState *st = d.st;
  if (st->status != kStatRunning) return;
  const bool odd = (st->iter & 1) != 0;
  const T *p_old = odd ? d.pb : d.pa;
  T *p_new = odd ? d.pa : d.pb;
  const T *v_old = odd ? d.vb : d.va;
  T *v_new = odd ? d.va : d.vb;
  const bool first = st->first != 0;
  const T beta = (T)st->beta;
  const T nomega = -(T)st->omega;
  const T *r = d.r;
  const T *M = d.M;
  auto pval = [&](int64_t c) -> T {
    if (first) return r[c];
    T t1 = nomega * v_old[c];
    T p1 = t1 + p_old[c];
    T t2 = beta * p1;
    return r[c] + t2;
  };
  const bool once = d.once != 0;  // p and phat already formed by k_bcg_p
  auto phv = [&](int64_t c) -> T {
    if (once) return d.phat[c];
    T p = pval(c);
    return M ? M[c] * p : p;
  };
  const int64_t qi = tile * 4 + (b & 3);
  const int64_t R0 = d.row0 + qi * kQuarter;
  const int64_t R1 = min(d.row0 + d.n, R0 + kQuarter);
  if (d.sell)
    tile_sell(d, R0, R1, qloc, v_new, phv);
  else
    tile_spmv(d, qi, R0, prod, qloc, v_new, sh, phv);
  const int t = threadIdx.x;
  T acc[2] = {0, 0};
  if (t < 64) {
    const int64_t i0 = R0 + 2 * t;
#pragma unroll
    for (int e = 0; e < 2; e++) {
      const int64_t i = i0 + e;
      if (i < R1) {
        if (!once) {
          T p = pval(i);
          p_new[i] = p;
          d.phat[i] = M ? M[i] * p : p;
        }
        acc[e] = qloc[2 * t + e] * d.r0[i];
      }
    }
  }
  T s = wave_tree(acc[0] + acc[1]);
  if (t == 0) publish_partial(d.P, qi, s);
  if (arrive_reduce<T, 1, 4>(d.tick, d.P, 0, d.S, d.ntiles, qi, prod, sh, res, &flag)) {
    if (t == 0) {
      const T vr0 = res[0];
      T rho = (T)st->rho;
      st->vr0 = (double)vr0;
      st->alpha = (double)(rho / vr0);
    }
  }
}

// K2: s = r - alpha v and shat = M s at each gather; t = A shat; <t,s>,
//     <t,t>; omega = <t,s>/<t,t>, breakdown if omega == 0 (bicgstab.cpp:83-98)
template <class T>
__global__ __launch_bounds__(kBlock) void k_bcg_spmv2(const SysDesc<T> *sys,
                                                      const TileRef *tiles,
                                                      int64_t nlaunch, int remap,
                                                      int32_t *stopped) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *prod = reinterpret_cast<T *>(smem_raw);
  __shared__ T qloc[kQuarter];
  __shared__ T sh[8];
  __shared__ T res[2];
  __shared__ int flag;
  const int64_t b = xcd_block(remap);  // consecutive quarters share an XCD's L2
  if (b >= nlaunch) return;
  int64_t tile;
  const SysDesc<T> &d = locate(sys, tiles, b >> 2, tile);
  SolveState *st = d.st;
  if (st->status != kStatRunning) return;
  const bool odd = (st->iter & 1) != 0;
  const T *v_new = odd ? d.va : d.vb;
  const T nalpha = -(T)st->alpha;
  const T *r = d.r;
  const T *M = d.M;
  auto sval = [&](int64_t c) -> T {
    T t1 = nalpha * v_new[c];
    return t1 + r[c];
  };
  const bool once = d.once != 0;  // s and shat already formed by k_bcg_s
  auto shv = [&](int64_t c) -> T {
    if (once) return d.shat[c];
    T s = sval(c);
    return M ? M[c] * s : s;
  };
  const int64_t qi = tile * 4 + (b & 3);
  const int64_t R0 = d.row0 + qi * kQuarter;
  const int64_t R1 = min(d.row0 + d.n, R0 + kQuarter);
  if (d.sell)
    tile_sell(d, R0, R1, qloc, d.t, shv);
  else
    tile_spmv(d, qi, R0, prod, qloc, d.t, sh, shv);
  const int t = threadIdx.x;
  T ts2[2] = {0, 0}, tt2[2] = {0, 0};
  if (t < 64) {
    const int64_t i0 = R0 + 2 * t;
#pragma unroll
    for (int e = 0; e < 2; e++) {
      const int64_t i = i0 + e;
      if (i < R1) {
        T sv;
        if (once) {
          sv = d.s[i];
        } else {
          sv = sval(i);
          d.s[i] = sv;
          d.shat[i] = M ? M[i] * sv : sv;
        }
        T tv = qloc[2 * t + e];
        ts2[e] = tv * sv;
        tt2[e] = tv * tv;
      }
    }
  }
  T ts_a = wave_tree(ts2[0] + ts2[1]);
  T tt_a = wave_tree(tt2[0] + tt2[1]);
  if (t == 0) {
    publish_partial(d.P, qi, ts_a);
    publish_partial(d.P, 4 * d.ntiles + qi, tt_a);
  }
  if (arrive_reduce<T, 2, 4>(d.tick, d.P, 4 * d.ntiles, d.S, d.ntiles, qi, prod, sh, res,
                             &flag)) {
    if (t == 0) {
      const T ts = res[0], tt = res[1];
      T omega = ts / tt;
      st->ts = (double)ts;
      st->tt = (double)tt;
      st->omega = (double)omega;
      if (omega == T(0)) {
        // breakdown: the residual reported is that of the current r
        st->iter = st->iter + 1;
        st->status = MGX_SOLVER_BREAKDOWN;
        if (stopped) atomicAdd(stopped, 1);
      }
    }
  }
}

// gather-once BiCGSTAB, before K1: p = r + beta (p - omega v) (p = r in the
// first iteration) and phat = M p over the system's rows, the arithmetic of
// K1's gathers (bicgstab.cpp:66-77)
template <class T>
__global__ __launch_bounds__(kBlock) void k_bcg_p(const SysDesc<T> *sys, const TileRef *tiles,
                                                  int64_t nlaunch) {
  const int64_t b = blockIdx.x;
  if (b >= nlaunch) return;
  int64_t tile;
  const SysDesc<T> &d = locate(sys, tiles, b, tile);
  if (!d.once) return;
  const SolveState *st = d.st;
  if (st->status != kStatRunning) return;
  const bool odd = (st->iter & 1) != 0;
  const T *p_old = odd ? d.pb : d.pa;
  T *p_new = odd ? d.pa : d.pb;
  const T *v_old = odd ? d.vb : d.va;
  const bool first = st->first != 0;
  const T beta = (T)st->beta;
  const T nomega = -(T)st->omega;
  const int64_t R1 = min(d.n, (tile + 1) * kTile);
  for (int64_t i = tile * kTile + threadIdx.x; i < R1; i += kBlock) {
    T p = d.r[i];
    if (!first) {
      T t1 = nomega * v_old[i];
      T p1 = t1 + p_old[i];
      T t2 = beta * p1;
      p = d.r[i] + t2;
    }
    p_new[i] = p;
    d.phat[i] = d.M ? d.M[i] * p : p;
  }
}

// gather-once BiCGSTAB, between K1 and K2: s = r - alpha v and shat = M s
// (bicgstab.cpp:83-86)
template <class T>
__global__ __launch_bounds__(kBlock) void k_bcg_s(const SysDesc<T> *sys, const TileRef *tiles,
                                                  int64_t nlaunch) {
  const int64_t b = blockIdx.x;
  if (b >= nlaunch) return;
  int64_t tile;
  const SysDesc<T> &d = locate(sys, tiles, b, tile);
  if (!d.once) return;
  const SolveState *st = d.st;
  if (st->status != kStatRunning) return;
  const bool odd = (st->iter & 1) != 0;
  const T *v_new = odd ? d.va : d.vb;
  const T nalpha = -(T)st->alpha;
  const int64_t R1 = min(d.n, (tile + 1) * kTile);
  for (int64_t i = tile * kTile + threadIdx.x; i < R1; i += kBlock) {
    T t1 = nalpha * v_new[i];
    T s = t1 + d.r[i];
    d.s[i] = s;
    d.shat[i] = d.M ? d.M[i] * s : s;
  }
}

// K3: x += alpha phat; x += omega shat; r = s - omega t; <r,r>, <r,r0>;
//     residual and stop; next rho, breakdown on rho == 0, next beta
//     (bicgstab.cpp:101-128, then :56-69 of the next iteration)
template <class T>
__global__ __launch_bounds__(kBlock) void k_bcg_update(const SysDesc<T> *sys,
                                                       const TileRef *tiles,
                                                       int64_t nlaunch,
                                                       int32_t *stopped) {
  __shared__ T sh[8];
  __shared__ T tp[2 * kBlock];
  __shared__ T res[2];
  __shared__ int flag;
  const int64_t b = blockIdx.x;
  if (b >= nlaunch) return;
  int64_t tile;
  const SysDesc<T> &d = locate(sys, tiles, b, tile);
  SolveState *st = d.st;
  if (st->status != kStatRunning) return;
  const T alpha = (T)st->alpha, omega = (T)st->omega, nomega = -omega;
  const int t = threadIdx.x;
  const int64_t R0 = d.row0 + tile * kTile;
  const int64_t R1 = min(d.row0 + d.n, R0 + kTile);
  const int64_t i0 = R0 + 2 * t;
  T rr2[2] = {0, 0}, rq2[2] = {0, 0};
  typedef T V2 __attribute__((ext_vector_type(2)));
  if (d.vec && i0 + 1 < R1) {
    const V2 ph = *reinterpret_cast<const V2 *>(d.phat + i0);
    const V2 sh_ = *reinterpret_cast<const V2 *>(d.shat + i0);
    const V2 tv = *reinterpret_cast<const V2 *>(d.t + i0);
    const V2 sv = *reinterpret_cast<const V2 *>(d.s + i0);
    const V2 r0v = *reinterpret_cast<const V2 *>(d.r0 + i0);
    V2 xv = *reinterpret_cast<const V2 *>(d.x + i0);
    V2 rv;
#pragma unroll
    for (int e = 0; e < 2; e++) {
      T a1 = alpha * ph[e];
      T x1 = a1 + xv[e];
      T a2 = omega * sh_[e];
      xv[e] = a2 + x1;
      T a3 = nomega * tv[e];
      T r_ = a3 + sv[e];
      rv[e] = r_;
      rr2[e] = r_ * r_;
      rq2[e] = r_ * r0v[e];
    }
    *reinterpret_cast<V2 *>(d.x + i0) = xv;
    *reinterpret_cast<V2 *>(d.r + i0) = rv;
  } else {
#pragma unroll
    for (int e = 0; e < 2; e++) {
      const int64_t i = i0 + e;
      if (i < R1) {
        T a1 = alpha * d.phat[i];
        T xv = a1 + d.x[i];
        T a2 = omega * d.shat[i];
        d.x[i] = a2 + xv;
        T a3 = nomega * d.t[i];
        T rv = a3 + d.s[i];
        d.r[i] = rv;
        rr2[e] = rv * rv;
        rq2[e] = rv * d.r0[i];
      }
    }
  }
  T rr_a = rr2[0] + rr2[1];
  T rq_a = rq2[0] + rq2[1];
  block_tree2(rr_a, rq_a, sh);
  if (t == 0) {
    publish_partial(d.P, tile, rr_a);
    publish_partial(d.P, d.ntiles + tile, rq_a);
  }
  if (arrive_reduce<T, 2, 1>(d.tick, d.P, d.ntiles, d.S, d.ntiles, tile, tp, sh, res,
                             &flag)) {
    if (t == 0) {
      const T rr = res[0], rho_next = res[1];
      const T resid = dsqrt(rr);
      const int64_t it = st->iter;
      if (d.rhist && it < d.rhist_len) d.rhist[it] = (double)resid;
      st->resid = (double)resid;
      st->rr = (double)rr;
      st->iter = it + 1;
      int32_t s = kStatRunning;
      if ((double)resid < st->tol && st->miniter <= it + 1) {
        s = MGX_OK;
      } else if (resid != resid) {
        s = MGX_SOLVER_RESIDUAL_NAN;
      } else if (it + 1 >= st->maxiter) {
        s = MGX_SOLVER_MAXITER;
      } else {
        const T rho = (T)st->rho;
        if (rho_next == T(0)) {
          s = MGX_SOLVER_BREAKDOWN;  // at the top of the next iteration
          st->iter = it + 2;
        } else {
          const T beta = (rho_next / rho) * (alpha / omega);
          st->beta = (double)beta;
          st->rho_old = (double)rho;
          st->rho = (double)rho_next;
          st->first = 0;
        }
      }
      st->status = s;
      if (s != kStatRunning && stopped) atomicAdd(stopped, 1);
    }
  }
}

// ------------------------------------------------------------- set-up ----
// The set-up of every system admitted since the last launch runs in a few
// launches over the list of their tiles (cg.cpp:35-45, bicgstab.cpp:41-52).

// copies staged descriptors and states into their slots
template <class T>
__global__ void k_scatter(const SysDesc<T> *ds, const SolveState *ss, const int32_t *slots,
                          SysDesc<T> *sys, SolveState *states) {
  const int i = blockIdx.x, slot = slots[i];
  const uint32_t *src = reinterpret_cast<const uint32_t *>(&ds[i]);
  uint32_t *dst = reinterpret_cast<uint32_t *>(&sys[slot]);
  for (int w = threadIdx.x; w < (int)(sizeof(SysDesc<T>) / 4); w += blockDim.x) dst[w] = src[w];
  const uint32_t *s2 = reinterpret_cast<const uint32_t *>(&ss[i]);
  uint32_t *d2 = reinterpret_cast<uint32_t *>(&states[slot]);
  for (int w = threadIdx.x; w < (int)(sizeof(SolveState) / 4); w += blockDim.x) d2[w] = s2[w];
}

// Jacobi: M = 1 / diag(A) for the tile's 512 rows (crs_getvec.cpp:8-55 then
// vml::reciprocal); the diagonal search is that of k_diag, per 256 rows
template <class T>
__global__ __launch_bounds__(kBlock) void k_setup_diag(const SysDesc<T> *sys,
                                                       const TileRef *tiles, int64_t nlaunch) {
  __shared__ int32_t srp[kBlock + 1];
  __shared__ int32_t best[kBlock];
  const int64_t b = blockIdx.x;
  if (b >= nlaunch) return;
  int64_t tile;
  const SysDesc<T> &d = locate(sys, tiles, b, tile);
  if (!d.M) return;
  T *M = const_cast<T *>(d.M);
  const int t = threadIdx.x;
  for (int half = 0; half < 2; half++) {
    const int64_t r0 = tile * kTile + half * kBlock;
    if (r0 >= d.n) break;
    const int nr = (int)(d.n - r0 < kBlock ? d.n - r0 : kBlock);
    if (t < nr) srp[t] = d.rp[r0 + t];
    if (t == 0) srp[nr] = d.rp[r0 + nr];
    best[t] = -1;
    __syncthreads();
    const int32_t e0 = srp[0], e1 = srp[nr];
    for (int32_t j = e0 + t; j < e1; j += kBlock) {
      int lo = 0, hi = nr - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (srp[mid] <= j)
          lo = mid;
        else
          hi = mid - 1;
      }
      if (d.ci[j] == (int32_t)(r0 + lo)) atomicMax(&best[lo], j);
    }
    __syncthreads();
    if (t < nr) {
      const int32_t bj = best[t];
      const T dv = bj >= 0 ? d.val[bj] : T(0);
      M[r0 + t] = T(1) / dv;
    }
    __syncthreads();
  }
}

// r = b - A x per 128-row quarter; CG: z = M r (or r); BiCGSTAB: r0 = r
template <class T>
__global__ __launch_bounds__(kBlock) void k_setup_r(const SysDesc<T> *sys, const TileRef *tiles,
                                                    int64_t nlaunch, int remap) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *prod = reinterpret_cast<T *>(smem_raw);
  __shared__ T qloc[kQuarter];
  __shared__ T sh[8];
  const int64_t b = xcd_block(remap);
  if (b >= nlaunch) return;
  int64_t tile;
  const SysDesc<T> &d = locate(sys, tiles, b >> 2, tile);
  const int64_t qi = tile * 4 + (b & 3);
  const int64_t R0 = qi * kQuarter;
  const int64_t R1 = min(d.n, R0 + kQuarter);
  const T *x = d.x;
  T *scratch = d.method == kMethodCG ? d.q : d.t;
  tile_spmv(d, qi, R0, prod, qloc, scratch, sh, [&](int64_t c) { return x[c]; });
  const int64_t i = R0 + threadIdx.x;
  if (threadIdx.x < kQuarter && i < R1) {
    const T rv = d.b[i] - qloc[threadIdx.x];
    d.r[i] = rv;
    if (d.method == kMethodCG)
      d.z[i] = d.M ? d.M[i] * rv : rv;
    else
      d.r0[i] = rv;
  }
}

// per-system tile-order dots of the set-up.  pass 0: tol = tol_scale *
// nrm2(b) for systems given a relative tolerance (others skip); pass 1:
// CG <z, r> into zr, BiCGSTAB <r, r0> (r0 = r) into rho
template <class T>
__global__ __launch_bounds__(kBlock) void k_sys_dot(const SysDesc<T> *sys,
                                                    const TileRef *tiles,
                                                    int64_t nlaunch, int pass) {
  __shared__ T sh[8];
  __shared__ T tp[2 * kBlock];
  __shared__ T res[2];
  __shared__ int flag;
  const int64_t b = blockIdx.x;
  if (b >= nlaunch) return;
  int64_t tile;
  const SysDesc<T> &d = locate(sys, tiles, b, tile);
  if (pass == 0 && d.tol_scale == 0.0) return;
  SolveState *st = d.st;
  const bool cg = d.method == kMethodCG;
  const T *xv = pass == 0 ? d.b : (cg ? d.z : d.r);
  const T *yv = pass == 0 ? d.b : d.r;
  const int64_t R0 = tile * kTile;
  const int64_t R1 = min(d.n, R0 + kTile);
  const int64_t i0 = R0 + 2 * threadIdx.x;
  T p0 = 0, p1 = 0;
  typedef T V2 __attribute__((ext_vector_type(2)));
  if (d.vec && i0 + 1 < R1) {
    const V2 a = *reinterpret_cast<const V2 *>(xv + i0);
    const V2 c = *reinterpret_cast<const V2 *>(yv + i0);
    p0 = a[0] * c[0];
    p1 = a[1] * c[1];
  } else {
    if (i0 < R1) p0 = xv[i0] * yv[i0];
    if (i0 + 1 < R1) p1 = xv[i0 + 1] * yv[i0 + 1];
  }
  T s = block_tree(p0 + p1, sh);
  if (threadIdx.x == 0) publish_partial(d.P, tile, s);
  if (arrive_reduce<T, 1, 1>(d.tick, d.P, 0, d.S, d.ntiles, tile, tp, sh, res, &flag)) {
    if (threadIdx.x == 0) {
      const T v = res[0];
      if (pass == 0)
        sst(&st->tol, d.tol_scale * (double)dsqrt(v));
      else if (cg)
        sst(&st->zr, (double)v);
      else
        sst(&st->rho, (double)v);
    }
  }
}

// BiCGSTAB set-up: rho = <r, r0> is in st->rho (r0 = r); the residual of
// r is sqrt(rho); rho == 0 is a breakdown at the first iteration
// (bicgstab.cpp:56-62).  One thread per BiCGSTAB system (its tile 0).
template <class T>
__global__ void k_bcg_init(const SysDesc<T> *sys, const TileRef *tiles, int64_t nlaunch,
                           int32_t *stopped) {
  const int64_t b = blockIdx.x;
  if (b >= nlaunch || threadIdx.x != 0) return;
  int64_t tile;
  const SysDesc<T> &d = locate(sys, tiles, b, tile);
  if (tile != 0 || d.method != kMethodBiCGSTAB) return;
  SolveState *st = d.st;
  const T rho = (T)st->rho;
  st->resid = (double)dsqrt(rho);
  if (rho == T(0)) {
    st->iter = 1;
    st->status = MGX_SOLVER_BREAKDOWN;
    if (stopped) atomicAdd(stopped, 1);
  }
}

}  // namespace

// ------------------------------------------------------------ launchers --
template <class T>
void batch_launch_setup(Ctx *c, const SysDesc<T> *stage_d, const SolveState *stage_s,
                        const int32_t *stage_slots, int nstage, SysDesc<T> *sys,
                        SolveState *states, const TileRef *tiles, int64_t ntiles, bool any_m,
                        bool any_rel, bool any_bcg, int lds_entries, int32_t *stopped) {
  if (nstage > 0)
    k_scatter<T><<<(unsigned)nstage, 64, 0, c->stream>>>(stage_d, stage_s, stage_slots, sys,
                                                         states);
  if (ntiles <= 0) return;
  const unsigned gt = (unsigned)ntiles;
  if (any_m) k_setup_diag<T><<<gt, kBlock, 0, c->stream>>>(sys, tiles, ntiles);
  const size_t lds = (size_t)std::max(lds_entries, 2 * kBlock) * sizeof(T);
  const int64_t nq = 4 * ntiles;
  const int remap = nq >= 64;
  const unsigned gq = (unsigned)(remap ? (nq + 7) / 8 * 8 : nq);
  k_setup_r<T><<<gq, kBlock, lds, c->stream>>>(sys, tiles, nq, remap);
  if (any_rel) k_sys_dot<T><<<gt, kBlock, 0, c->stream>>>(sys, tiles, ntiles, 0);
  k_sys_dot<T><<<gt, kBlock, 0, c->stream>>>(sys, tiles, ntiles, 1);
  if (any_bcg) k_bcg_init<T><<<gt, 64, 0, c->stream>>>(sys, tiles, ntiles, stopped);
  MGX_HIP(hipGetLastError());
}
template void batch_launch_setup<double>(Ctx *, const SysDesc<double> *, const SolveState *,
                                         const int32_t *, int, SysDesc<double> *, SolveState *,
                                         const TileRef *, int64_t, bool, bool, bool, int,
                                         int32_t *);
template void batch_launch_setup<float>(Ctx *, const SysDesc<float> *, const SolveState *,
                                        const int32_t *, int, SysDesc<float> *, SolveState *,
                                        const TileRef *, int64_t, bool, bool, bool, int,
                                        int32_t *);

template <class T>
void batch_launch_cg_iter(Ctx *c, const SysDesc<T> *sys, const TileRef *tiles,
                          int64_t nlaunch, int32_t *stopped, int lds_entries, bool any_fused) {
  if (nlaunch <= 0) return;
  const size_t lds = (size_t)std::max(lds_entries, 2 * kBlock) * sizeof(T);
  // SpMV stages run one block per 128-row quarter, updates one per tile
  if (any_fused) {
    const int64_t nq = 4 * nlaunch;
    const int remap = nq >= 64;
    const unsigned gq = (unsigned)(remap ? (nq + 7) / 8 * 8 : nq);
    k_cg_spmv<T><<<gq, kBlock, lds, c->stream>>>(sys, tiles, nq, remap, stopped);
  }
  k_cg_update<T><<<(unsigned)nlaunch, kBlock, 0, c->stream>>>(sys, tiles, nlaunch, stopped);
}

template <class T>
int batch_cg_persist_blocks(Ctx *c, int lds_entries) {
  const size_t lds = (size_t)std::max(lds_entries, 2 * kBlock) * sizeof(T);
  int per_cu = 0;
  MGX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void *>(&k_cg_persist<T>), kBlock, lds));
  return per_cu * c->num_cus;
}

template <class T>
void batch_launch_cg_persist(Ctx *c, const SysDesc<T> *sys, GridSync *gs, int grid,
                             int32_t *stopped, int lds_entries, int max_iters) {
  size_t lds = (size_t)std::max(lds_entries, 2 * kBlock) * sizeof(T);
  uint32_t ngroups = 1;
  while (ngroups * ngroups < (uint32_t)grid && ngroups < 64) ngroups++;
  void *args[] = {(void *)&sys, (void *)&gs, (void *)&ngroups, (void *)&stopped,
                  (void *)&max_iters};
  MGX_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&k_cg_persist<T>),
                                     dim3((unsigned)grid), dim3(kBlock), args, (unsigned)lds,
                                     c->stream));
}

template <class T>
void batch_launch_cg_p(Ctx *c, const SysDesc<T> *sys, const TileRef *tiles, int64_t nlaunch,
                       int32_t *stopped) {
  if (nlaunch > 0)
    k_cg_p<T><<<(unsigned)nlaunch, kBlock, 0, c->stream>>>(sys, tiles, nlaunch, stopped);
}

template <class T> void batch_launch_split(Ctx *c, const SysDesc<T> &d, int stage) {
  const unsigned grid = (unsigned)((d.ntiles + kSplitTiles - 1) / kSplitTiles);
  if (grid == 0) return;
  if (stage == 0)
    k_split_p<T><<<grid, kBlock, 0, c->stream>>>(d);
  else if (stage == 1)
    k_split_pq<T><<<grid, kBlock, 0, c->stream>>>(d);
  else if (sizeof(T) == 4 && d.acc64)
    k_split_update<T, double><<<grid, kBlock, 0, c->stream>>>(d);
  else
    k_split_update<T, T><<<grid, kBlock, 0, c->stream>>>(d);
}

template <class T>
void batch_launch_cg_finish(Ctx *c, const SysDesc<T> *sys, const int32_t *slots, int nslots,
                            int stage, int32_t *stopped) {
  if (nslots > 0)
    k_cg_finish<T><<<(unsigned)nslots, kBlock, 0, c->stream>>>(sys, slots, nslots, stage,
                                                               stopped);
}

template <class T>
void batch_launch_bicgstab_iter(Ctx *c, const SysDesc<T> *sys, const TileRef *tiles,
                                int64_t nlaunch, int32_t *stopped, int lds_entries,
                                bool any_once) {
  if (nlaunch <= 0) return;
  const size_t lds = (size_t)std::max(lds_entries, 2 * kBlock) * sizeof(T);
  const int64_t nq = 4 * nlaunch;
  const int remap = nq >= 64;
  const unsigned gq = (unsigned)(remap ? (nq + 7) / 8 * 8 : nq);
  if (any_once) k_bcg_p<T><<<(unsigned)nlaunch, kBlock, 0, c->stream>>>(sys, tiles, nlaunch);
  k_bcg_spmv1<T><<<gq, kBlock, lds, c->stream>>>(sys, tiles, nq, remap);
  if (any_once) k_bcg_s<T><<<(unsigned)nlaunch, kBlock, 0, c->stream>>>(sys, tiles, nlaunch);
  k_bcg_spmv2<T><<<gq, kBlock, lds, c->stream>>>(sys, tiles, nq, remap, stopped);
  k_bcg_update<T><<<(unsigned)nlaunch, kBlock, 0, c->stream>>>(sys, tiles, nlaunch, stopped);
}

template void batch_launch_cg_iter<double>(Ctx *, const SysDesc<double> *, const TileRef *,
                                           int64_t, int32_t *, int, bool);
template void batch_launch_cg_iter<float>(Ctx *, const SysDesc<float> *, const TileRef *,
                                          int64_t, int32_t *, int, bool);
template int batch_cg_persist_blocks<double>(Ctx *, int);
template int batch_cg_persist_blocks<float>(Ctx *, int);
template void batch_launch_cg_persist<double>(Ctx *, const SysDesc<double> *, GridSync *, int,
                                              int32_t *, int, int);
template void batch_launch_cg_persist<float>(Ctx *, const SysDesc<float> *, GridSync *, int,
                                             int32_t *, int, int);
template void batch_launch_cg_p<double>(Ctx *, const SysDesc<double> *, const TileRef *, int64_t,
                                        int32_t *);
template void batch_launch_cg_p<float>(Ctx *, const SysDesc<float> *, const TileRef *, int64_t,
                                       int32_t *);
template void batch_launch_split<double>(Ctx *, const SysDesc<double> &, int);
template void batch_launch_split<float>(Ctx *, const SysDesc<float> &, int);
template void batch_launch_cg_finish<double>(Ctx *, const SysDesc<double> *, const int32_t *, int,
                                             int, int32_t *);
template void batch_launch_cg_finish<float>(Ctx *, const SysDesc<float> *, const int32_t *, int,
                                            int, int32_t *);
template void batch_launch_bicgstab_iter<double>(Ctx *, const SysDesc<double> *,
                                                 const TileRef *, int64_t, int32_t *, int, bool);
template void batch_launch_bicgstab_iter<float>(Ctx *, const SysDesc<float> *, const TileRef *,
                                                int64_t, int32_t *, int, bool);

}  // namespace mgx
