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
    for (int32_t e = e0 + t; e < e1; e += kBlock) {
      T v = __builtin_nontemporal_load(&d.val[e]);
      int32_t c = __builtin_nontemporal_load(&d.ci[e]);
      prod[e - e0] = v * gv(c);
    }
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

// ---------------------------------------------------------------- CG ----
// p = z + beta p formed at each gather; q = A p; <p, q>; alpha = <z,r>/<p,q>
template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_spmv(const SysDesc<T> *sys,
                                                    const TileRef *tiles,
                                                    int64_t nlaunch) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *prod = reinterpret_cast<T *>(smem_raw);
  __shared__ T qloc[kTile];
  __shared__ T sh[8];
  __shared__ int flag;
  const int64_t b = blockIdx.x;
  if (b >= nlaunch) return;
  int64_t tile;
  const SysDesc<T> &d = locate(sys, tiles, b, tile);
  SolveState *st = d.st;
  if (st->status != kStatRunning) return;
  const bool odd = (st->iter & 1) != 0;
  const T *p_old = odd ? d.pb : d.pa;
  T *p_new = odd ? d.pa : d.pb;
  const bool first = st->first != 0;
  const T beta = (T)st->beta;
  const T *z = d.z;
  auto pv = [&](int64_t c) -> T {
    if (first) return z[c];
    T bp = beta * p_old[c];
    return z[c] + bp;
  };
  const int64_t R0 = d.row0 + tile * kTile;
  const int64_t R1 = min(d.row0 + d.n, R0 + kTile);
  tile_spmv(d, tile, R0, prod, qloc, d.q, sh, pv);
  const int t = threadIdx.x;
  const int64_t i0 = R0 + 2 * t;
  T acc0 = 0, acc1 = 0;
  if (i0 < R1) {
    T p = pv(i0);
    p_new[i0] = p;
    acc0 = p * qloc[2 * t];
  }
  if (i0 + 1 < R1) {
    T p = pv(i0 + 1);
    p_new[i0 + 1] = p;
    acc1 = p * qloc[2 * t + 1];
  }
  T s = block_tree(acc0 + acc1, sh);
  if (t == 0) publish_partial(d.P, tile, s);
  if (last_block_arrives(&st->ticket[0], (uint32_t)d.ntiles, &flag)) {
    T pq = reduce_partials(d.P, d.ntiles, sh);
    if (t == 0) {
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
  __shared__ int flag;
  const int64_t b = blockIdx.x;
  if (b >= nlaunch) return;
  int64_t tile;
  const SysDesc<T> &d = locate(sys, tiles, b, tile);
  SolveState *st = d.st;
  if (st->status != kStatRunning) return;
  const bool odd = (st->iter & 1) != 0;
  const T *p_new = odd ? d.pa : d.pb;
  const T alpha = (T)st->alpha;
  const T nalpha = -alpha;
  const int t = threadIdx.x;
  const int64_t R0 = d.row0 + tile * kTile;
  const int64_t R1 = min(d.row0 + d.n, R0 + kTile);
  const int64_t i0 = R0 + 2 * t;
  T zr2[2] = {0, 0}, rr2[2] = {0, 0};
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
  T zr_a = zr2[0] + zr2[1];
  T rr_a = rr2[0] + rr2[1];
  block_tree2(zr_a, rr_a, sh);
  if (t == 0) {
    publish_partial(d.P, tile, zr_a);
    publish_partial(d.P, d.ntiles + tile, rr_a);
  }
  if (last_block_arrives(&st->ticket[1], (uint32_t)d.ntiles, &flag)) {
    T zr_new = reduce_partials(d.P, d.ntiles, sh);
    __syncthreads();
    T rr = reduce_partials(d.P + d.ntiles, d.ntiles, sh);
    if (t == 0) {
      const T zr_old = (T)st->zr;
      const T beta = zr_new / zr_old;
      const T resid = dsqrt(rr);
      const int64_t it = st->iter;
      if (d.rhist && it < d.rhist_len) d.rhist[it] = (double)resid;
      st->beta = (double)beta;
      st->zr = (double)zr_new;
      st->rr = (double)rr;
      st->resid = (double)resid;
      st->first = 0;
      st->iter = it + 1;
      int32_t s = kStatRunning;
      if ((double)resid < st->tol && st->miniter <= it + 1)
        s = MGX_OK;
      else if (resid != resid)
        s = MGX_SOLVER_RESIDUAL_NAN;
      else if (it + 1 >= st->maxiter)
        s = MGX_SOLVER_MAXITER;
      st->status = s;
      if (s != kStatRunning && stopped) atomicAdd(stopped, 1);
    }
  }
}

// ---------------------------------------------------------- BiCGSTAB ----
// K1: p = r + beta (p - omega v) and phat = M p at each gather;
//     v = A phat; <v, r0>; alpha = rho / <v, r0>          (bicgstab.cpp:64-80)
template <class T>
__global__ __launch_bounds__(kBlock) void k_bcg_spmv1(const SysDesc<T> *sys,
                                                      const TileRef *tiles,
                                                      int64_t nlaunch) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *prod = reinterpret_cast<T *>(smem_raw);
  __shared__ T qloc[kTile];
  __shared__ T sh[8];
  __shared__ int flag;
  const int64_t b = blockIdx.x;
  if (b >= nlaunch) return;
  int64_t tile;
  const SysDesc<T> &d = locate(sys, tiles, b, tile);
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
  auto phv = [&](int64_t c) -> T {
    T p = pval(c);
    return M ? M[c] * p : p;
  };
  const int64_t R0 = d.row0 + tile * kTile;
  const int64_t R1 = min(d.row0 + d.n, R0 + kTile);
  tile_spmv(d, tile, R0, prod, qloc, v_new, sh, phv);
  const int t = threadIdx.x;
  const int64_t i0 = R0 + 2 * t;
  T acc[2] = {0, 0};
#pragma unroll
  for (int e = 0; e < 2; e++) {
    const int64_t i = i0 + e;
    if (i < R1) {
      T p = pval(i);
      p_new[i] = p;
      d.phat[i] = M ? M[i] * p : p;
      acc[e] = qloc[2 * t + e] * d.r0[i];
    }
  }
  T s = block_tree(acc[0] + acc[1], sh);
  if (t == 0) publish_partial(d.P, tile, s);
  if (last_block_arrives(&st->ticket[0], (uint32_t)d.ntiles, &flag)) {
    T vr0 = reduce_partials(d.P, d.ntiles, sh);
    if (t == 0) {
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
                                                      int64_t nlaunch,
                                                      int32_t *stopped) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *prod = reinterpret_cast<T *>(smem_raw);
  __shared__ T qloc[kTile];
  __shared__ T sh[8];
  __shared__ int flag;
  const int64_t b = blockIdx.x;
  if (b >= nlaunch) return;
  int64_t tile;
  const SysDesc<T> &d = locate(sys, tiles, b, tile);
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
  auto shv = [&](int64_t c) -> T {
    T s = sval(c);
    return M ? M[c] * s : s;
  };
  const int64_t R0 = d.row0 + tile * kTile;
  const int64_t R1 = min(d.row0 + d.n, R0 + kTile);
  tile_spmv(d, tile, R0, prod, qloc, d.t, sh, shv);
  const int t = threadIdx.x;
  const int64_t i0 = R0 + 2 * t;
  T ts2[2] = {0, 0}, tt2[2] = {0, 0};
#pragma unroll
  for (int e = 0; e < 2; e++) {
    const int64_t i = i0 + e;
    if (i < R1) {
      T sv = sval(i);
      d.s[i] = sv;
      d.shat[i] = M ? M[i] * sv : sv;
      T tv = qloc[2 * t + e];
      ts2[e] = tv * sv;
      tt2[e] = tv * tv;
    }
  }
  T ts_a = ts2[0] + ts2[1];
  T tt_a = tt2[0] + tt2[1];
  block_tree2(ts_a, tt_a, sh);
  if (t == 0) {
    publish_partial(d.P, tile, ts_a);
    publish_partial(d.P, d.ntiles + tile, tt_a);
  }
  if (last_block_arrives(&st->ticket[1], (uint32_t)d.ntiles, &flag)) {
    T ts = reduce_partials(d.P, d.ntiles, sh);
    __syncthreads();
    T tt = reduce_partials(d.P + d.ntiles, d.ntiles, sh);
    if (t == 0) {
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

// K3: x += alpha phat; x += omega shat; r = s - omega t; <r,r>, <r,r0>;
//     residual and stop; next rho, breakdown on rho == 0, next beta
//     (bicgstab.cpp:101-128, then :56-69 of the next iteration)
template <class T>
__global__ __launch_bounds__(kBlock) void k_bcg_update(const SysDesc<T> *sys,
                                                       const TileRef *tiles,
                                                       int64_t nlaunch,
                                                       int32_t *stopped) {
  __shared__ T sh[8];
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
  T rr_a = rr2[0] + rr2[1];
  T rq_a = rq2[0] + rq2[1];
  block_tree2(rr_a, rq_a, sh);
  if (t == 0) {
    publish_partial(d.P, tile, rr_a);
    publish_partial(d.P, d.ntiles + tile, rq_a);
  }
  if (last_block_arrives(&st->ticket[2], (uint32_t)d.ntiles, &flag)) {
    T rr = reduce_partials(d.P, d.ntiles, sh);
    __syncthreads();
    T rho_next = reduce_partials(d.P + d.ntiles, d.ntiles, sh);
    if (t == 0) {
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

// per-system tile-order dot for set-up: result (as double) into a SolveState
// field; the field is selected by byte offset
template <class T>
__global__ __launch_bounds__(kBlock) void k_sys_dot(const SysDesc<T> *sys,
                                                    const TileRef *tiles,
                                                    int64_t nlaunch, int which,
                                                    int field_off, double scale) {
  __shared__ T sh[8];
  __shared__ int flag;
  const int64_t b = blockIdx.x;
  if (b >= nlaunch) return;
  int64_t tile;
  const SysDesc<T> &d = locate(sys, tiles, b, tile);
  SolveState *st = d.st;
  const T *xv = which == 0 ? d.z : (which == 1 ? d.b : d.r);  // <z,r>, <b,b>, <r,r>
  const T *yv = which == 0 ? d.r : (which == 1 ? d.b : d.r);
  const int64_t R0 = d.row0 + tile * kTile;
  const int64_t R1 = min(d.row0 + d.n, R0 + kTile);
  const int64_t i0 = R0 + 2 * threadIdx.x;
  T p0 = 0, p1 = 0;
  if (i0 < R1) p0 = xv[i0] * yv[i0];
  if (i0 + 1 < R1) p1 = xv[i0 + 1] * yv[i0 + 1];
  T s = block_tree(p0 + p1, sh);
  if (threadIdx.x == 0) publish_partial(d.P, tile, s);
  if (last_block_arrives(&st->ticket[3], (uint32_t)d.ntiles, &flag)) {
    T v = reduce_partials(d.P, d.ntiles, sh);
    if (threadIdx.x == 0) {
      double *f = reinterpret_cast<double *>(reinterpret_cast<char *>(st) + field_off);
      // scale != 0: store scale * sqrt(v) (a tolerance relative to ||b||)
      *f = scale != 0.0 ? scale * (double)dsqrt(v) : (double)v;
    }
  }
}

// BiCGSTAB set-up: rho = <r, r0> is in st->rho (r0 = r); the residual of
// r is sqrt(rho); rho == 0 is a breakdown at the first iteration
template <class T>
__global__ void k_bcg_init(SolveState *st, int32_t *stopped) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
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
template <class T> void batch_launch_bcg_init(Ctx *c, SolveState *st, int32_t *stopped) {
  k_bcg_init<T><<<1, 64, 0, c->stream>>>(st, stopped);
}
template void batch_launch_bcg_init<double>(Ctx *, SolveState *, int32_t *);
template void batch_launch_bcg_init<float>(Ctx *, SolveState *, int32_t *);
template <class T>
void batch_launch_cg_iter(Ctx *c, const SysDesc<T> *sys, const TileRef *tiles,
                          int64_t nlaunch, int32_t *stopped, int lds_entries) {
  if (nlaunch <= 0) return;
  const size_t lds = (size_t)lds_entries * sizeof(T);
  k_cg_spmv<T><<<(unsigned)nlaunch, kBlock, lds, c->stream>>>(sys, tiles, nlaunch);
  k_cg_update<T><<<(unsigned)nlaunch, kBlock, 0, c->stream>>>(sys, tiles, nlaunch, stopped);
}

template <class T>
void batch_launch_bicgstab_iter(Ctx *c, const SysDesc<T> *sys, const TileRef *tiles,
                                int64_t nlaunch, int32_t *stopped, int lds_entries) {
  if (nlaunch <= 0) return;
  const size_t lds = (size_t)lds_entries * sizeof(T);
  k_bcg_spmv1<T><<<(unsigned)nlaunch, kBlock, lds, c->stream>>>(sys, tiles, nlaunch);
  k_bcg_spmv2<T><<<(unsigned)nlaunch, kBlock, lds, c->stream>>>(sys, tiles, nlaunch, stopped);
  k_bcg_update<T><<<(unsigned)nlaunch, kBlock, 0, c->stream>>>(sys, tiles, nlaunch, stopped);
}

template <class T>
void batch_launch_sys_dot(Ctx *c, const SysDesc<T> *sys, const TileRef *tiles,
                          int64_t nlaunch, int which, int field_off, double scale) {
  if (nlaunch <= 0) return;
  k_sys_dot<T><<<(unsigned)nlaunch, kBlock, 0, c->stream>>>(sys, tiles, nlaunch, which,
                                                           field_off, scale);
}

template void batch_launch_cg_iter<double>(Ctx *, const SysDesc<double> *, const TileRef *,
                                           int64_t, int32_t *, int);
template void batch_launch_cg_iter<float>(Ctx *, const SysDesc<float> *, const TileRef *,
                                          int64_t, int32_t *, int);
template void batch_launch_bicgstab_iter<double>(Ctx *, const SysDesc<double> *,
                                                 const TileRef *, int64_t, int32_t *, int);
template void batch_launch_bicgstab_iter<float>(Ctx *, const SysDesc<float> *, const TileRef *,
                                                int64_t, int32_t *, int);
template void batch_launch_sys_dot<double>(Ctx *, const SysDesc<double> *, const TileRef *,
                                           int64_t, int, int, double);
template void batch_launch_sys_dot<float>(Ctx *, const SysDesc<float> *, const TileRef *,
                                          int64_t, int, int, double);

}  // namespace mgx
