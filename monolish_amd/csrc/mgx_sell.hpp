// Sliced CRS layout shared by the SpMV kernels and the CG / BiCGSTAB engine:
// slice geometry, entry placement, and the lane-per-row walk of one row.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mgx {

// Sliced layout: rows in slices of kSlice = 64 and entries in tuples of
// W = 16 / sizeof(value) (one 16-byte value load per lane; the FP32 copy of
// an FP64 matrix gets a layout of its own with W = 4).  Slice s has a tuple
// region at soff[2s], nt_s = (soff[2s + 1] - soff[2s]) / (64 W) tuples deep,
// and a tail region at soff[2s + 1].  Row l of the slice, len entries, keeps
// its first t = min(nt_s, len / W) * W entries in tuples and the other
// len - t in its tail: entry j sits at soff[2s] + ((j / W) * 64 + l) * W +
// j % W for j < t, else at soff[2s + 1] + (j - t) * 64 + l.  nt_s is the
// deepest tuple region that keeps the slice as wide as its longest row.  One lane per
// row walks its row in storage order with no LDS: a wavefront loads 1 KiB
// of consecutive values per tuple instruction, a tail entry is one 8-byte
// (FP64) or 4-byte (FP32) load, and no lane reads a slot beyond its own
// row's entries (a 7-entry stencil row reads exactly its 7 values and
// indices); on banded and stencil matrices x is gathered from consecutive
// addresses.
constexpr int kSlice = 64;

// a row's length: FP64 layouts read their byte lengths (255: row_ptr
// decides), one byte per row instead of two row_ptr words (1-2 % on the
// 7-point FP64 SpMV); FP32 layouts read row_ptr (the byte lengths measured
// 0.5 % slower there)
template <class T>
__device__ __forceinline__ int32_t sell_len(const uint8_t *__restrict__ slen,
                                            const int32_t *__restrict__ rp, int64_t r) {
  if constexpr (sizeof(T) == 8) {
    const int32_t l = slen[r];
    return l < 255 ? l : rp[r + 1] - rp[r];
  } else {
    return rp[r + 1] - rp[r];
  }
}

// entries of a row kept in tuples: nt_s tuples at most
__host__ __device__ constexpr int32_t sell_tupled(int32_t len, int32_t nts, int w) {
  return (len / w < nts ? len / w : nts) * w;
}

__host__ __device__ constexpr int64_t sell_index(int64_t tup, int64_t tail, int32_t j,
                                                 int32_t t, int l, int w) {
  return j < t ? tup + ((int64_t)(j / w) * kSlice + l) * w + j % w
               : tail + (int64_t)(j - t) * kSlice + l;
}

// the tail entries k0 + [0, KT) of a row (those below rem): values and
// gathered x
template <class T, int KT>
__device__ __forceinline__ void sell_tail_load(int64_t tb, int32_t k0, int32_t rem,
                                               const int32_t *__restrict__ sci,
                                               const T *__restrict__ sval, T *tv,
                                               int32_t *tc) {
#pragma unroll
  for (int k = 0; k < KT; k++)
    if (k0 + k < rem) {
      tv[k] = __builtin_nontemporal_load(sval + tb + (int64_t)(k0 + k) * kSlice);
      tc[k] = __builtin_nontemporal_load(sci + tb + (int64_t)(k0 + k) * kSlice);
    }
}

template <class T, int KT>
__device__ __forceinline__ void sell_tail_add(int32_t k0, int32_t rem, const T *tv,
                                              const T *tx, T &s) {
#pragma unroll
  for (int k = 0; k < KT; k++)
    if (k0 + k < rem) {
      T p = tv[k] * tx[k];
      s = s + p;
    }
}

// one row of the sliced layout: its tuples eight entries at a time
// (nontemporal 16-byte loads), the first KT tail entries loaded with
// the first batch, the vector gathered through gv (gv(c): its value at
// column c), the products added in storage order (tuples, then the tail),
// from 0: the CPU row loop's order (crs_matvec.hpp:100-108)
template <class T, int W, class G>
__device__ __forceinline__ T sell_row(int64_t r, int32_t len, const int64_t *__restrict__ soff,
                                      const int32_t *__restrict__ sci,
                                      const T *__restrict__ sval, G gv) {
  constexpr int QT = 8 / W;  // tuples per step: 8 entries
  constexpr int KT = W - 1;  // tail entries loaded with the first tuples
  typedef T VT __attribute__((ext_vector_type(W)));
  typedef int32_t VI __attribute__((ext_vector_type(W)));
  const int64_t s2 = 2 * (r / kSlice);
  const int l = (int)(r % kSlice);
  const int64_t tup = soff[s2], tail = soff[s2 + 1];
  const int32_t nt = sell_tupled(len, (int32_t)((tail - tup) / (kSlice * W)), W) / W;
  const int32_t rem = len - nt * W;
  const int64_t base = tup + (int64_t)l * W, tb = tail + l;
  const VT *v = reinterpret_cast<const VT *>(sval + base);
  const VI *c = reinterpret_cast<const VI *>(sci + base);
  T tv[KT], tx[KT];
  int32_t tc[KT];
  sell_tail_load<T, KT>(tb, 0, rem, sci, sval, tv, tc);
  T s = 0;
  for (int32_t q0 = 0; q0 < nt; q0 += QT) {
    VT vv[QT];
    T xv[QT * W];
#pragma unroll
    for (int q = 0; q < QT; q++)
      if (q0 + q < nt) {
        const int64_t o = (int64_t)(q0 + q) * kSlice;
        vv[q] = __builtin_nontemporal_load(v + o);
        const VI cc = __builtin_nontemporal_load(c + o);
#pragma unroll
        for (int e = 0; e < W; e++) xv[q * W + e] = gv(cc[e]);
      }
    if (q0 == 0) {
#pragma unroll
      for (int k = 0; k < KT; k++)
        if (k < rem) tx[k] = gv(tc[k]);
    }
#pragma unroll
    for (int q = 0; q < QT; q++)
      if (q0 + q < nt)
#pragma unroll
        for (int e = 0; e < W; e++) {
          T p = vv[q][e] * xv[q * W + e];
          s = s + p;
        }
  }
  if (nt == 0) {
#pragma unroll
    for (int k = 0; k < KT; k++)
      if (k < rem) tx[k] = gv(tc[k]);
  }
  sell_tail_add<T, KT>(0, rem, tv, tx, s);
#pragma unroll 1
  for (int32_t k0 = KT; k0 < rem; k0 += KT) {  // longer tails
    sell_tail_load<T, KT>(tb, k0, rem, sci, sval, tv, tc);
#pragma unroll
    for (int k = 0; k < KT; k++)
      if (k0 + k < rem) tx[k] = gv(tc[k]);
    sell_tail_add<T, KT>(k0, rem, tv, tx, s);
  }
  return s;
}

}  // namespace mgx
