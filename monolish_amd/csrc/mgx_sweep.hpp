// mgx_sweep.hpp -- level-scheduled row sweeps on the device.
//
// A sweep computes each row from rows computed before it (a triangular
// solve, an SOR sweep, an incomplete factorisation).  Rows are grouped into
// levels (level of a row = 1 + the largest level of the rows it reads; see
// build_levels); rows of one level are independent, so a level runs in
// parallel while each row still does its own arithmetic in the serial
// loop's order and sees the values the serial loop would have seen.
// Schedules whose levels are at most kSweepWgRows wide run inside one
// 1024-thread workgroup with a barrier between levels (one launch per
// sweep); wider schedules, and large ones with wide levels on average,
// launch once per level.
#pragma once

#include "mgx_device.hpp"

namespace mgx {

constexpr int kSweepWgThreads = 1024;
constexpr int kSweepWgRows = 4096;  // widest level the one-workgroup form takes

template <class Op>
__global__ __launch_bounds__(kSweepWgThreads) void k_sweep_wg(const int32_t *__restrict__ lptr,
                                                              int32_t nlev,
                                                              const int32_t *__restrict__ rows,
                                                              Op op) {
  for (int32_t L = 0; L < nlev; L++) {
    const int32_t k1 = lptr[L + 1];
    for (int32_t k = lptr[L] + (int32_t)threadIdx.x; k < k1; k += kSweepWgThreads) op(rows[k]);
    __syncthreads();  // the level's rows are written before the next reads them
  }
}

template <class Op>
__global__ __launch_bounds__(kBlock) void k_sweep_level(const int32_t *__restrict__ rows,
                                                        int32_t count, Op op) {
  const int32_t k = (int32_t)(blockIdx.x * kBlock + threadIdx.x);
  if (k < count) op(rows[k]);
}

// run op(row) over every row of the schedule L, level by level, on c's stream
// schedules of at least kSweepChipRows rows whose levels average
// kSweepChipWidth rows or more launch once per level even when every level
// fits the workgroup: one CU would carry all of their rows
constexpr int64_t kSweepChipRows = 1 << 17, kSweepChipWidth = 256;

// true when run_sweep launches once per level (else one workgroup sweeps)
inline bool sweep_launches_per_level(const SorLevels &L) {
  if (L.nlev == 0) return false;
  const int64_t n = L.ptr[L.nlev];
  const bool chip = n >= kSweepChipRows && n >= kSweepChipWidth * (int64_t)L.nlev;
  return L.max_width > kSweepWgRows || chip;
}

template <class Op> void run_sweep(Ctx *c, const SorLevels &L, const Op &op) {
  if (L.nlev == 0) return;
  if (!sweep_launches_per_level(L)) {
    k_sweep_wg<Op><<<1, kSweepWgThreads, 0, c->stream>>>(L.d_ptr, L.nlev, L.d_rows, op);
  } else {
    for (int32_t l = 0; l < L.nlev; l++) {
      const int32_t cnt = L.ptr[l + 1] - L.ptr[l];
      k_sweep_level<Op><<<(unsigned)((cnt + kBlock - 1) / kBlock), kBlock, 0, c->stream>>>(
          L.d_rows + L.ptr[l], cnt, op);
    }
  }
  MGX_HIP(hipGetLastError());
}

}  // namespace mgx
