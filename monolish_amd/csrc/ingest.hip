// ingest.hip -- COO triplets and MatrixMarket files straight to a device CRS.
//
// monolish reads a MatrixMarket file with COO<T>::input_mm
// (src/utils/IO/IO_coo.cpp:69-153), optionally orders it with
// COO<T>::sort(merge) (src/utils/at_insert_sort/at_insert_sort_coo.cpp:
// 119-144) and converts it with CRS<T>(COO&) (src/utils/convert/
// convert_crs.cpp:9-43), all serially on the host.  Here the text is parsed
// by host threads and everything after it runs on the device: a stable LSD
// radix sort of the (row, col) keys, the duplicate merge, and row_ptr.
//
// Sort keys: key = row << cbits | col, cbits = bits of (cols - 1); 8-bit
// digits, one pass per digit of the key's width.  A pass is three launches
// over 4096-key tiles: a per-tile digit histogram, an exclusive scan of the
// digit-major histogram table, and a stable scatter in which each wavefront
// ranks its keys among equal digits with ballots.  The permutation rides
// along as the payload, so values move once, after the last pass.
#include <sys/mman.h>
#include <sys/stat.h>
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "mgx_device.hpp"

namespace mgx {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int64_t kSortTile = (int64_t)kSortThreads * kSortItems;  // keys per tile
constexpr int kScanItems = 8;
constexpr int64_t kScanTile = (int64_t)kBlock * kScanItems;

// ------------------------------------------------------------ scans ------
// exclusive scan of each 2048-element tile in place; sums[tile] = its total
__global__ __launch_bounds__(kBlock) void k_scan_tiles(uint32_t *a, int64_t n, uint32_t *sums) {
  __shared__ uint32_t s[kBlock];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems], run = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    v[k] = base + k < n ? a[base + k] : 0u;
    run += v[k];
  }
  s[threadIdx.x] = run;
  __syncthreads();
  // Hillis-Steele over the 256 thread totals
  for (int off = 1; off < kBlock; off <<= 1) {
    const uint32_t add = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += add;
    __syncthreads();
  }
  uint32_t acc = s[threadIdx.x] - run;  // exclusive prefix of this thread
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    if (base + k < n) a[base + k] = acc;
    acc += v[k];
  }
  if (threadIdx.x == kBlock - 1) sums[blockIdx.x] = s[kBlock - 1];
}

__global__ void k_scan_add(uint32_t *a, int64_t n, const uint32_t *sums) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) a[i] += sums[i / kScanTile];
}

// in-place exclusive scan of n uint32 (totals must fit 32 bits)
void scan_exclusive(Ctx *c, uint32_t *a, int64_t n) {
  if (n <= 0) return;
  const int64_t nt = (n + kScanTile - 1) / kScanTile;
  uint32_t *sums = (uint32_t *)dev_alloc((size_t)nt * 4);
  try {
    k_scan_tiles<<<(unsigned)nt, kBlock, 0, c->stream>>>(a, n, sums);
    MGX_HIP(hipGetLastError());
    if (nt > 1) {
      scan_exclusive(c, sums, nt);
      k_scan_add<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, c->stream>>>(a, n, sums);
      MGX_HIP(hipGetLastError());
    }
  } catch (...) {
    (void)hipStreamSynchronize(c->stream);
    dev_free(sums);
    throw;
  }
  MGX_HIP(hipStreamSynchronize(c->stream));
  dev_free(sums);
}

// ------------------------------------------------------- radix sort ------
__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const uint64_t *keys, int64_t n,
                                                            int shift, uint32_t *hist,
                                                            int64_t ntiles) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
  for (int k = 0; k < kSortItems; k++) {
    const int64_t i = base + (int64_t)k * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[(unsigned)(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one tile.  Items are taken in index order (round, wave,
// lane); within a wavefront the lanes with the same digit are found with 8
// ballots, so a key's position is: the tile's running offset for its digit
// + the counts of earlier wavefronts of this round + its rank in the wave.
__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(
    const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin, uint64_t *__restrict__ kout,
    uint32_t *__restrict__ vout, int64_t n, int shift, const uint32_t *__restrict__ offs,
    int64_t ntiles) {
  constexpr int kWaves = kSortThreads / 64;
  __shared__ uint32_t run[256];
  __shared__ uint32_t wcnt[kWaves][256];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  run[t] = offs[(int64_t)t * ntiles + blockIdx.x];
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
  for (int k = 0; k < kSortItems; k++) {
    const int64_t i = base + (int64_t)k * kSortThreads + t;
    const bool ok = i < n;
    const uint64_t key = ok ? kin[i] : 0ull;
    const uint32_t d = (uint32_t)(key >> shift) & 255u;
#pragma unroll
    for (int q = 0; q < 256 / 64; q++) wcnt[w][lane + 64 * q] = 0;
    uint64_t same = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      same &= bit ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(same & below);
    __syncthreads();  // run[] of the previous round and the cleared counts
    if (ok && rank == 0) wcnt[w][d] = (uint32_t)__popcll(same);
    __syncthreads();
    if (ok) {
      uint32_t pos = run[d] + rank;
      for (int v = 0; v < w; v++) pos += wcnt[v][d];
      kout[pos] = key;
      vout[pos] = vin[i];
    }
    __syncthreads();  // every read of run[] and wcnt[] of this round is done
    uint32_t tot = 0;
#pragma unroll
    for (int v = 0; v < kWaves; v++) tot += wcnt[v][t];
    run[t] += tot;
    __syncthreads();  // before the next round clears wcnt[]
  }
}

__global__ void k_iota(uint32_t *v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) v[i] = (uint32_t)i;
}

__global__ void k_make_keys(const int32_t *row, const int32_t *col, int64_t n, int cbits,
                            uint64_t *keys) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) keys[i] = ((uint64_t)(uint32_t)row[i] << cbits) | (uint32_t)col[i];
}

// sorted keys back to rows / columns; values gathered through the permutation
template <class T>
__global__ void k_unpack(const uint64_t *keys, const uint32_t *perm, const T *vin, int64_t n,
                         int cbits, int32_t *row, int32_t *col, T *vout) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  row[i] = (int32_t)(k >> cbits);
  col[i] = (int32_t)(k & ((1ull << cbits) - 1));
  vout[i] = vin[perm[i]];
}

// merge: entry i is kept when it is the last of its (row, col) group, so the
// group's value is the last one in input order (the stable sort kept it)
__global__ void k_last_of_group(const uint64_t *keys, int64_t n, uint32_t *flag) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) flag[i] = (i + 1 == n || keys[i] != keys[i + 1]) ? 1u : 0u;
}

template <class T>
__global__ void k_compact(const int32_t *row, const int32_t *col, const T *val,
                          const uint32_t *flag, const uint32_t *pos, int64_t n, int32_t *row_o,
                          int32_t *col_o, T *val_o) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const uint32_t p = pos[i];
  row_o[p] = row[i];
  col_o[p] = col[i];
  val_o[p] = val[i];
}

// row_ptr of row-ordered entries: the entry that starts row r (or the end,
// i = n) writes row_ptr[prev + 1 .. r], prev = the row of entry i - 1; every
// row_ptr index is written exactly once
__global__ void k_row_ptr(const int32_t *row, int64_t n, int64_t rows, int32_t *rp) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i > n) return;
  const int64_t prev = i > 0 ? row[i - 1] : -1;
  const int64_t cur = i < n ? row[i] : rows;
  if (i < n && cur == prev) return;
  for (int64_t r = prev + 1; r <= cur && r <= rows; r++) rp[r] = (int32_t)i;
}

int bits_for(int64_t v) {  // bits to hold 0..v-1 (at least 1)
  int b = 1;
  while (b < 62 && (1ll << b) < v) b++;
  return b;
}

unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>((n + kBlock - 1) / kBlock, 1); }

// stable sort of n (key, payload) pairs on the low `bits` bits of the keys;
// the result is left in (k0, v0), (k1, v1) is scratch of the same size
void radix_sort(Ctx *c, uint64_t *&k0, uint32_t *&v0, uint64_t *&k1, uint32_t *&v1, int64_t n,
                int bits) {
  const int64_t ntiles = (n + kSortTile - 1) / kSortTile;
  uint32_t *hist = (uint32_t *)dev_alloc((size_t)std::max<int64_t>(ntiles, 1) * 256 * 4);
  try {
    for (int shift = 0; shift < bits; shift += 8) {
      k_sort_hist<<<(unsigned)ntiles, kSortThreads, 0, c->stream>>>(k0, n, shift, hist, ntiles);
      MGX_HIP(hipGetLastError());
      scan_exclusive(c, hist, ntiles * 256);
      k_sort_scatter<<<(unsigned)ntiles, kSortThreads, 0, c->stream>>>(k0, v0, k1, v1, n, shift,
                                                                      hist, ntiles);
      MGX_HIP(hipGetLastError());
      std::swap(k0, k1);
      std::swap(v0, v1);
    }
  } catch (...) {
    (void)hipStreamSynchronize(c->stream);
    dev_free(hist);
    throw;
  }
  MGX_HIP(hipStreamSynchronize(c->stream));
  dev_free(hist);
}

template <class T>
void sort_entries(Ctx *c, int64_t rows, int64_t cols, int64_t nnz, int32_t *d_row, int32_t *d_col,
                  T *d_val) {
  const int cbits = bits_for(cols);
  const int bits = bits_for(rows) + cbits;
  std::vector<void *> tmp;
  auto take = [&](size_t b) {
    void *p = dev_alloc(b);
    tmp.push_back(p);
    return p;
  };
  try {
    uint64_t *k0 = (uint64_t *)take((size_t)nnz * 8), *k1 = (uint64_t *)take((size_t)nnz * 8);
    uint32_t *v0 = (uint32_t *)take((size_t)nnz * 4), *v1 = (uint32_t *)take((size_t)nnz * 4);
    T *vals = (T *)take((size_t)nnz * sizeof(T));
    k_make_keys<<<grid_for(nnz), kBlock, 0, c->stream>>>(d_row, d_col, nnz, cbits, k0);
    k_iota<<<grid_for(nnz), kBlock, 0, c->stream>>>(v0, nnz);
    MGX_HIP(hipGetLastError());
    radix_sort(c, k0, v0, k1, v1, nnz, bits);
    MGX_HIP(hipMemcpyAsync(vals, d_val, (size_t)nnz * sizeof(T), hipMemcpyDeviceToDevice,
                           c->stream));
    k_unpack<T><<<grid_for(nnz), kBlock, 0, c->stream>>>(k0, v0, vals, nnz, cbits, d_row, d_col,
                                                         d_val);
    MGX_HIP(hipGetLastError());
    MGX_HIP(hipStreamSynchronize(c->stream));
  } catch (...) {
    (void)hipStreamSynchronize(c->stream);
    for (void *p : tmp) dev_free(p);
    throw;
  }
  for (void *p : tmp) dev_free(p);
}

// row index of every entry of a CRS (one thread per row)
__global__ void k_entry_rows(const int32_t *rp, int64_t rows, int32_t *row) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= rows) return;
  for (int32_t j = rp[r]; j < rp[r + 1]; j++) row[j] = (int32_t)r;
}

// A^T on the device: entries sorted stably by (col, row), so row c of A^T
// lists column c's entries in ascending original row -- the order in which
// the reference's serial transposed product (crs_matvec.hpp:94-97) adds them
template <class T> Csr *transpose_device(Csr *A) {
  Ctx *c = A->ctx;
  std::vector<int32_t> rp(A->rows + 1);
  MGX_HIP(hipStreamSynchronize(c->stream));
  MGX_HIP(hipMemcpy(rp.data(), A->rp, rp.size() * 4, hipMemcpyDeviceToHost));
  const int64_t used = rp[A->rows];
  Csr *Tm = new mgx_csr();
  Tm->ctx = c;
  Tm->rows = A->cols;
  Tm->cols = A->rows;
  Tm->nnz = used;
  Tm->dtype = A->dtype;
  int32_t *d_row = nullptr, *d_col = nullptr;
  try {
    const size_t ne = (size_t)std::max<int64_t>(used, 1);
    d_row = (int32_t *)dev_alloc(ne * 4);  // becomes A^T's row = A's column
    d_col = (int32_t *)dev_alloc(ne * 4);
    Tm->rp = (int32_t *)dev_alloc((size_t)(Tm->rows + 1) * 4);
    alloc_entries(Tm, used);
    if (used > 0) {
      k_entry_rows<<<grid_for(A->rows), kBlock, 0, c->stream>>>(A->rp, A->rows, d_col);
      MGX_HIP(hipGetLastError());
      MGX_HIP(hipMemcpyAsync(d_row, A->ci, (size_t)used * 4, hipMemcpyDeviceToDevice, c->stream));
      MGX_HIP(hipMemcpyAsync(Tm->val, A->val, (size_t)used * sizeof(T), hipMemcpyDeviceToDevice,
                             c->stream));
      sort_entries<T>(c, Tm->rows, Tm->cols, used, d_row, d_col, static_cast<T *>(Tm->val));
      MGX_HIP(hipMemcpyAsync(Tm->ci, d_col, (size_t)used * 4, hipMemcpyDeviceToDevice, c->stream));
    }
    k_row_ptr<<<grid_for(used + 1), kBlock, 0, c->stream>>>(d_row, used, Tm->rows, Tm->rp);
    MGX_HIP(hipGetLastError());
    std::vector<int32_t> trp(Tm->rows + 1);
    MGX_HIP(hipMemcpyAsync(trp.data(), Tm->rp, trp.size() * 4, hipMemcpyDeviceToHost, c->stream));
    MGX_HIP(hipStreamSynchronize(c->stream));
    spmv_plan_build(c, Tm, trp.data());
  } catch (...) {
    dev_free(d_row);
    dev_free(d_col);
    csr_free(Tm);
    throw;
  }
  dev_free(d_row);
  dev_free(d_col);
  return Tm;
}

// duplicates (equal row and col, adjacent after the sort) collapse to the
// last one; returns the new entry count
template <class T>
int64_t merge_entries(Ctx *c, int64_t cols, int64_t nnz, int32_t *d_row, int32_t *d_col,
                      T *d_val) {
  if (nnz == 0) return 0;
  const int cbits = bits_for(cols);
  std::vector<void *> tmp;
  auto take = [&](size_t b) {
    void *p = dev_alloc(b);
    tmp.push_back(p);
    return p;
  };
  int64_t m = 0;
  try {
    uint64_t *keys = (uint64_t *)take((size_t)nnz * 8);
    uint32_t *flag = (uint32_t *)take((size_t)nnz * 4), *pos = (uint32_t *)take((size_t)nnz * 4);
    int32_t *r2 = (int32_t *)take((size_t)nnz * 4), *c2 = (int32_t *)take((size_t)nnz * 4);
    T *v2 = (T *)take((size_t)nnz * sizeof(T));
    k_make_keys<<<grid_for(nnz), kBlock, 0, c->stream>>>(d_row, d_col, nnz, cbits, keys);
    k_last_of_group<<<grid_for(nnz), kBlock, 0, c->stream>>>(keys, nnz, flag);
    MGX_HIP(hipGetLastError());
    MGX_HIP(hipMemcpyAsync(pos, flag, (size_t)nnz * 4, hipMemcpyDeviceToDevice, c->stream));
    scan_exclusive(c, pos, nnz);
    uint32_t last_pos = 0, last_flag = 0;
    MGX_HIP(hipMemcpy(&last_pos, pos + nnz - 1, 4, hipMemcpyDeviceToHost));
    MGX_HIP(hipMemcpy(&last_flag, flag + nnz - 1, 4, hipMemcpyDeviceToHost));
    m = (int64_t)last_pos + last_flag;
    k_compact<T><<<grid_for(nnz), kBlock, 0, c->stream>>>(d_row, d_col, d_val, flag, pos, nnz, r2,
                                                          c2, v2);
    MGX_HIP(hipGetLastError());
    MGX_HIP(hipMemcpyAsync(d_row, r2, (size_t)m * 4, hipMemcpyDeviceToDevice, c->stream));
    MGX_HIP(hipMemcpyAsync(d_col, c2, (size_t)m * 4, hipMemcpyDeviceToDevice, c->stream));
    MGX_HIP(hipMemcpyAsync(d_val, v2, (size_t)m * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
    MGX_HIP(hipStreamSynchronize(c->stream));
  } catch (...) {
    (void)hipStreamSynchronize(c->stream);
    for (void *p : tmp) dev_free(p);
    throw;
  }
  for (void *p : tmp) dev_free(p);
  return m;
}

template <class F> void host_parallel(int64_t n, F fn) {
  unsigned hw = std::thread::hardware_concurrency();
  const int nt = (int)std::min<int64_t>(std::max(1u, std::min(hw, 16u)),
                                        std::max<int64_t>(1, n / 65536));
  if (nt <= 1) {
    fn(0, (int64_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  const int64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; t++) {
    const int64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b < e) th.emplace_back(fn, t, b, e);
  }
  for (auto &t : th) t.join();
}

// rows of input-ordered entries must not decrease (CRS(COO&) reads them so)
bool rows_sorted(const int32_t *row, int64_t nnz) {
  std::atomic<bool> ok(true);
  host_parallel(nnz, [&](int, int64_t b, int64_t e) {
    for (int64_t i = std::max<int64_t>(b, 1); i < e && ok.load(std::memory_order_relaxed); i++)
      if (row[i] < row[i - 1]) ok.store(false);
  });
  return ok.load();
}

void check_indices(const int32_t *row, const int32_t *col, int64_t nnz, int64_t rows,
                   int64_t cols) {
  std::atomic<int64_t> bad(-1);
  host_parallel(nnz, [&](int, int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++)
      if (row[i] < 0 || row[i] >= rows || col[i] < 0 || col[i] >= cols) {
        bad.store(i);
        return;
      }
  });
  MGX_REQUIRE(bad.load() < 0, "entry %lld: (row, col) = (%d, %d) outside %lld x %lld",
              (long long)bad.load(), bad.load() >= 0 ? row[bad.load()] : 0,
              bad.load() >= 0 ? col[bad.load()] : 0, (long long)rows, (long long)cols);
}

// host COO (row-ordered unless sort) -> device CRS
template <class T>
Csr *csr_from_host_coo(Ctx *c, int64_t rows, int64_t cols, int64_t nnz, const int32_t *row,
                       const int32_t *col, const T *val, int sort) {
  check_indices(row, col, nnz, rows, cols);
  MGX_REQUIRE(sort != 0 || rows_sorted(row, nnz),
              "COO rows are not in ascending order: CRS(COO&) needs a row-sorted COO "
              "(pass sort = 1 or 2)");
  int32_t *d_row = nullptr, *d_col = nullptr;
  T *d_val = nullptr;
  Csr *A = new mgx_csr();
  A->ctx = c;
  A->rows = rows;
  A->cols = cols;
  A->dtype = std::is_same<T, double>::value ? MGX_F64 : MGX_F32;
  try {
    const size_t ne = (size_t)std::max<int64_t>(nnz, 1);
    d_row = (int32_t *)dev_alloc(ne * 4);
    d_col = (int32_t *)dev_alloc(ne * 4);
    d_val = (T *)dev_alloc(ne * sizeof(T));
    if (nnz > 0) {
      MGX_HIP(hipMemcpyAsync(d_row, row, (size_t)nnz * 4, hipMemcpyHostToDevice, c->stream));
      MGX_HIP(hipMemcpyAsync(d_col, col, (size_t)nnz * 4, hipMemcpyHostToDevice, c->stream));
      MGX_HIP(hipMemcpyAsync(d_val, val, (size_t)nnz * sizeof(T), hipMemcpyHostToDevice,
                             c->stream));
      MGX_HIP(hipStreamSynchronize(c->stream));
    }
    int64_t m = nnz;
    if (sort && nnz > 1) sort_entries<T>(c, rows, cols, nnz, d_row, d_col, d_val);
    if (sort == 2) m = merge_entries<T>(c, cols, nnz, d_row, d_col, d_val);
    A->nnz = m;
    A->rp = (int32_t *)dev_alloc((size_t)(rows + 1) * 4);
    alloc_entries(A, m);
    k_row_ptr<<<grid_for(m + 1), kBlock, 0, c->stream>>>(d_row, m, rows, A->rp);
    MGX_HIP(hipGetLastError());
    if (m > 0) {
      MGX_HIP(hipMemcpyAsync(A->ci, d_col, (size_t)m * 4, hipMemcpyDeviceToDevice, c->stream));
      MGX_HIP(hipMemcpyAsync(A->val, d_val, (size_t)m * sizeof(T), hipMemcpyDeviceToDevice,
                             c->stream));
    }
    std::vector<int32_t> rp(rows + 1);
    MGX_HIP(hipMemcpyAsync(rp.data(), A->rp, (size_t)(rows + 1) * 4, hipMemcpyDeviceToHost,
                           c->stream));
    int32_t last_row = -1;
    if (m > 0)
      MGX_HIP(hipMemcpyAsync(&last_row, d_row + m - 1, 4, hipMemcpyDeviceToHost, c->stream));
    MGX_HIP(hipStreamSynchronize(c->stream));
    // CRS(COO&) leaves row_ptr[r] = 0 for r > last non-empty row + 1
    if (last_row + 2 <= rows) {
      A->rp_raw = rp;
      for (int64_t r = (int64_t)last_row + 2; r <= rows; r++) A->rp_raw[r] = 0;
    }
    spmv_plan_build(c, A, rp.data());
    if (rows == cols && rows > 0) fused_plan_build(A, rp.data());
    if (m >= 64) {  // a contiguous sample of the columns from the middle
      const int64_t k = std::min<int64_t>(m, 4096), e0 = (m - k) / 2;
      std::vector<int32_t> smp(k);
      MGX_HIP(hipMemcpy(smp.data(), A->ci + e0, (size_t)k * 4, hipMemcpyDeviceToHost));
      A->gather_lines = sample_gather_lines(smp.data(), k, sizeof(T));
    }
  } catch (...) {
    dev_free(d_row);
    dev_free(d_col);
    dev_free(d_val);
    csr_free(A);
    throw;
  }
  dev_free(d_row);
  dev_free(d_col);
  dev_free(d_val);
  return A;
}

// ------------------------------------------------- MatrixMarket text -----
struct Mapped {
  const char *p = nullptr;
  size_t n = 0;
  int fd = -1;
  ~Mapped() {
    if (p && n) munmap(const_cast<char *>(p), n);
    if (fd >= 0) close(fd);
  }
};

inline bool is_space(char ch) { return ch == ' ' || ch == '\t' || ch == '\r'; }

// a value from_chars reports out of range (subnormal or beyond the largest
// finite value): the C library's strtod / strtof result, as the stream
// reader's conversion gives it
template <class T> T parse_range(const char *b, const char *t, T) {
  char buf[128];
  const size_t n = std::min<size_t>((size_t)(t - b), sizeof buf - 1);
  memcpy(buf, b, n);
  buf[n] = 0;
  if (std::is_same<T, double>::value) return (T)strtod(buf, nullptr);
  return (T)strtof(buf, nullptr);
}

// one whitespace-separated token of [s, e) starting at s; s advanced past it
inline bool token(const char *&s, const char *e, const char *&b, const char *&t) {
  while (s < e && is_space(*s)) s++;
  b = s;
  while (s < e && !is_space(*s) && *s != '\n') s++;
  t = s;
  return b < t;
}

template <class T> struct Coo {
  int64_t rows = 0, cols = 0, nnz = 0;
  std::vector<int32_t> row, col;
  std::vector<T> val;
};

// COO<T>::input_mm: the banner must read "%%MatrixMarket matrix coordinate
// real general", '%' lines up to the size line are skipped, then nnz lines
// "i j value" (1-based) follow; a value is read as T directly (the FP32
// reader rounds the decimal text once, to float).  Errors are reported
// instead of the reference's exit().
template <class T> void read_mm(const char *path, Coo<T> &out) {
  Mapped f;
  f.fd = open(path, O_RDONLY);
  MGX_REQUIRE(f.fd >= 0, "input_mm: cannot open file %s", path);
  struct stat st;
  MGX_REQUIRE(fstat(f.fd, &st) == 0, "input_mm: cannot stat %s", path);
  f.n = (size_t)st.st_size;
  MGX_REQUIRE(f.n > 0, "input_mm: %s is empty", path);
  void *m = mmap(nullptr, f.n, PROT_READ, MAP_PRIVATE, f.fd, 0);
  MGX_REQUIRE(m != MAP_FAILED, "input_mm: cannot map %s", path);
  f.p = static_cast<const char *>(m);
  const char *s = f.p, *e = f.p + f.n;
  auto line_end = [&](const char *q) {
    const void *nl = memchr(q, '\n', (size_t)(e - q));
    return nl ? static_cast<const char *>(nl) : e;
  };
  // banner
  {
    const char *le = line_end(s);
    static const char *want[5] = {"%%MatrixMarket", "matrix", "coordinate", "real", "general"};
    static const char *what[5] = {"not MM format", "not matrix type", "not coordinate format",
                                  "not real", "not general"};
    const char *q = s;
    for (int k = 0; k < 5; k++) {
      const char *b, *t;
      const bool got = token(q, le, b, t);
      MGX_REQUIRE(got && (size_t)(t - b) == strlen(want[k]) && !memcmp(b, want[k], t - b),
                  "input_mm: %s: %s", path, what[k]);
    }
    s = le < e ? le + 1 : e;
  }
  // comments, then the size line
  while (s < e && *s == '%') {
    const char *le = line_end(s);
    s = le < e ? le + 1 : e;
  }
  MGX_REQUIRE(s < e, "input_mm: %s: no size line", path);
  {
    const char *le = line_end(s);
    const char *q = s;
    uint64_t v[3];
    for (int k = 0; k < 3; k++) {
      const char *b, *t;
      MGX_REQUIRE(token(q, le, b, t), "input_mm: %s: bad size line", path);
      auto r = std::from_chars(b, t, v[k]);
      MGX_REQUIRE(r.ec == std::errc() && r.ptr == t, "input_mm: %s: bad size line", path);
    }
    MGX_REQUIRE(v[1] > 0, "input_mm: Matrix size should be positive");
    MGX_REQUIRE(v[0] < (uint64_t)INT32_MAX && v[1] < (uint64_t)INT32_MAX &&
                    v[2] < (uint64_t)INT32_MAX,
                "input_mm: %s: sizes must fit int32 indices", path);
    out.rows = (int64_t)v[0];
    out.cols = (int64_t)v[1];
    out.nnz = (int64_t)v[2];
    s = le < e ? le + 1 : e;
  }
  const int64_t nnz = out.nnz;
  out.row.resize(nnz);
  out.col.resize(nnz);
  out.val.resize(nnz);
  if (nnz == 0) return;
  // split the body into chunks at line starts; count lines per chunk, then
  // parse lines [0, nnz) into their positions
  const int64_t body = e - s;
  const int nt = (int)std::min<int64_t>(16, std::max<int64_t>(1, body / (1 << 20)));
  std::vector<const char *> cb(nt + 1);
  cb[0] = s;
  cb[nt] = e;
  for (int k = 1; k < nt; k++) {
    const char *q = s + body * k / nt;
    if (q < cb[k - 1]) q = cb[k - 1];
    const char *le = line_end(q > s ? q - 1 : q);
    cb[k] = q == s ? s : (le < e ? le + 1 : e);
    if (cb[k] < cb[k - 1]) cb[k] = cb[k - 1];
  }
  std::vector<int64_t> lines(nt + 1, 0);
  {
    std::vector<std::thread> th;
    for (int k = 0; k < nt; k++)
      th.emplace_back([&, k] {
        int64_t cnt = 0;
        const char *q = cb[k];
        while (q < cb[k + 1]) {
          const char *le = static_cast<const char *>(memchr(q, '\n', (size_t)(cb[k + 1] - q)));
          cnt++;
          if (!le) break;
          q = le + 1;
        }
        lines[k + 1] = cnt;
      });
    for (auto &t : th) t.join();
  }
  for (int k = 0; k < nt; k++) lines[k + 1] += lines[k];
  MGX_REQUIRE(lines[nt] >= nnz, "input_mm: %s: %lld entries announced, %lld lines present",
              path, (long long)nnz, (long long)lines[nt]);
  std::atomic<int64_t> bad(-1);
  {
    std::vector<std::thread> th;
    for (int k = 0; k < nt; k++)
      th.emplace_back([&, k] {
        int64_t li = lines[k];
        const char *q = cb[k];
        while (q < cb[k + 1] && li < nnz) {
          const char *le = static_cast<const char *>(memchr(q, '\n', (size_t)(cb[k + 1] - q)));
          if (!le) le = cb[k + 1];
          const char *b, *t, *p = q;
          uint64_t ix = 0, jx = 0;
          T v = 0;
          bool ok = token(p, le, b, t) && std::from_chars(b, t, ix).ptr == t;
          ok = ok && token(p, le, b, t) && std::from_chars(b, t, jx).ptr == t;
          ok = ok && token(p, le, b, t);
          if (ok) {
            if (*b == '+' && t - b > 1 && b[1] != '-') b++;
            auto r = std::from_chars(b, t, v);
            ok = r.ptr == t && (r.ec == std::errc() || r.ec == std::errc::result_out_of_range);
            if (ok && r.ec != std::errc()) v = parse_range(b, t, v);
          }
          ok = ok && ix >= 1 && jx >= 1 && ix <= (uint64_t)out.rows && jx <= (uint64_t)out.cols;
          if (!ok) {
            int64_t cur = bad.load();
            while ((cur < 0 || li < cur) && !bad.compare_exchange_weak(cur, li)) {
            }
            return;
          }
          out.row[li] = (int32_t)(ix - 1);
          out.col[li] = (int32_t)(jx - 1);
          out.val[li] = v;
          li++;
          q = le + 1;
        }
      });
    for (auto &t : th) t.join();
  }
  MGX_REQUIRE(bad.load() < 0, "input_mm: %s: entry %lld is not \"i j value\" with 1 <= i <= %lld, "
              "1 <= j <= %lld", path, (long long)bad.load() + 1, (long long)out.rows,
              (long long)out.cols);
}

// COO<T>::output_mm (src/utils/IO/IO_coo.cpp:50-66) of COO(A): banner, size
// line, then "i j value" per entry in row order, 1-based, values in
// std::scientific with max_digits10 digits (printf %.17e / %.9e)
template <class T> void write_mm(Csr *A, const char *path) {
  const int64_t rows = A->rows, nnz = A->nnz;
  std::vector<int32_t> rp(rows + 1), ci(nnz);
  std::vector<T> val(nnz);
  MGX_HIP(hipStreamSynchronize(A->ctx->stream));
  MGX_HIP(hipMemcpy(rp.data(), A->rp, (size_t)(rows + 1) * 4, hipMemcpyDeviceToHost));
  if (nnz) {
    MGX_HIP(hipMemcpy(ci.data(), A->ci, (size_t)nnz * 4, hipMemcpyDeviceToHost));
    MGX_HIP(hipMemcpy(val.data(), A->val, (size_t)nnz * sizeof(T), hipMemcpyDeviceToHost));
  }
  const int prec = std::numeric_limits<T>::max_digits10;
  const int nt = (int)std::min<int64_t>(16, std::max<int64_t>(1, rows / 4096));
  std::vector<std::string> part(nt);
  {
    std::vector<std::thread> th;
    for (int k = 0; k < nt; k++)
      th.emplace_back([&, k] {
        const int64_t r0 = rows * k / nt, r1 = rows * (k + 1) / nt;
        std::string &o = part[k];
        o.reserve((size_t)(rp[r1] - rp[r0]) * (prec + 28));
        char buf[96];
        for (int64_t r = r0; r < r1; r++)
          for (int32_t j = rp[r]; j < rp[r + 1]; j++) {
            char *q = std::to_chars(buf, buf + sizeof buf, (int64_t)r + 1).ptr;
            *q++ = ' ';
            q = std::to_chars(q, buf + sizeof buf, (int64_t)ci[j] + 1).ptr;
            *q++ = ' ';
            q = std::to_chars(q, buf + sizeof buf, val[j], std::chars_format::scientific, prec).ptr;
            *q++ = '\n';
            o.append(buf, q);
          }
      });
    for (auto &t : th) t.join();
  }
  FILE *fp = fopen(path, "wb");
  MGX_REQUIRE(fp, "output_mm: cannot open %s for writing", path);
  std::string head = "%%MatrixMarket matrix coordinate real general\n" + std::to_string(rows) +
                     " " + std::to_string(A->cols) + " " + std::to_string(nnz) + "\n";
  bool ok = fwrite(head.data(), 1, head.size(), fp) == head.size();
  for (auto &s : part) ok = ok && fwrite(s.data(), 1, s.size(), fp) == s.size();
  ok = (fclose(fp) == 0) && ok;
  MGX_REQUIRE(ok, "output_mm: writing %s failed", path);
}

}  // namespace

Csr *csr_transpose(Csr *A) {
  return A->dtype == MGX_F64 ? transpose_device<double>(A) : transpose_device<float>(A);
}

namespace {

void register_csr(Ctx *c, Csr *A) {
  std::lock_guard<std::mutex> lk(c->objs_mu);
  c->live_csrs.insert(A);
}

}  // namespace

}  // namespace mgx

using namespace mgx;

extern "C" {

int mgx_csr_from_coo(mgx_ctx *ctx, int64_t rows, int64_t cols, int64_t nnz, const int32_t *row,
                     const int32_t *col, const void *val, mgx_dtype dtype, int sort,
                     mgx_csr **out) {
  return guard([&] {
    MGX_REQUIRE(ctx && out, "mgx_csr_from_coo: NULL argument");
    MGX_REQUIRE(dtype == MGX_F64 || dtype == MGX_F32, "mgx_csr_from_coo: bad dtype");
    MGX_REQUIRE(sort >= 0 && sort <= 2, "mgx_csr_from_coo: sort must be 0, 1 or 2");
    MGX_REQUIRE(rows >= 0 && cols >= 0 && nnz >= 0, "mgx_csr_from_coo: negative size");
    MGX_REQUIRE(rows < INT32_MAX && cols < INT32_MAX && nnz < INT32_MAX,
                "mgx_csr_from_coo: sizes must fit int32 indices");
    MGX_REQUIRE(nnz == 0 || (row && col && val), "mgx_csr_from_coo: NULL arrays");
    MGX_REQUIRE(nnz == 0 || cols > 0, "mgx_csr_from_coo: entries in a matrix without columns");
    activate(ctx);
    Csr *A = dtype == MGX_F64
                 ? csr_from_host_coo<double>(ctx, rows, cols, nnz, row, col,
                                             static_cast<const double *>(val), sort)
                 : csr_from_host_coo<float>(ctx, rows, cols, nnz, row, col,
                                            static_cast<const float *>(val), sort);
    register_csr(ctx, A);
    *out = static_cast<mgx_csr *>(A);
    return MGX_OK;
  });
}

int mgx_csr_write_mm(mgx_csr *A, const char *path) {
  return guard([&] {
    MGX_REQUIRE(A && path, "mgx_csr_write_mm: NULL argument");
    activate(A->ctx);
    if (A->dtype == MGX_F64)
      write_mm<double>(A, path);
    else
      write_mm<float>(A, path);
    return MGX_OK;
  });
}

int mgx_csr_read_mm(mgx_ctx *ctx, const char *path, mgx_dtype dtype, int sort, mgx_csr **out) {
  return guard([&] {
    MGX_REQUIRE(ctx && path && out, "mgx_csr_read_mm: NULL argument");
    MGX_REQUIRE(dtype == MGX_F64 || dtype == MGX_F32, "mgx_csr_read_mm: bad dtype");
    MGX_REQUIRE(sort >= 0 && sort <= 2, "mgx_csr_read_mm: sort must be 0, 1 or 2");
    activate(ctx);
    Csr *A = nullptr;
    if (dtype == MGX_F64) {
      Coo<double> coo;
      read_mm(path, coo);
      A = csr_from_host_coo<double>(ctx, coo.rows, coo.cols, coo.nnz, coo.row.data(),
                                    coo.col.data(), coo.val.data(), sort);
    } else {
      Coo<float> coo;
      read_mm(path, coo);
      A = csr_from_host_coo<float>(ctx, coo.rows, coo.cols, coo.nnz, coo.row.data(),
                                   coo.col.data(), coo.val.data(), sort);
    }
    register_csr(ctx, A);
    *out = static_cast<mgx_csr *>(A);
    return MGX_OK;
  });
}

}  // extern "C"
