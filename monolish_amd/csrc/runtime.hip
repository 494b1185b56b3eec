// runtime.hip -- contexts, device residency of vectors and CRS matrices, and
// the BLAS entry points of the C-ABI (include/monolish_amd.h).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <thread>

#include "mgx_batch.hpp"
#include "mgx_internal.hpp"

namespace mgx {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

double seconds_now() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

void *dev_alloc(size_t bytes) {
  void *p = nullptr;
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    throw Error{MGX_ERR_NOMEM};
  }
  return p;
}

void dev_free(void *p) {
  if (p) (void)hipFree(p);
}

void *ws_get(Ctx *c, size_t bytes) {
  bytes = (bytes + 255) & ~(size_t)255;
  if (bytes == 0) bytes = 256;
  size_t best = (size_t)-1;
  int bi = -1;
  for (size_t i = 0; i < c->ws_free.size(); i++)
    if (c->ws_free[i].second >= bytes && c->ws_free[i].second < best) {
      best = c->ws_free[i].second;
      bi = (int)i;
    }
  if (bi >= 0) {
    auto e = c->ws_free[bi];
    c->ws_free.erase(c->ws_free.begin() + bi);
    c->ws_used.push_back(e);
    return e.first;
  }
  void *p = dev_alloc(bytes);
  c->ws_used.push_back({p, bytes});
  return p;
}

void ws_release_to(Ctx *c, size_t mark) {
  while (c->ws_used.size() > mark) {
    c->ws_free.push_back(c->ws_used.back());
    c->ws_used.pop_back();
  }
}

void ws_release_all(Ctx *c) {
  for (auto &e : c->ws_used) c->ws_free.push_back(e);
  c->ws_used.clear();
}

void ensure_partials(Ctx *c, size_t bytes) {
  if (bytes <= c->partials_bytes) return;
  size_t nb = std::max(bytes, (size_t)1 << 20);
  // the stream may still be reading the old buffer
  MGX_HIP(hipStreamSynchronize(c->stream));
  dev_free(c->d_partials);
  c->d_partials = nullptr;
  c->partials_bytes = 0;
  c->d_partials = dev_alloc(nb);
  c->partials_bytes = nb;
}


// run fn(begin, end) over [0, n) on up to 16 host threads
template <class F> static void parallel_for(int64_t n, F fn) {
  int nt = (int)std::min<int64_t>(16, std::max<int64_t>(1, n / (1 << 20)));
  unsigned hw = std::thread::hardware_concurrency();
  if (hw > 0) nt = std::min<int>(nt, (int)hw);
  if (nt <= 1) {
    fn((int64_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  const int64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; t++) {
    int64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back(fn, b, e);
  }
  for (auto &t : th) t.join();
}

static Vec *V(mgx_vec *v) { return static_cast<Vec *>(v); }
static Csr *M(mgx_csr *a) { return static_cast<Csr *>(a); }

static void check_vec(const mgx_vec *v, const char *what) {
  MGX_REQUIRE(v != nullptr, "%s: NULL vector", what);
}

static void same_kind(const Vec *a, const Vec *b, const char *what) {
  MGX_REQUIRE(a->ctx == b->ctx, "%s: vectors belong to different contexts", what);
  MGX_REQUIRE(a->dtype == b->dtype, "%s: dtype mismatch", what);
  MGX_REQUIRE(a->n == b->n, "%s: size mismatch (%lld vs %lld)", what,
              (long long)a->n, (long long)b->n);
}

static bool overlap(const Vec *a, const Vec *b) {
  const char *a0 = (const char *)a->d, *b0 = (const char *)b->d;
  const char *a1 = a0 + a->n * a->elem(), *b1 = b0 + b->n * b->elem();
  return a->n > 0 && b->n > 0 && a0 < b1 && b0 < a1;
}

// col_ind / val with kEntryPad zeroed entries after the last one: the SpMV
// reads whole 16-byte vectors, which may extend past the final entry
constexpr int64_t kEntryPad = 4;
void alloc_entries(Csr *A, int64_t nnz) {
  const size_t cnt = (size_t)(std::max<int64_t>(nnz, 0) + kEntryPad);
  A->ci = (int32_t *)dev_alloc(cnt * 4);
  A->val = dev_alloc(cnt * A->elem());
  MGX_HIP(hipMemsetAsync(A->ci + nnz, 0, kEntryPad * 4, A->ctx->stream));
  MGX_HIP(hipMemsetAsync((char *)A->val + nnz * A->elem(), 0, kEntryPad * A->elem(),
                         A->ctx->stream));
}

void csr_free(Csr *A) {
  if (!A) return;
  if (A->ctx) (void)hipStreamSynchronize(A->ctx->stream);
  dev_free(A->rp);
  dev_free(A->ci);
  dev_free(A->val);
  dev_free(A->val32);
  dev_free(A->dcorr32);
  dev_free(A->plan.d_blocks);
  spmv_panels_free(A);
  spmv_sell_free(A);
  spmv_engine_free(A);
  spmv_lpanel_free(A);
  tri_free(A);
  ilu_free(A);
  dev_free(A->d_chunks);
  dev_free(A->d_tile_chunk);
  if (A->trans) csr_free(A->trans);
  delete static_cast<mgx_csr *>(A);
}

}  // namespace mgx

using namespace mgx;

extern "C" {

const char *mgx_last_error(void) { return g_err; }

int mgx_version(int *major, int *minor, int *patch) {
  if (major) *major = 0;
  if (minor) *minor = 1;
  if (patch) *patch = 0;
  return MGX_OK;
}

int mgx_device_count(int *count) {
  return guard([&] {
    MGX_REQUIRE(count, "mgx_device_count: NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = e == hipSuccess ? n : 0;
    return MGX_OK;
  });
}

int mgx_init(int device, mgx_ctx **out) {
  return guard([&] {
    MGX_REQUIRE(out, "mgx_init: NULL out");
    int n = 0;
    MGX_HIP(hipGetDeviceCount(&n));
    MGX_REQUIRE(device >= 0 && device < n, "mgx_init: device %d of %d", device, n);
    MGX_HIP(hipSetDevice(device));
    auto *c = new mgx_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    MGX_HIP(hipGetDeviceProperties(&prop, device));
    c->num_cus = prop.multiProcessorCount;
    c->hbm_bytes = prop.totalGlobalMem;
    snprintf(c->name, sizeof c->name, "%s (%s)", prop.name, prop.gcnArchName);
    MGX_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->own_stream = c->stream;
    c->d_state = (SolveState *)dev_alloc(sizeof(SolveState));
    MGX_HIP(hipMemset(c->d_state, 0, sizeof(SolveState)));
    c->d_tick = (uint32_t *)dev_alloc(kTickWords * sizeof(uint32_t));
    MGX_HIP(hipMemset(c->d_tick, 0, kTickWords * sizeof(uint32_t)));
    MGX_HIP(hipHostMalloc((void **)&c->h_pinned, 64 * sizeof(double), hipHostMallocDefault));
    MGX_HIP(hipEventCreateWithFlags(&c->ev, hipEventDisableTiming));
    ensure_partials(c, 1 << 20);
    noalloc_scratch_init(device);
    *out = c;
    return MGX_OK;
  });
}

int mgx_finalize(mgx_ctx *ctx) {
  return guard([&] {
    if (!ctx) return MGX_OK;
    activate(ctx);
    (void)hipStreamSynchronize(ctx->stream);
    ctx_free_engines(ctx);
    {
      std::lock_guard<std::mutex> lk(ctx->objs_mu);
      for (Csr *A : ctx->live_csrs) csr_free(A);
      for (Vec *v : ctx->live_vecs) {
        if (v->owner) dev_free(v->d);
        delete static_cast<mgx_vec *>(v);
      }
      ctx->live_csrs.clear();
      ctx->live_vecs.clear();
    }
    for (auto &e : ctx->ws_used) dev_free(e.first);
    for (auto &e : ctx->ws_free) dev_free(e.first);
    dev_free(ctx->d_partials);
    dev_free(ctx->d_state);
    dev_free(ctx->d_tick);
    if (ctx->h_pinned) (void)hipHostFree(ctx->h_pinned);
    if (ctx->ev) (void)hipEventDestroy(ctx->ev);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return MGX_OK;
  });
}

int mgx_sync(mgx_ctx *ctx) {
  return guard([&] {
    MGX_REQUIRE(ctx, "mgx_sync: NULL ctx");
    MGX_HIP(hipStreamSynchronize(ctx->stream));
    return MGX_OK;
  });
}

void *mgx_stream(mgx_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int mgx_set_stream(mgx_ctx *ctx, void *stream) {
  return guard([&] {
    MGX_REQUIRE(ctx, "mgx_set_stream: NULL ctx");
    activate(ctx);
    // work queued so far finishes before later work starts on the new stream
    MGX_HIP(hipStreamSynchronize(ctx->stream));
    ctx->stream = stream ? static_cast<hipStream_t>(stream) : ctx->own_stream;
    return MGX_OK;
  });
}

int mgx_device_info(mgx_ctx *ctx, char *name, size_t len, int *cus,
                    size_t *hbm) {
  return guard([&] {
    MGX_REQUIRE(ctx, "mgx_device_info: NULL ctx");
    if (name && len) {
      strncpy(name, ctx->name, len - 1);
      name[len - 1] = 0;
    }
    if (cus) *cus = ctx->num_cus;
    if (hbm) *hbm = ctx->hbm_bytes;
    return MGX_OK;
  });
}

int mgx_set_spmv_tuning(mgx_ctx *ctx, int rows_per_block, int nnz_per_block,
                        int long_row, int xcd_remap) {
  return guard([&] {
    MGX_REQUIRE(ctx, "mgx_set_spmv_tuning: NULL ctx");
    MGX_REQUIRE(rows_per_block >= 1 && rows_per_block <= kBlock,
                "rows_per_block must be in [1, %d]", kBlock);
    MGX_REQUIRE(nnz_per_block >= 64 && nnz_per_block <= 8192,
                "nnz_per_block must be in [64, 8192]");
    MGX_REQUIRE(long_row >= 1 && long_row <= nnz_per_block,
                "long_row must be in [1, nnz_per_block]");
    ctx->spmv_rows = rows_per_block;
    ctx->spmv_nnz = nnz_per_block;
    ctx->spmv_long = long_row;
    ctx->xcd_remap = xcd_remap ? 1 : 0;
    return MGX_OK;
  });
}

int mgx_set_graphs(mgx_ctx *ctx, int enable) {
  return guard([&] {
    MGX_REQUIRE(ctx, "mgx_set_graphs: NULL ctx");
    ctx->use_graphs = enable ? 1 : 0;
    return MGX_OK;
  });
}

int mgx_csr_autotune(mgx_csr *A, int reps, mgx_spmv_tuning *out) {
  return guard([&] {
    MGX_REQUIRE(A, "mgx_csr_autotune: NULL matrix");
    MGX_REQUIRE(reps >= 1 && reps <= 10000, "mgx_csr_autotune: reps must be in [1, 10000]");
    activate(A->ctx);
    MGX_HIP(hipStreamSynchronize(A->ctx->stream));
    spmv_autotune(A, reps, out);
    return MGX_OK;
  });
}

int mgx_csr_set_plan(mgx_csr *A, int rows_per_block, int nnz_per_block, int xcd_remap,
                     int row_walk, int column_panels) {
  return guard([&] {
    MGX_REQUIRE(A, "mgx_csr_set_plan: NULL matrix");
    MGX_REQUIRE(rows_per_block >= 1 && rows_per_block <= kBlock,
                "rows_per_block must be in [1, %d]", kBlock);
    MGX_REQUIRE(row_walk >= 0 && row_walk <= 3, "row_walk must be 0, 1, 2 or 3");
    MGX_REQUIRE(xcd_remap >= 0 && xcd_remap <= kMaxRemapRun,
                "xcd_remap must be 0, 1 or a run length in [2, %d]", kMaxRemapRun);
    // the row walk stages 12 bytes per FP64 entry in LDS: 4096 entries fit 64 KiB
    MGX_REQUIRE(nnz_per_block >= 64 && nnz_per_block <= (row_walk == 1 ? 4096 : 8192),
                "nnz_per_block must be in [64, %d]", row_walk == 1 ? 4096 : 8192);
    MGX_REQUIRE(row_walk == 3 ? column_panels >= 1 && column_panels <= 65536
                              : column_panels == 0 || (column_panels >= 2 && column_panels <= 64),
                "column_panels must be 0 or in [2, 64] (in [1, 65536] with row_walk 3)");
    activate(A->ctx);
    MGX_HIP(hipStreamSynchronize(A->ctx->stream));
    std::vector<int32_t> rp(A->rows + 1);
    MGX_HIP(hipMemcpy(rp.data(), A->rp, (size_t)(A->rows + 1) * 4, hipMemcpyDeviceToHost));
    const int long_t = A->plan.long_row > 0 ? A->plan.long_row : A->ctx->spmv_long;
    MGX_REQUIRE(long_t <= nnz_per_block, "nnz_per_block below the long-row threshold %d",
                long_t);
    MGX_REQUIRE(row_walk != 2 || column_panels == 0,
                "mgx_csr_set_plan: the sliced layout and column panels exclude each other");
    spmv_plan_build_with(A, rp.data(), rows_per_block, nnz_per_block, long_t, xcd_remap,
                         row_walk == 1);
    if (row_walk == 3)
      MGX_REQUIRE(spmv_lpanel_build(A, column_panels),
                  "mgx_csr_set_plan: L2 panels need ascending columns within rows and at "
                  "most 2^20 columns per panel");
    else if (column_panels)
      MGX_REQUIRE(spmv_panels_build(A, column_panels),
                  "mgx_csr_set_plan: column panels need ascending columns within rows");
    if (row_walk == 2)
      MGX_REQUIRE(spmv_sell_build(A, rp.data()),
                  "mgx_csr_set_plan: the sliced layout would pad the entries by more than 30 %%");
    return MGX_OK;
  });
}

int mgx_set_persistent(mgx_ctx *ctx, int enable) {
  return guard([&] {
    MGX_REQUIRE(ctx, "mgx_set_persistent: NULL ctx");
    MGX_REQUIRE(enable >= 0, "mgx_set_persistent: negative value");
    ctx->use_persistent = enable;
    return MGX_OK;
  });
}

// ------------------------------------------------------------- vectors ----
int mgx_vec_create(mgx_ctx *ctx, int64_t n, mgx_dtype dtype, const void *host,
                   mgx_vec **out) {
  return guard([&] {
    MGX_REQUIRE(ctx && out, "mgx_vec_create: NULL argument");
    MGX_REQUIRE(n >= 0, "mgx_vec_create: negative size");
    MGX_REQUIRE(dtype == MGX_F64 || dtype == MGX_F32, "mgx_vec_create: bad dtype");
    activate(ctx);
    auto *v = new mgx_vec();
    v->ctx = ctx;
    v->n = n;
    v->dtype = dtype;
    try {
      v->d = dev_alloc((size_t)std::max<int64_t>(n, 1) * v->elem());
      if (n > 0) {
        if (host)
          MGX_HIP(hipMemcpy(v->d, host, (size_t)n * v->elem(), hipMemcpyHostToDevice));
        else
          MGX_HIP(hipMemsetAsync(v->d, 0, (size_t)n * v->elem(), ctx->stream));
      }
      MGX_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
      dev_free(v->d);
      delete v;
      throw;
    }
    {
      std::lock_guard<std::mutex> lk(ctx->objs_mu);
      ctx->live_vecs.insert(v);
    }
    *out = v;
    return MGX_OK;
  });
}

int mgx_vec_view(mgx_vec *base, int64_t offset, int64_t n, mgx_vec **out) {
  return guard([&] {
    MGX_REQUIRE(base && out, "mgx_vec_view: NULL argument");
    MGX_REQUIRE(offset >= 0 && n >= 0 && offset + n <= base->n,
                "mgx_vec_view: [%lld, %lld) outside a vector of %lld",
                (long long)offset, (long long)(offset + n), (long long)base->n);
    auto *v = new mgx_vec();
    v->ctx = base->ctx;
    v->n = n;
    v->dtype = base->dtype;
    v->d = (char *)base->d + offset * (int64_t)base->elem();
    v->owner = false;
    {
      std::lock_guard<std::mutex> lk(v->ctx->objs_mu);
      v->ctx->live_vecs.insert(v);
    }
    *out = v;
    return MGX_OK;
  });
}

int mgx_vec_upload(mgx_vec *v, const void *host) {
  return guard([&] {
    check_vec(v, "mgx_vec_upload");
    MGX_REQUIRE(host || v->n == 0, "mgx_vec_upload: NULL host");
    activate(v->ctx);
    MGX_HIP(hipStreamSynchronize(v->ctx->stream));
    if (v->n) MGX_HIP(hipMemcpy(v->d, host, (size_t)v->n * v->elem(), hipMemcpyHostToDevice));
    return MGX_OK;
  });
}

int mgx_vec_download(mgx_vec *v, void *host) {
  return guard([&] {
    check_vec(v, "mgx_vec_download");
    MGX_REQUIRE(host || v->n == 0, "mgx_vec_download: NULL host");
    activate(v->ctx);
    MGX_HIP(hipStreamSynchronize(v->ctx->stream));
    if (v->n) MGX_HIP(hipMemcpy(host, v->d, (size_t)v->n * v->elem(), hipMemcpyDeviceToHost));
    return MGX_OK;
  });
}

int mgx_vec_destroy(mgx_vec *v) {
  return guard([&] {
    if (!v) return MGX_OK;
    {
      std::lock_guard<std::mutex> lk(v->ctx->objs_mu);
      v->ctx->live_vecs.erase(v);
    }
    activate(v->ctx);
    if (v->owner) {
      (void)hipStreamSynchronize(v->ctx->stream);
      dev_free(v->d);
    }
    delete v;
    return MGX_OK;
  });
}

int mgx_vec_fill(mgx_vec *v, double value) {
  return guard([&] {
    check_vec(v, "mgx_vec_fill");
    activate(v->ctx);
    launch_fill(v->ctx, v->dtype, v->d, v->n, value);
    return MGX_OK;
  });
}

int64_t mgx_vec_size(const mgx_vec *v) { return v ? v->n : -1; }
void *mgx_vec_device_ptr(mgx_vec *v) { return v ? v->d : nullptr; }

// ----------------------------------------------------------------- CRS ----
int mgx_csr_create(mgx_ctx *ctx, int64_t rows, int64_t cols, int64_t nnz,
                   const int32_t *row_ptr, const int32_t *col_ind,
                   const void *val, mgx_dtype dtype, mgx_csr **out) {
  return guard([&] {
    MGX_REQUIRE(ctx && out && row_ptr, "mgx_csr_create: NULL argument");
    MGX_REQUIRE(dtype == MGX_F64 || dtype == MGX_F32, "mgx_csr_create: bad dtype");
    MGX_REQUIRE(rows >= 0 && cols >= 0 && nnz >= 0, "mgx_csr_create: negative size");
    MGX_REQUIRE(rows < INT32_MAX && cols < INT32_MAX && nnz < INT32_MAX,
                "mgx_csr_create: sizes must fit int32 indices");
    MGX_REQUIRE(nnz == 0 || (col_ind && val), "mgx_csr_create: NULL col_ind/val");
    activate(ctx);
    // effective row_ptr: non-decreasing; a drop is accepted only as the
    // all-zero tail CRS(COO&) leaves after the last non-empty row
    std::vector<int32_t> rp(row_ptr, row_ptr + rows + 1);
    bool raw_differs = false;
    for (int64_t i = 0; i <= rows; i++)
      MGX_REQUIRE(rp[i] >= 0 && rp[i] <= nnz, "row_ptr[%lld] = %d outside [0, %lld]",
                  (long long)i, rp[i], (long long)nnz);
    for (int64_t i = 0; i < rows; i++) {
      if (rp[i + 1] < rp[i]) {
        for (int64_t k = i + 1; k <= rows; k++)
          MGX_REQUIRE(row_ptr[k] == 0, "row_ptr decreases at row %lld", (long long)i);
        for (int64_t k = i + 1; k <= rows; k++) rp[k] = rp[i];
        raw_differs = true;
        break;
      }
    }
    const int64_t used = rp[rows];
    std::atomic<int64_t> bad_at(-1);
    if (used > 0) {
      parallel_for(used, [&](int64_t b, int64_t e) {
        for (int64_t j = b; j < e; j++)
          if (col_ind[j] < 0 || col_ind[j] >= cols) {
            bad_at.store(j);
            return;
          }
      });
    }
    MGX_REQUIRE(bad_at.load() < 0, "col_ind[%lld] outside [0, %lld)",
                (long long)bad_at.load(), (long long)cols);
    auto *A = new mgx_csr();
    A->ctx = ctx;
    A->rows = rows;
    A->cols = cols;
    A->nnz = nnz;
    A->dtype = dtype;
    try {
      A->rp = (int32_t *)dev_alloc((size_t)(rows + 1) * 4);
      alloc_entries(A, nnz);
      MGX_HIP(hipMemcpy(A->rp, rp.data(), (size_t)(rows + 1) * 4, hipMemcpyHostToDevice));
      if (nnz > 0) {
        MGX_HIP(hipMemcpy(A->ci, col_ind, (size_t)nnz * 4, hipMemcpyHostToDevice));
        MGX_HIP(hipMemcpy(A->val, val, (size_t)nnz * A->elem(), hipMemcpyHostToDevice));
      }
      if (raw_differs) A->rp_raw.assign(row_ptr, row_ptr + rows + 1);
      A->gather_lines = sample_gather_lines(col_ind, used, A->elem());
      spmv_plan_build(ctx, A, rp.data());
      if (rows == cols) fused_plan_build(A, rp.data());
    } catch (...) {
      csr_free(A);
      throw;
    }
    {
      std::lock_guard<std::mutex> lk(ctx->objs_mu);
      ctx->live_csrs.insert(A);
    }
    *out = A;
    return MGX_OK;
  });
}

int mgx_csr_update_values(mgx_csr *A, const void *val) {
  return guard([&] {
    MGX_REQUIRE(A && (val || A->nnz == 0), "mgx_csr_update_values: NULL argument");
    activate(A->ctx);
    MGX_HIP(hipStreamSynchronize(A->ctx->stream));
    if (A->nnz) MGX_HIP(hipMemcpy(A->val, val, (size_t)A->nnz * A->elem(), hipMemcpyHostToDevice));
    if (A->val32) {
      launch_convert(A->ctx, A->dtype, MGX_F32, A->nnz, A->val, A->val32);
      diag_corr_build(A);
    }
    spmv_panels_free(A);  // their value copies would be stale
    spmv_lpanel_free(A);
    spmv_sell_values(A);
    spmv_engine_values(A);
    if (A->trans) {
      csr_free(A->trans);
      A->trans = nullptr;
    }
    MGX_HIP(hipStreamSynchronize(A->ctx->stream));
    return MGX_OK;
  });
}

int mgx_csr_download(mgx_csr *A, int32_t *row_ptr, int32_t *col_ind, void *val) {
  return guard([&] {
    MGX_REQUIRE(A, "mgx_csr_download: NULL matrix");
    activate(A->ctx);
    MGX_HIP(hipStreamSynchronize(A->ctx->stream));
    if (row_ptr) {
      if (!A->rp_raw.empty())
        std::memcpy(row_ptr, A->rp_raw.data(), A->rp_raw.size() * 4);
      else
        MGX_HIP(hipMemcpy(row_ptr, A->rp, (size_t)(A->rows + 1) * 4, hipMemcpyDeviceToHost));
    }
    if (col_ind && A->nnz) MGX_HIP(hipMemcpy(col_ind, A->ci, (size_t)A->nnz * 4, hipMemcpyDeviceToHost));
    if (val && A->nnz) MGX_HIP(hipMemcpy(val, A->val, (size_t)A->nnz * A->elem(), hipMemcpyDeviceToHost));
    return MGX_OK;
  });
}

int mgx_csr_destroy(mgx_csr *A) {
  return guard([&] {
    if (!A) return MGX_OK;
    {
      std::lock_guard<std::mutex> lk(A->ctx->objs_mu);
      A->ctx->live_csrs.erase(A);
    }
    activate(A->ctx);
    csr_free(A);
    return MGX_OK;
  });
}

int mgx_csr_diag(mgx_csr *A, mgx_vec *d) {
  return guard([&] {
    MGX_REQUIRE(A && d, "mgx_csr_diag: NULL argument");
    MGX_REQUIRE(d->n == std::min(A->rows, A->cols), "mgx_csr_diag: size mismatch");
    MGX_REQUIRE(d->dtype == A->dtype && d->ctx == A->ctx, "mgx_csr_diag: dtype/context mismatch");
    activate(A->ctx);
    launch_diag(A, d->d);
    return MGX_OK;
  });
}

int mgx_csr_stats_get(mgx_csr *A, mgx_csr_stats *s) {
  return guard([&] {
    MGX_REQUIRE(A && s, "mgx_csr_stats_get: NULL argument");
    s->rows = A->rows;
    s->cols = A->cols;
    s->nnz = A->nnz;
    s->max_row = A->plan.max_row;
    s->min_row = A->plan.min_row;
    s->mean_row = A->plan.mean_row;
    s->stream_blocks = A->plan.nblocks - A->plan.nlong;
    s->long_rows = A->plan.nlong;
    s->gather_lines = A->gather_lines;
    return MGX_OK;
  });
}

// ---------------------------------------------------------------- BLAS ----
static void matvec_checks(Csr *A, Vec *x, Vec *y, bool trans, const char *w) {
  MGX_REQUIRE(A && x && y, "%s: NULL argument", w);
  MGX_REQUIRE(x->ctx == A->ctx && y->ctx == A->ctx, "%s: context mismatch", w);
  MGX_REQUIRE(x->dtype == A->dtype && y->dtype == A->dtype, "%s: dtype mismatch", w);
  MGX_REQUIRE(x->n == (trans ? A->rows : A->cols) && y->n == (trans ? A->cols : A->rows),
              "%s: size mismatch (A %lldx%lld, x %lld, y %lld)", w, (long long)A->rows,
              (long long)A->cols, (long long)x->n, (long long)y->n);
  MGX_REQUIRE(!overlap(x, y), "%s: x and y overlap", w);
}

int mgx_matvec(mgx_csr *A, mgx_vec *x, mgx_vec *y) {
  return guard([&] {
    matvec_checks(A, x, y, false, "mgx_matvec");
    activate(A->ctx);
    launch_spmv(A, A->val, A->dtype, x->d, y->d, 1.0, 0.0, false);
    return MGX_OK;
  });
}

int mgx_matvec_ab(double alpha, mgx_csr *A, mgx_vec *x, double beta, mgx_vec *y) {
  return guard([&] {
    matvec_checks(A, x, y, false, "mgx_matvec_ab");
    activate(A->ctx);
    launch_spmv(A, A->val, A->dtype, x->d, y->d, alpha, beta, true);
    return MGX_OK;
  });
}

int mgx_matvec_t(mgx_csr *A, mgx_vec *x, mgx_vec *y) {
  return guard([&] {
    matvec_checks(A, x, y, true, "mgx_matvec_t");
    activate(A->ctx);
    if (!A->trans) A->trans = csr_transpose(A);
    launch_spmv(A->trans, A->trans->val, A->dtype, x->d, y->d, 1.0, 0.0, false);
    return MGX_OK;
  });
}

int mgx_dot(mgx_vec *x, mgx_vec *y, double *out) {
  return guard([&] {
    MGX_REQUIRE(x && y && out, "mgx_dot: NULL argument");
    same_kind(x, y, "mgx_dot");
    activate(x->ctx);
    *out = dot_host(x->ctx, x->dtype, x->n, x->d, y->d);
    return MGX_OK;
  });
}

int mgx_dot_device(mgx_vec *x, mgx_vec *y, double *d_out) {
  return guard([&] {
    MGX_REQUIRE(x && y && d_out, "mgx_dot_device: NULL argument");
    same_kind(x, y, "mgx_dot_device");
    activate(x->ctx);
    launch_dot(x->ctx, x->dtype, x->n, x->d, y->d, d_out, x->ctx->d_tick);
    return MGX_OK;
  });
}

int mgx_nrm2_device(mgx_vec *x, double *d_out) {
  return guard([&] {
    MGX_REQUIRE(x && d_out, "mgx_nrm2_device: NULL argument");
    activate(x->ctx);
    launch_dot(x->ctx, x->dtype, x->n, x->d, x->d, d_out, x->ctx->d_tick, true);
    return MGX_OK;
  });
}

int mgx_nrm2(mgx_vec *x, double *out) {
  return guard([&] {
    MGX_REQUIRE(x && out, "mgx_nrm2: NULL argument");
    activate(x->ctx);
    double d = dot_host(x->ctx, x->dtype, x->n, x->d, x->d);
    *out = x->dtype == MGX_F64 ? std::sqrt(d) : (double)std::sqrt((float)d);
    return MGX_OK;
  });
}

int mgx_axpy(double a, mgx_vec *x, mgx_vec *y) {
  return guard([&] {
    MGX_REQUIRE(x && y, "mgx_axpy: NULL argument");
    same_kind(x, y, "mgx_axpy");
    activate(x->ctx);
    launch_axpy(x->ctx, x->dtype, x->n, a, x->d, y->d);
    return MGX_OK;
  });
}

int mgx_xpay(double a, mgx_vec *x, mgx_vec *y) {
  return guard([&] {
    MGX_REQUIRE(x && y, "mgx_xpay: NULL argument");
    same_kind(x, y, "mgx_xpay");
    activate(x->ctx);
    launch_xpay(x->ctx, x->dtype, x->n, a, x->d, y->d);
    return MGX_OK;
  });
}

int mgx_scal(double a, mgx_vec *x) {
  return guard([&] {
    MGX_REQUIRE(x, "mgx_scal: NULL argument");
    activate(x->ctx);
    launch_scal(x->ctx, x->dtype, x->n, a, x->d);
    return MGX_OK;
  });
}

int mgx_copy(mgx_vec *x, mgx_vec *y) {
  return guard([&] {
    MGX_REQUIRE(x && y, "mgx_copy: NULL argument");
    same_kind(x, y, "mgx_copy");
    activate(x->ctx);
    launch_copy(x->ctx, x->dtype, x->n, x->d, y->d);
    return MGX_OK;
  });
}

int mgx_axpyz(double a, mgx_vec *x, mgx_vec *y, mgx_vec *z) {
  return guard([&] {
    MGX_REQUIRE(x && y && z, "mgx_axpyz: NULL argument");
    same_kind(x, y, "mgx_axpyz");
    same_kind(x, z, "mgx_axpyz");
    activate(x->ctx);
    launch_axpyz(x->ctx, x->dtype, x->n, a, x->d, y->d, z->d);
    return MGX_OK;
  });
}

static int vv(int op, mgx_vec *a, mgx_vec *b, mgx_vec *y, const char *w) {
  return guard([&] {
    MGX_REQUIRE(a && b && y, "%s: NULL argument", w);
    same_kind(a, b, w);
    same_kind(a, y, w);
    activate(a->ctx);
    launch_vv(a->ctx, a->dtype, op, a->n, a->d, b->d, y->d);
    return MGX_OK;
  });
}

int mgx_vadd(mgx_vec *a, mgx_vec *b, mgx_vec *y) { return vv(0, a, b, y, "mgx_vadd"); }
int mgx_vsub(mgx_vec *a, mgx_vec *b, mgx_vec *y) { return vv(1, a, b, y, "mgx_vsub"); }
int mgx_vmul(mgx_vec *a, mgx_vec *b, mgx_vec *y) { return vv(2, a, b, y, "mgx_vmul"); }

int mgx_vreciprocal(mgx_vec *a, mgx_vec *y) {
  return guard([&] {
    MGX_REQUIRE(a && y, "mgx_vreciprocal: NULL argument");
    same_kind(a, y, "mgx_vreciprocal");
    activate(a->ctx);
    launch_reciprocal(a->ctx, a->dtype, a->n, a->d, y->d);
    return MGX_OK;
  });
}

}  // extern "C"
