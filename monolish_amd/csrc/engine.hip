// engine.hip -- host side of the multi-system fused engine: device memory
// pool, system admission (set-up kernels), iteration launches and polling.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>

#include "mgx_batch.hpp"

namespace mgx {

// ------------------------------------------------------------ DevicePool --
size_t DevicePool::round(size_t bytes) {
  if (bytes <= 4096) return 4096;
  // 1/8 steps between powers of two keep waste under 12.5%
  size_t p = 4096;
  while (p < bytes) p <<= 1;
  const size_t step = p >> 4;
  return ((bytes + step - 1) / step) * step;
}

DevicePool::~DevicePool() {
  for (auto &f : free_)
    for (void *p : f.second)
      if (!carved_.count(p)) dev_free(p);
  for (auto &l : live_)
    if (!carved_.count(l.first)) dev_free(l.first);
  for (void *s : slabs_) dev_free(s);
}

bool DevicePool::reserve(const std::vector<size_t> &bytes) {
  // what the free lists already hold, per rounded size
  std::unordered_map<size_t, size_t> have, want;
  for (size_t b : bytes) want[round(b)]++;
  size_t total = 0;
  for (auto &w : want) {
    auto it = free_.find(w.first);
    const size_t h = it == free_.end() ? 0 : it->second.size();
    if (w.second > h) total += (w.second - h) * w.first;
  }
  if (total == 0) return true;
  if (in_use_ + total > budget_) return false;
  void *slab = nullptr;
  if (hipMalloc(&slab, total) != hipSuccess) return false;
  slabs_.push_back(slab);
  reserved_ += total;
  char *p = static_cast<char *>(slab);
  for (auto &w : want) {
    auto &fl = free_[w.first];
    const size_t h = fl.size();
    for (size_t k = h; k < w.second; k++) {
      fl.push_back(p);
      carved_.insert(p);
      p += w.first;
    }
  }
  return true;
}

// free blocks are kept per rounded size; lookups and returns are O(1)
void *DevicePool::get(size_t bytes) {
  const size_t sz = round(bytes);
  auto it = free_.find(sz);
  if (it != free_.end() && !it->second.empty()) {
    void *p = it->second.back();
    it->second.pop_back();
    live_.emplace(p, sz);
    in_use_ += sz;
    return p;
  }
  if (in_use_ + sz > budget_) return nullptr;
  // drop cached blocks of other sizes while they keep us over budget
  for (auto f = free_.begin(); reserved_ + sz > budget_ && f != free_.end(); ++f) {
    auto &fl = f->second;
    for (size_t k = fl.size(); k-- > 0 && reserved_ + sz > budget_;) {
      if (carved_.count(fl[k])) continue;  // slab blocks go with their slab
      dev_free(fl[k]);
      fl.erase(fl.begin() + (std::ptrdiff_t)k);
      reserved_ -= f->first;
    }
  }
  void *p = nullptr;
  if (hipMalloc(&p, sz) != hipSuccess) return nullptr;
  reserved_ += sz;
  live_.emplace(p, sz);
  in_use_ += sz;
  return p;
}

void DevicePool::put(void *p) {
  if (!p) return;
  auto it = live_.find(p);
  if (it == live_.end()) return;
  free_[it->second].push_back(p);
  in_use_ -= it->second;
  live_.erase(it);
}

// ---------------------------------------------------------------- Engine --
template <class T> struct DTof;
template <> struct DTof<double> { static constexpr mgx_dtype v = MGX_F64; };
template <> struct DTof<float> { static constexpr mgx_dtype v = MGX_F32; };

template <class T>
Engine<T>::Engine(Ctx *c, DevicePool *pool, int capacity) : c_(c), pool_(pool), cap_(capacity) {
  d_sys_ = (SysDesc<T> *)dev_alloc(sizeof(SysDesc<T>) * cap_);
  d_states_ = (SolveState *)dev_alloc(sizeof(SolveState) * cap_);
  d_stopped_ = (int32_t *)dev_alloc(sizeof(int32_t));
  MGX_HIP(hipMemset(d_stopped_, 0, sizeof(int32_t)));
  d_ticks_ = (uint32_t *)dev_alloc(sizeof(uint32_t) * kTickWords * cap_);
  MGX_HIP(hipMemset(d_ticks_, 0, sizeof(uint32_t) * kTickWords * cap_));
  rel_.assign(cap_, 0);
  h_sys_.resize(cap_);
  method_.assign(cap_, kMethodCG);
  bufs_.resize(cap_);
  h_states_.resize(cap_);
  reported_.assign(cap_, 0);
  csr_.assign(cap_, nullptr);
  val_.assign(cap_, nullptr);
  for (int s = cap_ - 1; s >= 0; s--) free_slots_.push_back(s);
}

template <class T> Engine<T>::~Engine() {
  (void)hipStreamSynchronize(c_->stream);
  if (gexec_) (void)hipGraphExecDestroy(gexec_);
  for (int s : active_)
    for (void *p : bufs_[s]) pool_->put(p);
  dev_free(d_sys_);
  dev_free(d_states_);
  dev_free(d_stopped_);
  dev_free(d_tiles_[0]);
  dev_free(d_tiles_[1]);
  dev_free(d_gs_);
  dev_free(d_ticks_);
  dev_free(d_stage_);
  dev_free(d_ptiles_);
  dev_free(d_split_slots_);
}

template <class T> size_t Engine<T>::bytes_needed(const Csr *A, int method) const {
  const int64_t n = std::max<int64_t>(A->rows, 1);
  const size_t v = DevicePool::round((size_t)n * sizeof(T));
  const int nvec = method == kMethodCG ? 6 : 11;
  return nvec * v + DevicePool::round(((size_t)8 * ntiles_of(n) + 2 * kRedGroups) * sizeof(T));
}

template <class T>
void Engine<T>::sizes_of(const Csr *A, const SolveCfg &cfg, std::vector<size_t> &out) {
  const int64_t n = A->rows;
  const size_t vb = (size_t)n * sizeof(T);
  const int nvec = 3 + (cfg.precond == MGX_PRECOND_JACOBI ? 1 : 0) +
                   (cfg.method == kMethodCG ? 2 : 7);
  for (int k = 0; k < nvec; k++) out.push_back(vb);
  out.push_back((size_t)(8 * ntiles_of(n) + 2 * kRedGroups) * sizeof(T));
}

template <class T>
int Engine<T>::add(Csr *A, const void *val, T *x, const T *b, const SolveCfg &cfg) {
  MGX_REQUIRE(A->rows == A->cols && A->rows > 0, "engine: square, non-empty system required");
  if (free_slots_.empty()) return -1;
  fused_plan_build(A);  // normally built at upload
  const int64_t n = A->rows, nt = ntiles_of(n);
  const int slot = free_slots_.back();
  std::vector<void *> got;
  auto take = [&](size_t bytes) -> T * {
    void *p = pool_->get(bytes);
    if (p) got.push_back(p);
    return (T *)p;
  };
  const size_t vb = (size_t)n * sizeof(T);
  SysDesc<T> d;
  std::memset(&d, 0, sizeof d);
  bool ok = true;
  auto need = [&](T *&dst) {
    dst = take(vb);
    ok = ok && dst != nullptr;
  };
  need(d.r);
  need(d.pa);
  need(d.pb);
  if (cfg.precond == MGX_PRECOND_JACOBI) {
    T *m = nullptr;
    need(m);
    d.M = m;
  }
  if (cfg.method == kMethodCG) {
    need(d.z);
    need(d.q);
  } else {
    need(d.r0);
    need(d.s);
    need(d.shat);
    need(d.phat);
    need(d.t);
    need(d.va);
    need(d.vb);
  }
  // partials: per quarter for the SpMV stages (CG 1 sum, BiCGSTAB 2), per
  // tile for the updates (2 sums), per quarter for the persistent CG (2
  // sums); then the stage-2 lane sums
  const int64_t np = 8 * nt;
  d.P = take((size_t)(np + 2 * kRedGroups) * sizeof(T));
  ok = ok && d.P != nullptr;
  if (ok) d.S = d.P + np;
  if (!ok) {
    for (void *p : got) pool_->put(p);
    return -1;
  }
  free_slots_.pop_back();
  bufs_[slot] = got;
  method_[slot] = cfg.method;
  reported_[slot] = 0;
  d.rp = A->rp;
  d.ci = A->ci;
  d.val = static_cast<const T *>(val);
  d.chunks = A->d_chunks;
  d.tile_chunk = A->d_tile_chunk;
  d.x = x;
  d.b = b;
  d.st = d_states_ + slot;
  d.tick = d_ticks_ + (size_t)slot * kTickWords;
  d.rhist = cfg.rhist;
  d.rhist_len = cfg.rhist ? cfg.rhist_len : 0;
  d.row0 = 0;
  d.n = n;
  d.ntiles = nt;
  // A lone large CG system runs the split iteration (its own tuned SpMV
  // plan); in multi-system launches (gather_once) a large system joins the
  // shared tile launches instead of adding its own SpMV launch per iteration.
  d.split = cfg.method == kMethodCG && n >= kSplitRows && !cfg.gather_once ? 1 : 0;
  d.once = cfg.method == kMethodBiCGSTAB ||
                   (!d.split && (cfg.gather_once || A->gather_lines > kScatteredLines))
               ? 1
               : 0;
  d.pq_fused = d.split && spmv_sell_pq_ok(A, val, DTof<T>::v) ? 1 : 0;
  d.acc64 = cfg.acc64 && d.pq_fused && sizeof(T) == 4 ? 1 : 0;
  d.dcorr = d.acc64 ? static_cast<const T *>(cfg.dcorr) : nullptr;
  // the fused SpMV stages walk a sliced copy of the matrix when it has one
  // or qualifies for the engine's own (every row at most kEngineSellRow
  // entries, slices padded by at most 30 %)
  if (const SpmvPlan::Sliced *sl = !d.split ? spmv_sell_layout(A, val, DTof<T>::v) : nullptr) {
    d.sell = 1;
    d.slen = sl->d_len;
    d.soff = sl->d_soff;
    d.sci = sl->d_sci;
    d.sval = static_cast<const T *>(sl->d_val);
  }
  d.lf = cfg.method == kMethodCG && !d.split && nt <= kRedGroups && !c_->use_persistent ? 1 : 0;
  d.vec = ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(b)) & 15) == 0;
  d.method = cfg.method;
  d.tol_scale = cfg.tol_relative ? cfg.tol : 0.0;
  csr_[slot] = A;
  val_[slot] = val;
  lds_entries_ = std::max(lds_entries_, A->fused_nnz);
  SolveState &hs = h_states_[slot];
  std::memset(&hs, 0, sizeof hs);
  hs.tol = cfg.tol;
  hs.miniter = cfg.miniter;
  hs.maxiter = cfg.maxiter;
  hs.status = kStatRunning;
  hs.first = 1;
  hs.fresh = 1;
  hs.ia = hs.ib = 0;
  h_sys_[slot] = d;
  rel_[slot] = cfg.tol_relative ? 1 : 0;
  pending_.push_back(slot);
  active_.push_back(slot);
  host_running_++;
  dirty_ = true;
  listed_ok_ = false;
  return slot;
}

template <class T> void Engine<T>::flush_pending() {
  if (pending_.empty()) return;
  // the staging buffers of the previous flush may still be in use
  MGX_HIP(hipStreamSynchronize(c_->stream));
  const int np = (int)pending_.size();
  const size_t sd = sizeof(SysDesc<T>) * np, ss = sizeof(SolveState) * np,
               sl = sizeof(int32_t) * np;
  const size_t bytes = sd + ss + sl;
  h_stage_.resize(bytes);
  std::vector<TileRef> tl;
  bool any_m = false, any_rel = false, any_bcg = false;
  for (int i = 0; i < np; i++) {
    const int slot = pending_[i];
    const SysDesc<T> &d = h_sys_[slot];
    std::memcpy(h_stage_.data() + sizeof(SysDesc<T>) * i, &d, sizeof d);
    std::memcpy(h_stage_.data() + sd + sizeof(SolveState) * i, &h_states_[slot],
                sizeof(SolveState));
    const int32_t s32 = slot;
    std::memcpy(h_stage_.data() + sd + ss + sizeof(int32_t) * i, &s32, sizeof s32);
    for (int64_t k = 0; k < d.ntiles; k++) tl.push_back(TileRef{slot, (int32_t)k});
    any_m = any_m || d.M != nullptr;
    any_rel = any_rel || rel_[slot];
    any_bcg = any_bcg || method_[slot] == kMethodBiCGSTAB;
  }
  if (bytes > stage_cap_) {
    dev_free(d_stage_);
    d_stage_ = nullptr;
    stage_cap_ = std::max(bytes, 2 * stage_cap_);
    d_stage_ = dev_alloc(stage_cap_);
  }
  const int64_t nt = (int64_t)tl.size();
  if (nt > ptiles_cap_) {
    dev_free(d_ptiles_);
    d_ptiles_ = nullptr;
    ptiles_cap_ = std::max<int64_t>(nt, 2 * ptiles_cap_);
    d_ptiles_ = (TileRef *)dev_alloc(sizeof(TileRef) * ptiles_cap_);
  }
  MGX_HIP(hipMemcpyAsync(d_stage_, h_stage_.data(), bytes, hipMemcpyHostToDevice, c_->stream));
  MGX_HIP(hipMemcpyAsync(d_ptiles_, tl.data(), sizeof(TileRef) * nt, hipMemcpyHostToDevice,
                         c_->stream));
  char *base = static_cast<char *>(d_stage_);
  batch_launch_setup<T>(c_, reinterpret_cast<const SysDesc<T> *>(base),
                        reinterpret_cast<const SolveState *>(base + sd),
                        reinterpret_cast<const int32_t *>(base + sd + ss), np, d_sys_, d_states_,
                        d_ptiles_, nt, any_m, any_rel, any_bcg, lds_entries_, d_stopped_);
  // tl is a local: the copies above must finish before it goes away
  MGX_HIP(hipStreamSynchronize(c_->stream));
  pending_.clear();
}

template <class T> void Engine<T>::rebuild_lists() {
  std::vector<TileRef> lst[2];
  any_fused_ = false;
  any_once_ = false;
  any_bcg_once_ = false;
  split_slots_.clear();
  for (int s : active_) {
    if (reported_[s]) continue;
    const int m = method_[s];
    if (m == kMethodCG) {
      if (h_sys_[s].split) {
        // launched one system at a time (batch_launch_split), not listed
        split_slots_.push_back(s);
        continue;
      }
      any_fused_ = true;
      any_once_ = any_once_ || h_sys_[s].once;
    } else {
      any_bcg_once_ = any_bcg_once_ || h_sys_[s].once;
    }
    for (int64_t k = 0; k < h_sys_[s].ntiles; k++) lst[m].push_back(TileRef{s, (int32_t)k});
  }
  MGX_HIP(hipStreamSynchronize(c_->stream));
  for (int m = 0; m < 2; m++) {
    ntiles_[m] = (int64_t)lst[m].size();
    if (ntiles_[m] > tiles_cap_[m]) {
      dev_free(d_tiles_[m]);
      d_tiles_[m] = nullptr;
      tiles_cap_[m] = std::max<int64_t>(ntiles_[m], 2 * tiles_cap_[m]);
      d_tiles_[m] = (TileRef *)dev_alloc(sizeof(TileRef) * tiles_cap_[m]);
    }
    if (ntiles_[m])
      MGX_HIP(hipMemcpy(d_tiles_[m], lst[m].data(), sizeof(TileRef) * ntiles_[m],
                        hipMemcpyHostToDevice));
  }
  const int64_t ns = (int64_t)split_slots_.size();
  if (ns > split_cap_) {
    dev_free(d_split_slots_);
    d_split_slots_ = nullptr;
    split_cap_ = std::max<int64_t>(ns, 2 * split_cap_);
    d_split_slots_ = (int32_t *)dev_alloc(sizeof(int32_t) * split_cap_);
  }
  if (ns) {
    std::vector<int32_t> sl(split_slots_.begin(), split_slots_.end());
    MGX_HIP(hipMemcpy(d_split_slots_, sl.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice));
  }
  dirty_ = false;
}

template <class T> void Engine<T>::launch_iters(int iters) {
  for (int i = 0; i < iters; i++) {
    if (!split_slots_.empty()) {
      // split form, system by system: p = z + beta p; q = A p with <p, q>
      for (int s : split_slots_) batch_launch_split<T>(c_, h_sys_[s], 0);
      for (int s : split_slots_) {
        const SysDesc<T> &h = h_sys_[s];
        if (h.pq_fused) {
          // the sliced layout was there when the system was added, so the
          // fused SpMV forms the <p, q> partials
          MGX_REQUIRE(spmv_sell_pq(csr_[s], val_[s], DTof<T>::v, h.pa, h.q, h.P,
                                   &d_states_[s].status, h.acc64 != 0,
                                   reinterpret_cast<const float *>(h.dcorr)),
                      "solver: the matrix's SpMV plan changed during a solve");
        } else {
          launch_spmv(csr_[s], val_[s], DTof<T>::v, h.pa, h.q, 1.0, 0.0, false,
                      &d_states_[s].status);
          batch_launch_split<T>(c_, h, 1);
        }
      }
      batch_launch_cg_finish<T>(c_, d_sys_, d_split_slots_, (int)split_slots_.size(), 0,
                                d_stopped_);
      for (int s : split_slots_) batch_launch_split<T>(c_, h_sys_[s], 2);
      batch_launch_cg_finish<T>(c_, d_sys_, d_split_slots_, (int)split_slots_.size(), 1,
                                d_stopped_);
    }
    if (ntiles_[0]) {
      // p = z + beta p of the gather-once systems
      if (any_once_) batch_launch_cg_p<T>(c_, d_sys_, d_tiles_[0], ntiles_[0], d_stopped_);
      batch_launch_cg_iter<T>(c_, d_sys_, d_tiles_[0], ntiles_[0], d_stopped_, lds_entries_,
                              any_fused_);
    }
    if (ntiles_[1])
      batch_launch_bicgstab_iter<T>(c_, d_sys_, d_tiles_[1], ntiles_[1], d_stopped_,
                                    lds_entries_, any_bcg_once_);
  }
}

template <class T> void Engine<T>::step(int iters) {
  // the kernels read the sliced copies directly: they must still be the
  // matrices' own (a plan change frees them)
  for (int s : active_) {
    const SysDesc<T> &h = h_sys_[s];
    if (!h.sell || reported_[s]) continue;
    const SpmvPlan::Sliced *sl = spmv_sell_find(csr_[s], val_[s], DTof<T>::v);
    MGX_REQUIRE(sl && sl->d_soff == h.soff && sl->d_sci == h.sci && sl->d_val == h.sval &&
                    sl->d_len == h.slen,
                "solver: the matrix's SpMV plan changed during a solve");
  }
  flush_pending();
  if (dirty_) {
    // a slot that stopped and was resumed before the next step leaves the
    // lists as they were
    std::vector<int> run;
    for (int s : active_)
      if (!reported_[s]) run.push_back(s);
    if (listed_ok_ && run == listed_) {
      dirty_ = false;
    } else {
      rebuild_lists();
      listed_ = run;
      listed_ok_ = true;
      gstale_ = true;
    }
  }
  if (!c_->use_graphs) {
    launch_iters(iters);
    MGX_HIP(hipGetLastError());
    return;
  }
  // kernels only (no memsets) are captured; re-capture when the tile lists
  // or the group size change
  if (!gexec_ || gstale_ || giters_ != iters) {
    if (gexec_) {
      MGX_HIP(hipGraphExecDestroy(gexec_));
      gexec_ = nullptr;
    }
    hipGraph_t g = nullptr;
    MGX_HIP(hipStreamBeginCapture(c_->stream, hipStreamCaptureModeThreadLocal));
    launch_iters(iters);
    hipError_t le = hipGetLastError();
    hipError_t ce = hipStreamEndCapture(c_->stream, &g);
    MGX_HIP(le);
    MGX_HIP(ce);
    hipError_t ie = hipGraphInstantiate(&gexec_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    MGX_HIP(ie);
    giters_ = iters;
    gstale_ = false;
  }
  MGX_HIP(hipGraphLaunch(gexec_, c_->stream));
}

template <class T> bool Engine<T>::persistent_ok(int slot) const {
  if (!c_->use_persistent || active_.size() != 1 || active_[0] != slot) return false;
  const SysDesc<T> &d = h_sys_[slot];
  return method_[slot] == kMethodCG && !d.split && !d.once && !d.lf;
}

template <class T> void Engine<T>::run_persistent(int slot) {
  MGX_REQUIRE(persistent_ok(slot), "engine: persistent solve needs a single CG system");
  flush_pending();
  if (!d_gs_) d_gs_ = (GridSync *)dev_alloc(sizeof(GridSync));
  const int64_t nq = 4 * h_sys_[slot].ntiles;
  const int cap = batch_cg_persist_blocks<T>(c_, lds_entries_);
  MGX_REQUIRE(cap > 0, "engine: persistent kernel does not fit on the device");
  int64_t grid = std::min<int64_t>(nq, cap);
  if (c_->use_persistent > 1) grid = std::min<int64_t>(grid, c_->use_persistent);
  // each launch runs at most this many iterations, then the host looks at
  // the state and launches again if the solve is still running
  const int per_launch = 4096;
  for (;;) {
    MGX_HIP(hipMemsetAsync(d_gs_, 0, sizeof(GridSync), c_->stream));
    batch_launch_cg_persist<T>(c_, d_sys_ + slot, d_gs_, (int)grid, d_stopped_, lds_entries_,
                               per_launch);
    uint32_t err = 0;
    MGX_HIP(hipMemcpyAsync(c_->h_pinned, &d_gs_->err, sizeof(uint32_t), hipMemcpyDeviceToHost,
                           c_->stream));
    MGX_HIP(hipStreamSynchronize(c_->stream));
    std::memcpy(&err, c_->h_pinned, sizeof err);
    MGX_REQUIRE(err == 0, "persistent CG: a grid barrier wait timed out");
    if (!poll().empty()) return;
  }
}

template <class T> std::vector<int> Engine<T>::poll() {
  std::vector<int> out;
  int32_t stopped = 0;
  MGX_HIP(hipMemcpyAsync(c_->h_pinned, d_stopped_, sizeof(int32_t), hipMemcpyDeviceToHost,
                         c_->stream));
  MGX_HIP(hipStreamSynchronize(c_->stream));
  std::memcpy(&stopped, c_->h_pinned, sizeof stopped);
  if (stopped == host_stopped_) return out;
  host_stopped_ = stopped;
  MGX_HIP(hipMemcpy(h_states_.data(), d_states_, sizeof(SolveState) * cap_,
                    hipMemcpyDeviceToHost));
  for (int s : active_)
    if (!reported_[s] && h_states_[s].status != kStatRunning) {
      reported_[s] = 1;
      out.push_back(s);
      host_running_--;
    }
  if (!out.empty()) dirty_ = true;
  return out;
}

template <class T> SolveState Engine<T>::state(int slot) { return h_states_[slot]; }

template <class T> void Engine<T>::resume(int slot, const SolveState &s) {
  MGX_REQUIRE(slot >= 0 && slot < cap_ && reported_[slot] &&
                  std::find(active_.begin(), active_.end(), slot) != active_.end(),
              "engine: resume needs a stopped, unreleased slot");
  MGX_REQUIRE(s.status == kStatRunning, "engine: resume needs a running state");
  MGX_HIP(hipStreamSynchronize(c_->stream));
  h_states_[slot] = s;
  // kernel-finished reductions: the next iteration takes beta and zr from
  // the state instead of the last update's partials
  h_states_[slot].fresh = 2;
  h_states_[slot].ia = s.iter;
  MGX_HIP(hipMemcpy(d_states_ + slot, &h_states_[slot], sizeof(SolveState),
                    hipMemcpyHostToDevice));
  reported_[slot] = 0;
  host_running_++;
  dirty_ = true;
}

template <class T> void Engine<T>::resume_device(int slot) {
  MGX_REQUIRE(slot >= 0 && slot < cap_ && reported_[slot] &&
                  std::find(active_.begin(), active_.end(), slot) != active_.end(),
              "engine: resume needs a stopped, unreleased slot");
  reported_[slot] = 0;
  host_running_++;
  dirty_ = true;
}

template <class T> void Engine<T>::fetch_states() {
  MGX_HIP(hipStreamSynchronize(c_->stream));
  MGX_HIP(hipMemcpy(h_states_.data(), d_states_, sizeof(SolveState) * cap_,
                    hipMemcpyDeviceToHost));
}

template <class T> void Engine<T>::release(int slot) {
  MGX_HIP(hipStreamSynchronize(c_->stream));
  auto it = std::find(active_.begin(), active_.end(), slot);
  if (it == active_.end()) return;
  if (!reported_[slot]) host_running_--;
  active_.erase(it);
  pending_.erase(std::remove(pending_.begin(), pending_.end(), slot), pending_.end());
  for (void *p : bufs_[slot]) pool_->put(p);
  bufs_[slot].clear();
  free_slots_.push_back(slot);
  dirty_ = true;
  listed_ok_ = false;
}

template class Engine<double>;
template class Engine<float>;

DevicePool &ctx_pool(Ctx *c) {
  if (!c->pool) c->pool = new DevicePool((size_t)-1);
  return *static_cast<DevicePool *>(c->pool);
}

template <> Engine<double> &ctx_engine<double>(Ctx *c) {
  if (!c->eng64) c->eng64 = new Engine<double>(c, &ctx_pool(c), 4);
  return *static_cast<Engine<double> *>(c->eng64);
}

template <> Engine<float> &ctx_engine<float>(Ctx *c) {
  if (!c->eng32) c->eng32 = new Engine<float>(c, &ctx_pool(c), 4);
  return *static_cast<Engine<float> *>(c->eng32);
}

void ctx_free_engines(Ctx *c) {
  delete static_cast<Engine<double> *>(c->eng64);
  delete static_cast<Engine<float> *>(c->eng32);
  delete static_cast<DevicePool *>(c->pool);
  c->eng64 = c->eng32 = c->pool = nullptr;
}

}  // namespace mgx
