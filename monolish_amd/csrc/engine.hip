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
  for (auto &f : free_) dev_free(f.second);
  for (auto &l : live_) dev_free(l.first);
}

void *DevicePool::get(size_t bytes) {
  const size_t sz = round(bytes);
  for (size_t i = 0; i < free_.size(); i++)
    if (free_[i].first == sz) {
      void *p = free_[i].second;
      free_.erase(free_.begin() + i);
      live_.push_back({p, sz});
      in_use_ += sz;
      return p;
    }
  if (in_use_ + sz > budget_) return nullptr;
  // drop cached blocks of other sizes if they keep us over budget
  while (reserved_ + sz > budget_ && !free_.empty()) {
    dev_free(free_.back().second);
    reserved_ -= free_.back().first;
    free_.pop_back();
  }
  void *p = nullptr;
  if (hipMalloc(&p, sz) != hipSuccess) return nullptr;
  reserved_ += sz;
  live_.push_back({p, sz});
  in_use_ += sz;
  return p;
}

void DevicePool::put(void *p) {
  if (!p) return;
  for (size_t i = 0; i < live_.size(); i++)
    if (live_[i].first == p) {
      free_.push_back({live_[i].second, p});
      in_use_ -= live_[i].second;
      live_.erase(live_.begin() + i);
      return;
    }
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
}

template <class T> size_t Engine<T>::bytes_needed(const Csr *A, int method) const {
  const int64_t n = std::max<int64_t>(A->rows, 1);
  const size_t v = DevicePool::round((size_t)n * sizeof(T));
  const int nvec = method == kMethodCG ? 6 : 11;
  return nvec * v + DevicePool::round(((size_t)(method == kMethodCG ? 4 : 8) * ntiles_of(n) +
                                       2 * kRedGroups) *
                                      sizeof(T));
}

template <class T>
int Engine<T>::add(Csr *A, const void *val, T *x, const T *b, const SolveCfg &cfg) {
  MGX_REQUIRE(A->rows == A->cols && A->rows > 0, "engine: square, non-empty system required");
  if (free_slots_.empty()) return -1;
  MGX_HIP(hipStreamSynchronize(c_->stream));
  fused_plan_build(A);
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
  // tile for the updates (2 sums); then the stage-2 lane sums
  const int64_t np = (cfg.method == kMethodCG ? 4 : 8) * nt;
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
  d.rhist = cfg.rhist;
  d.rhist_len = cfg.rhist ? cfg.rhist_len : 0;
  d.row0 = 0;
  d.n = n;
  d.ntiles = nt;
  d.split = cfg.method == kMethodCG && n >= kSplitRows ? 1 : 0;
  d.vec = ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(b)) & 15) == 0;
  csr_[slot] = A;
  val_[slot] = val;
  lds_entries_ = std::max(lds_entries_, A->fused_nnz);
  const mgx_dtype dt = DTof<T>::v;
  Ctx *c = c_;
  if (d.M) {
    launch_diag_v(A, val, dt, (void *)d.M);
    launch_reciprocal(c, dt, n, d.M, (void *)d.M);
  }
  // r = b - A x (cg.cpp:39-40, bicgstab.cpp:47-48)
  T *tmp = cfg.method == kMethodCG ? d.q : d.t;
  launch_spmv(A, val, dt, x, tmp, 1, 0, false);
  launch_vv(c, dt, 1, n, b, tmp, d.r);
  if (cfg.method == kMethodCG) {
    if (d.M)
      launch_vv(c, dt, 2, n, d.M, d.r, d.z);
    else
      launch_copy(c, dt, n, d.r, d.z);
  } else {
    launch_copy(c, dt, n, d.r, d.r0);
  }
  SolveState hs;
  std::memset(&hs, 0, sizeof hs);
  hs.tol = cfg.tol;
  hs.miniter = cfg.miniter;
  hs.maxiter = cfg.maxiter;
  hs.status = kStatRunning;
  hs.first = 1;
  h_sys_[slot] = d;
  MGX_HIP(hipMemcpy(d_states_ + slot, &hs, sizeof hs, hipMemcpyHostToDevice));
  MGX_HIP(hipMemcpy(d_sys_ + slot, &d, sizeof d, hipMemcpyHostToDevice));
  if (cfg.tol_relative)
    batch_launch_sys_dot<T>(c, d_sys_ + slot, nullptr, nt, 1, (int)offsetof(SolveState, tol),
                            cfg.tol);
  if (cfg.method == kMethodCG) {
    batch_launch_sys_dot<T>(c, d_sys_ + slot, nullptr, nt, 0, (int)offsetof(SolveState, zr), 0.0);
  } else {
    batch_launch_sys_dot<T>(c, d_sys_ + slot, nullptr, nt, 2, (int)offsetof(SolveState, rho), 0.0);
    batch_launch_bcg_init<T>(c, d_states_ + slot, d_stopped_);
  }
  MGX_HIP(hipGetLastError());
  active_.push_back(slot);
  host_running_++;
  dirty_ = true;
  return slot;
}

template <class T> void Engine<T>::rebuild_lists() {
  std::vector<TileRef> lst[2];
  any_fused_ = false;
  split_slots_.clear();
  for (int s : active_) {
    if (reported_[s]) continue;
    const int m = method_[s];
    if (m == kMethodCG) {
      if (h_sys_[s].split)
        split_slots_.push_back(s);
      else
        any_fused_ = true;
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
  dirty_ = false;
}

template <class T> void Engine<T>::launch_iters(int iters) {
  for (int i = 0; i < iters; i++) {
    if (ntiles_[0]) {
      if (!split_slots_.empty()) {
        // split form: p, then q = A p per system, then <p, q>
        batch_launch_cg_p<T>(c_, d_sys_, d_tiles_[0], ntiles_[0]);
        for (int s : split_slots_)
          launch_spmv(csr_[s], val_[s], DTof<T>::v, h_sys_[s].pa, h_sys_[s].q, 1.0, 0.0, false,
                      &d_states_[s].status);
        batch_launch_cg_pq<T>(c_, d_sys_, d_tiles_[0], ntiles_[0]);
      }
      batch_launch_cg_iter<T>(c_, d_sys_, d_tiles_[0], ntiles_[0], d_stopped_, lds_entries_,
                              any_fused_);
    }
    if (ntiles_[1])
      batch_launch_bicgstab_iter<T>(c_, d_sys_, d_tiles_[1], ntiles_[1], d_stopped_,
                                    lds_entries_);
  }
}

template <class T> void Engine<T>::step(int iters) {
  if (dirty_) {
    rebuild_lists();
    gstale_ = true;
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

template <class T> void Engine<T>::release(int slot) {
  MGX_HIP(hipStreamSynchronize(c_->stream));
  auto it = std::find(active_.begin(), active_.end(), slot);
  if (it == active_.end()) return;
  if (!reported_[slot]) host_running_--;
  active_.erase(it);
  for (void *p : bufs_[slot]) pool_->put(p);
  bufs_[slot].clear();
  free_slots_.push_back(slot);
  dirty_ = true;
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
