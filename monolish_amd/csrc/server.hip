// server.hip -- batched solves over many systems (mgx_batch_solve) and the
// solve server: a request loop that keeps matrices and Krylov state resident
// in HBM, admits arriving systems in order while slots and the HBM budget
// allow, iterates every resident system together in fused launches
// (continuous batching: a system joins or leaves between launches) and
// reports latency percentiles beside throughput.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "mgx_batch.hpp"

namespace mgx {
namespace {

SolveCfg cfg_of(const mgx_solve_params *p, int method) {
  SolveCfg cfg;
  cfg.method = method;
  cfg.precond = p->precond;
  cfg.tol = p->tol;
  cfg.tol_relative = p->tol_is_relative;
  cfg.miniter = p->miniter;
  cfg.maxiter = p->maxiter <= 0 ? INT64_MAX / 4 : p->maxiter;
  cfg.gather_once = 1;  // batches and the server run many systems per launch
  return cfg;
}

template <class T>
int batch_solve_t(Ctx *c, int method, int64_t nsys, mgx_csr *const *A, mgx_vec *const *x,
                  mgx_vec *const *b, const mgx_solve_params *p, mgx_solve_result *res) {
  const double t0 = seconds_now();
  Engine<T> eng(c, &ctx_pool(c), (int)nsys);
  const SolveCfg cfg = cfg_of(p, method);
  {
    // every system's work vectors from one slab when the pool lacks them
    std::vector<size_t> sizes;
    for (int64_t i = 0; i < nsys; i++) Engine<T>::sizes_of(A[i], cfg, sizes);
    (void)ctx_pool(c).reserve(sizes);  // on failure add() allocates per block
  }
  std::vector<int> slot_of(nsys, -1);
  std::vector<int64_t> sys_of(nsys, -1);
  for (int64_t i = 0; i < nsys; i++) {
    const int s = eng.add(A[i], A[i]->val, (T *)x[i]->d, (const T *)b[i]->d, cfg);
    if (s < 0) {
      set_error("mgx_batch_solve: no device memory for system %lld", (long long)i);
      throw Error{MGX_ERR_NOMEM};
    }
    slot_of[i] = s;
    sys_of[s] = i;
  }
  // iterations per poll: given, or tuned online while the batch runs (the
  // schedule changes no bits): after one warm-up step, two steps under each
  // candidate, then the one with the most system-iterations per second
  int every = p->check_every > 0 ? p->check_every : 8;
  const bool tune = p->check_every <= 0;
  static constexpr int kCand[] = {4, 8, 16, 32};
  constexpr int kNCand = (int)(sizeof kCand / sizeof kCand[0]);
  int ci = tune ? -1 : kNCand, meas = 0;
  double acc_t = 0, acc_w = 0, best_rate = -1;
  int best_every = every;
  while (eng.running() > 0) {
    const int running = eng.running();
    const int e = ci >= 0 && ci < kNCand ? kCand[ci] : every;
    const double ts = seconds_now();
    eng.step(e);
    const std::vector<int> stopped = eng.poll();  // waits for the step
    const double dt = seconds_now() - ts;
    if (ci < kNCand) {
      if (ci < 0) {
        ci = 0;  // the warm-up step carried the set-up launches
      } else {
        acc_t += dt;
        acc_w += (double)running * e;
        if (++meas == 2) {
          const double rate = acc_w / std::max(acc_t, 1e-9);
          if (rate > best_rate) {
            best_rate = rate;
            best_every = kCand[ci];
          }
          ci++;
          meas = 0;
          acc_t = acc_w = 0;
          if (ci == kNCand) every = best_every;
        }
      }
    }
    for (int s : stopped) {
      const SolveState st = eng.state(s);
      mgx_solve_result &r = res[sys_of[s]];
      r.status = st.status;
      r.final_iter = st.iter;
      r.final_resid = st.resid;
      r.seconds = seconds_now() - t0;
      r.inner_iter = r.outer_iter = 0;
    }
  }
  for (int64_t i = 0; i < nsys; i++) eng.release(slot_of[i]);
  c->batch_iters_per_poll = every;
  return MGX_OK;
}

// ------------------------------------------------------------- server ----
struct Request {
  int64_t id = 0;
  int method = kMethodCG;
  int64_t n = 0, nnz = 0;
  std::vector<int32_t> rp, ci;
  std::vector<double> val, b, x0;
  SolveCfg cfg;
  double t_submit = 0, t_admit = 0, t_done = 0;
  // resident state
  Csr *A = nullptr;
  double *x = nullptr, *bd = nullptr;
  size_t bytes = 0;
  int slot = -1;
  // submitted by handle: A, b (and x when given) belong to the caller
  bool by_handle = false;
  bool own_x = true;
  // FP32-refined request (mgx_ir_cg_solve's arithmetic): phase 0 waits for
  // its FP64 residual pass, phase 1 runs its FP32 correction (slot32)
  bool ir = false;
  int ir_phase = 0;
  bool ir_first = true;
  double thr = 0, rn = 0;
  int64_t outer = 0, inner = 0;
  float *r32 = nullptr, *d32 = nullptr;
  int slot32 = -1;
  // control: deadline (absolute seconds_now() time, 0: none) and cancel
  double deadline = 0;
  bool cancel = false;
  bool active = false;  // admitted: holds an engine slot
  // outcome
  bool done = false;
  mgx_solve_result res{};
  std::vector<double> x_out;
};

}  // namespace

struct Server {
  Ctx *ctx;
  mgx_server_config cfg;
  std::unique_ptr<DevicePool> pool;
  std::unique_ptr<Engine<double>> eng;
  // FP32-refined requests: FP64 residual passes and FP32 corrections
  std::unique_ptr<Engine<double>> eres;
  std::unique_ptr<Engine<float>> e32;
  std::vector<std::shared_ptr<Request>> by_slot32;
  std::vector<std::shared_ptr<Request>> ir_wait;  // phase 0, admitted
  double ir_inner_tol = 1e-5;
  int64_t ir_max_outer = 100;
  IrSys *d_ir = nullptr;
  TileRef *d_irtl = nullptr;
  int64_t irtl_cap = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::unordered_map<int64_t, std::shared_ptr<Request>> all;
  std::vector<std::shared_ptr<Request>> by_slot;
  std::thread worker;
  bool stop = false;
  int64_t next_id = 1, submitted = 0, completed = 0, cancelled = 0, expired = 0;
  size_t resident = 0, peak = 0, budget = 0;
  std::vector<double> lat, qwait;
  double first_submit = -1, last_done = 0;
  std::string fatal;

  size_t request_bytes(const Request &r) const {
    Csr probe;
    probe.rows = probe.cols = r.n;
    size_t work = eng->bytes_needed(&probe, r.method);
    if (r.ir)  // FP32 correction engine, r32 / d32, the FP32 copy of the values
      work += work / 2 + DevicePool::round((size_t)r.n * 4) * 2 + (size_t)(r.nnz + 4) * 4;
    if (r.by_handle) return work + (r.own_x ? (size_t)r.n * 8 : 0);
    const size_t csr = (size_t)(r.n + 1) * 4 + (size_t)(r.nnz + 4) * 12 +
                       (size_t)(r.n / 256 + 2) * 4 * 2;
    return csr + (size_t)r.n * 16 + work;
  }

  void upload(Request &r) {
    if (r.by_handle) {
      if (r.own_x) {
        r.x = (double *)dev_alloc((size_t)r.n * 8);
        MGX_HIP(hipMemsetAsync(r.x, 0, (size_t)r.n * 8, ctx->stream));
      }
      return;
    }
    mgx_csr *A = nullptr;
    int rc = mgx_csr_create(static_cast<mgx_ctx *>(ctx), r.n, r.n, r.nnz, r.rp.data(),
                            r.ci.data(), r.val.data(), MGX_F64, &A);
    if (rc != MGX_OK) throw Error{rc};
    r.A = A;
    r.x = (double *)dev_alloc((size_t)r.n * 8);
    r.bd = (double *)dev_alloc((size_t)r.n * 8);
    if (!r.x0.empty())
      MGX_HIP(hipMemcpy(r.x, r.x0.data(), (size_t)r.n * 8, hipMemcpyHostToDevice));
    else
      MGX_HIP(hipMemset(r.x, 0, (size_t)r.n * 8));
    MGX_HIP(hipMemcpy(r.bd, r.b.data(), (size_t)r.n * 8, hipMemcpyHostToDevice));
    // host copies are no longer needed
    std::vector<int32_t>().swap(r.rp);
    std::vector<int32_t>().swap(r.ci);
    std::vector<double>().swap(r.val);
    std::vector<double>().swap(r.b);
    std::vector<double>().swap(r.x0);
  }

  void free_resident(Request &r) {
    if (r.r32 || r.d32) {
      (void)hipStreamSynchronize(ctx->stream);
      pool->put(r.r32);
      pool->put(r.d32);
      r.r32 = r.d32 = nullptr;
    }
    if (r.by_handle) {
      if (r.own_x) {
        (void)hipStreamSynchronize(ctx->stream);
        dev_free(r.x);
      }
      r.x = nullptr;
      return;
    }
    if (r.A) (void)mgx_csr_destroy(static_cast<mgx_csr *>(r.A));
    dev_free(r.x);
    dev_free(r.bd);
    r.A = nullptr;
    r.x = r.bd = nullptr;
  }

  void finish(const std::shared_ptr<Request> &r, int status) {
    std::lock_guard<std::mutex> lk(mu);
    finish_locked(r, status);
  }

  void finish_locked(const std::shared_ptr<Request> &r, int status) {
    r->t_done = seconds_now();
    r->res.status = status;
    r->res.seconds = r->t_done - r->t_submit;
    r->done = true;
    r->active = false;
    if (status == MGX_ERR_CANCELLED || status == MGX_ERR_TIMEOUT) {
      (status == MGX_ERR_CANCELLED ? cancelled : expired)++;
    } else {
      completed++;
      lat.push_back(r->t_done - r->t_submit);
      qwait.push_back(r->t_admit > 0 ? r->t_admit - r->t_submit : 0);
      last_done = r->t_done;
    }
    cv_done.notify_all();
  }

  // queued requests that were cancelled or whose deadline passed leave the
  // queue (called with mu held)
  void drop_queued_locked(double now) {
    for (auto it = queue.begin(); it != queue.end();) {
      const auto &r = *it;
      const bool late = r->deadline > 0 && now >= r->deadline;
      if (r->cancel || late) {
        auto keep = r;
        it = queue.erase(it);
        finish_locked(keep, keep->cancel ? MGX_ERR_CANCELLED : MGX_ERR_TIMEOUT);
      } else {
        ++it;
      }
    }
  }

  // running requests that were cancelled or whose deadline passed give up
  // their slot; x keeps the last iterate (by-handle requests) or is copied
  // out (requests submitted with their arrays)
  void drop_active(double now) {
    std::vector<std::shared_ptr<Request>> out, out_ir;
    {
      std::lock_guard<std::mutex> lk(mu);
      auto ends = [&](const std::shared_ptr<Request> &r) {
        return r && !r->done && (r->cancel || (r->deadline > 0 && now >= r->deadline));
      };
      for (auto &r : by_slot)
        if (ends(r)) out.push_back(r);
      for (auto &r : by_slot32)
        if (ends(r)) out_ir.push_back(r);
      for (auto &r : ir_wait)
        if (ends(r)) out_ir.push_back(r);
    }
    // FP32-refined requests: x holds the last refined iterate
    for (auto &r : out_ir) {
      if (r->slot32 >= 0) {
        e32->release(r->slot32);
        by_slot32[r->slot32].reset();
        r->slot32 = -1;
      } else {
        ir_wait.erase(std::remove(ir_wait.begin(), ir_wait.end(), r), ir_wait.end());
      }
      r->res.final_iter = r->inner;
      r->res.final_resid = r->rn;
      r->res.inner_iter = r->inner;
      r->res.outer_iter = r->outer;
      retire(r, r->cancel ? MGX_ERR_CANCELLED : MGX_ERR_TIMEOUT);
    }
    if (!out.empty()) eng->fetch_states();  // waits for the stream: x is final
    for (auto &r : out) {
      const int code = r->cancel ? MGX_ERR_CANCELLED : MGX_ERR_TIMEOUT;
      eng->release(r->slot);
      const SolveState st = eng->state(r->slot);
      r->res.final_iter = st.iter;
      r->res.final_resid = st.resid;
      by_slot[r->slot].reset();
      retire(r, code);
    }
  }

  int ir_active = 0;  // admitted FP32-refined requests not yet finished
  int busy() const { return eng->active() + ir_active; }

  // uploads an IrSys list and its tiles (waits for the stream first: the
  // previous list may still be read)
  int64_t ir_upload(const std::vector<IrSys> &hs, const std::vector<TileRef> &tl) {
    MGX_HIP(hipStreamSynchronize(ctx->stream));
    if (!d_ir) d_ir = (IrSys *)dev_alloc(sizeof(IrSys) * (size_t)cfg.max_active);
    if ((int64_t)tl.size() > irtl_cap) {
      dev_free(d_irtl);
      d_irtl = nullptr;
      irtl_cap = std::max<int64_t>((int64_t)tl.size(), 2 * irtl_cap);
      d_irtl = (TileRef *)dev_alloc(sizeof(TileRef) * irtl_cap);
    }
    if (!hs.empty()) MGX_HIP(hipMemcpy(d_ir, hs.data(), sizeof(IrSys) * hs.size(),
                                       hipMemcpyHostToDevice));
    if (!tl.empty()) MGX_HIP(hipMemcpy(d_irtl, tl.data(), sizeof(TileRef) * tl.size(),
                                       hipMemcpyHostToDevice));
    return (int64_t)tl.size();
  }

  // a finished request leaves: x copied out (requests with their arrays),
  // resident memory returned, outcome recorded
  void retire(const std::shared_ptr<Request> &r, int status) {
    if (!r->by_handle) {
      MGX_HIP(hipStreamSynchronize(ctx->stream));
      r->x_out.resize(r->n);
      MGX_HIP(hipMemcpy(r->x_out.data(), r->x, (size_t)r->n * 8, hipMemcpyDeviceToHost));
    }
    free_resident(*r);
    std::lock_guard<std::mutex> lk(mu);
    resident -= r->bytes;
    if (r->ir) ir_active--;
    finish_locked(r, status);
  }

  // FP32-refined requests waiting for their residual: one FP64 set-up pass
  // for all of them gives r = b - A x and ||r|| (mgx_ir_cg_solve's steps);
  // the stop tests; then r32 = r / ||r||, d32 = 0 and the corrections join
  // the FP32 engine
  void ir_residuals() {
    if (ir_wait.empty()) return;
    std::vector<int> rs(ir_wait.size());
    for (size_t k = 0; k < ir_wait.size(); k++) {
      const auto &r = ir_wait[k];
      SolveCfg rc;
      rc.method = kMethodCG;
      rc.precond = MGX_PRECOND_NONE;
      rc.tol = r->cfg.tol;
      rc.tol_relative = r->cfg.tol_relative;
      rs[k] = eres->add(r->A, r->A->val, r->x, r->bd, rc);
      MGX_REQUIRE(rs[k] >= 0, "server: no device memory for a residual pass");
    }
    eres->prepare();
    eres->fetch_states();
    std::vector<IrSys> hs;
    std::vector<TileRef> tl;
    std::vector<std::shared_ptr<Request>> go;
    std::vector<std::pair<std::shared_ptr<Request>, int>> done;
    for (size_t k = 0; k < ir_wait.size(); k++) {
      const auto &r = ir_wait[k];
      const SolveState st = eres->state(rs[k]);
      if (r->ir_first) r->thr = st.tol;
      r->ir_first = false;
      r->rn = std::sqrt(st.zr);
      int status = kStatRunning;
      if (r->rn < r->thr)
        status = MGX_OK;
      else if (std::isnan(r->rn))
        status = MGX_SOLVER_RESIDUAL_NAN;
      else if (r->outer >= ir_max_outer || r->inner >= r->cfg.maxiter)
        status = MGX_SOLVER_MAXITER;
      if (status != kStatRunning) {
        r->res.final_iter = r->inner;
        r->res.final_resid = r->rn;
        r->res.inner_iter = r->inner;
        r->res.outer_iter = r->outer;
        done.emplace_back(r, status);
        continue;
      }
      const int32_t idx = (int32_t)hs.size();
      hs.push_back(IrSys{eres->desc(rs[k]).r, r->r32, r->d32, r->x, 1.0 / r->rn, r->rn, r->n});
      for (int64_t t = 0; t < ntiles_of(r->n); t++) tl.push_back(TileRef{idx, (int32_t)t});
      go.push_back(r);
    }
    // ir_upload may (re)allocate d_ir / d_irtl: read them only after it returns
    const int64_t nscale = ir_upload(hs, tl);
    launch_ir_scale(ctx, d_ir, d_irtl, nscale);
    for (const auto &r : go) {
      SolveCfg ic;
      ic.method = kMethodCG;
      ic.precond = r->cfg.precond;
      ic.tol = ir_inner_tol;  // absolute, on the unit-norm scaled residual
      ic.tol_relative = 0;
      ic.maxiter = r->cfg.maxiter - r->inner;
      ic.gather_once = 1;
      r->slot32 = e32->add(r->A, r->A->val32, r->d32, r->r32, ic);
      MGX_REQUIRE(r->slot32 >= 0, "server: no device memory for an FP32 correction");
      by_slot32[r->slot32] = r;
      r->ir_phase = 1;
    }
    for (int s : rs) eres->release(s);  // after the scale read their residuals
    ir_wait.clear();
    for (auto &d : done) retire(d.first, d.second);
  }

  // finished FP32 corrections: x += ||r|| d in FP64, then another residual
  void ir_corrections_done(const std::vector<int> &stopped) {
    if (stopped.empty()) return;
    std::vector<IrSys> hs;
    std::vector<TileRef> tl;
    for (int s : stopped) {
      auto r = by_slot32[s];
      r->inner += e32->state(s).iter;
      const int32_t idx = (int32_t)hs.size();
      hs.push_back(IrSys{nullptr, r->r32, r->d32, r->x, 0.0, r->rn, r->n});
      for (int64_t t = 0; t < ntiles_of(r->n); t++) tl.push_back(TileRef{idx, (int32_t)t});
      e32->release(s);
      by_slot32[s].reset();
      r->slot32 = -1;
      r->outer++;
      r->ir_phase = 0;
      ir_wait.push_back(r);
    }
    const int64_t nupdate = ir_upload(hs, tl);  // before d_ir / d_irtl are read
    launch_ir_update(ctx, d_ir, d_irtl, nupdate);
  }

  void loop() {
    activate(ctx);
    for (;;) {
      // admission, in arrival order, while slots and the budget allow
      std::vector<std::shared_ptr<Request>> admit;
      {
        std::unique_lock<std::mutex> lk(mu);
        drop_queued_locked(seconds_now());
        while (!queue.empty() && busy() + (int)admit.size() < cfg.max_active) {
          auto r = queue.front();
          if (r->bytes > budget) {  // can never fit
            queue.pop_front();
            r->res.status = MGX_ERR_BUDGET;
            finish_locked(r, MGX_ERR_BUDGET);
            continue;
          }
          if (resident + r->bytes > budget) break;
          queue.pop_front();
          resident += r->bytes;
          peak = std::max(peak, resident);
          admit.push_back(r);
        }
        if (admit.empty() && busy() == 0) {
          if (stop && queue.empty()) return;
          cv_work.wait_for(lk, std::chrono::milliseconds(2));
          continue;
        }
      }
      for (auto &r : admit) {
        try {
          upload(*r);
          if (r->ir) {
            ensure_val32(r->A);
            r->r32 = (float *)pool->get((size_t)r->n * 4);
            r->d32 = (float *)pool->get((size_t)r->n * 4);
            if (!r->r32 || !r->d32) throw Error{MGX_ERR_BUDGET};
            r->ir_phase = 0;
            ir_wait.push_back(r);
          } else {
            r->slot = eng->add(r->A, r->A->val, r->x, r->bd, r->cfg);
            if (r->slot < 0) throw Error{MGX_ERR_BUDGET};
            by_slot[r->slot] = r;
          }
          r->t_admit = seconds_now();
          std::lock_guard<std::mutex> lk(mu);
          r->active = true;
          if (r->ir) ir_active++;
        } catch (const Error &e) {
          free_resident(*r);
          {
            std::lock_guard<std::mutex> lk(mu);
            resident -= r->bytes;
          }
          finish(r, e.code);
        }
      }
      if (busy() == 0) continue;
      drop_active(seconds_now());
      ir_residuals();
      const bool run64 = eng->active() > 0, run32 = e32->active() > 0;
      if (run64) eng->step(cfg.iters_per_poll);
      if (run32) e32->step(cfg.iters_per_poll);
      if (run64) {
        for (int s : eng->poll()) {
          auto r = by_slot[s];
          const SolveState st = eng->state(s);
          r->res.final_iter = st.iter;
          r->res.final_resid = st.resid;
          eng->release(s);
          by_slot[s].reset();
          retire(r, st.status);
        }
      }
      if (run32) ir_corrections_done(e32->poll());
    }
  }

  void run() {
    try {
      loop();
    } catch (const Error &) {
      std::lock_guard<std::mutex> lk(mu);
      fatal = mgx_last_error();
      stop = true;
      cv_done.notify_all();
    }
  }
};

double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0;
  std::sort(v.begin(), v.end());
  size_t k = (size_t)std::ceil(q * (double)v.size());
  if (k == 0) k = 1;
  return v[std::min(k, v.size()) - 1];
}

}  // namespace mgx

struct mgx_server : mgx::Server {};

using namespace mgx;

extern "C" {

int mgx_batch_solve(mgx_ctx *ctx, int method, int64_t nsys, mgx_csr *const *A,
                    mgx_vec *const *x, mgx_vec *const *b, const mgx_solve_params *p,
                    mgx_solve_result *results) {
  return guard([&] {
    MGX_REQUIRE(ctx && p && results && (nsys == 0 || (A && x && b)),
                "mgx_batch_solve: NULL argument");
    MGX_REQUIRE(method == MGX_METHOD_CG || method == MGX_METHOD_BICGSTAB,
                "mgx_batch_solve: unknown method %d", method);
    MGX_REQUIRE(nsys >= 0 && nsys < (1 << 30), "mgx_batch_solve: bad system count");
    MGX_REQUIRE(p->precond >= MGX_PRECOND_DEFAULT && p->precond <= MGX_PRECOND_JACOBI,
                "mgx_batch_solve: preconditioner must be default, none or Jacobi");
    if (nsys == 0) return MGX_OK;
    const mgx_dtype dt = A[0]->dtype;
    for (int64_t i = 0; i < nsys; i++) {
      MGX_REQUIRE(A[i] && x[i] && b[i], "mgx_batch_solve: NULL system %lld", (long long)i);
      MGX_REQUIRE(A[i]->ctx == ctx && x[i]->ctx == ctx && b[i]->ctx == ctx,
                  "mgx_batch_solve: context mismatch in system %lld", (long long)i);
      MGX_REQUIRE(A[i]->rows == A[i]->cols && A[i]->rows > 0,
                  "mgx_batch_solve: system %lld is not square and non-empty", (long long)i);
      MGX_REQUIRE(A[i]->dtype == dt && x[i]->dtype == dt && b[i]->dtype == dt,
                  "mgx_batch_solve: dtype mismatch in system %lld", (long long)i);
      MGX_REQUIRE(x[i]->n == A[i]->rows && b[i]->n == A[i]->rows,
                  "mgx_batch_solve: size mismatch in system %lld", (long long)i);
    }
    activate(ctx);
    if (dt == MGX_F64) return batch_solve_t<double>(ctx, method, nsys, A, x, b, p, results);
    return batch_solve_t<float>(ctx, method, nsys, A, x, b, p, results);
  });
}

int mgx_server_create(mgx_ctx *ctx, const mgx_server_config *cfg, mgx_server **out) {
  return guard([&] {
    MGX_REQUIRE(ctx && out, "mgx_server_create: NULL argument");
    activate(ctx);
    auto *s = new mgx_server();
    s->ctx = ctx;
    std::memset(&s->cfg, 0, sizeof s->cfg);
    if (cfg) s->cfg = *cfg;
    if (s->cfg.max_active <= 0) s->cfg.max_active = 1024;
    if (s->cfg.iters_per_poll <= 0) s->cfg.iters_per_poll = 8;
    MGX_REQUIRE(s->cfg.ir_inner_tol >= 0 && s->cfg.ir_inner_tol < 1,
                "mgx_server_create: ir_inner_tol in [0, 1)");
    if (s->cfg.ir_inner_tol > 0) s->ir_inner_tol = s->cfg.ir_inner_tol;
    if (s->cfg.ir_max_outer > 0) s->ir_max_outer = s->cfg.ir_max_outer;
    size_t freeb = 0, total = 0;
    MGX_HIP(hipMemGetInfo(&freeb, &total));
    s->budget = s->cfg.hbm_budget ? s->cfg.hbm_budget : freeb / 4 * 3;
    s->pool.reset(new DevicePool(s->budget));
    s->eng.reset(new Engine<double>(ctx, s->pool.get(), s->cfg.max_active));
    s->eres.reset(new Engine<double>(ctx, s->pool.get(), s->cfg.max_active));
    s->e32.reset(new Engine<float>(ctx, s->pool.get(), s->cfg.max_active));
    s->by_slot.resize(s->cfg.max_active);
    s->by_slot32.resize(s->cfg.max_active);
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return MGX_OK;
  });
}

int mgx_server_submit(mgx_server *s, int method, int64_t n, int64_t nnz, const int32_t *row_ptr,
                      const int32_t *col_ind, const double *val, const double *b,
                      const double *x0, const mgx_solve_params *p, int64_t *ticket) {
  return guard([&] {
    MGX_REQUIRE(s && row_ptr && b && p && ticket && (nnz == 0 || (col_ind && val)),
                "mgx_server_submit: NULL argument");
    MGX_REQUIRE(method == MGX_METHOD_CG || method == MGX_METHOD_BICGSTAB,
                "mgx_server_submit: unknown method %d", method);
    MGX_REQUIRE(n > 0 && nnz >= 0, "mgx_server_submit: empty system");
    MGX_REQUIRE(p->precond >= MGX_PRECOND_DEFAULT && p->precond <= MGX_PRECOND_JACOBI,
                "mgx_server_submit: preconditioner must be default, none or Jacobi");
    auto r = std::make_shared<Request>();
    r->method = method;
    r->n = n;
    r->nnz = nnz;
    r->rp.assign(row_ptr, row_ptr + n + 1);
    r->ci.assign(col_ind, col_ind + nnz);
    r->val.assign(val, val + nnz);
    r->b.assign(b, b + n);
    if (x0) r->x0.assign(x0, x0 + n);
    r->cfg = cfg_of(p, method);
    r->ir = p->mode == MGX_MODE_MIXED_IR;
    MGX_REQUIRE(!r->ir || method == MGX_METHOD_CG,
                "mgx_server_submit: FP32-refined requests are CG requests");
    r->bytes = s->request_bytes(*r);
    std::lock_guard<std::mutex> lk(s->mu);
    MGX_REQUIRE(!s->stop, "mgx_server_submit: server stopped (%s)", s->fatal.c_str());
    r->id = s->next_id++;
    r->t_submit = seconds_now();
    if (s->first_submit < 0) s->first_submit = r->t_submit;
    s->submitted++;
    s->all[r->id] = r;
    s->queue.push_back(r);
    s->cv_work.notify_one();
    *ticket = r->id;
    return MGX_OK;
  });
}

int mgx_server_submit_resident(mgx_server *s, int method, mgx_csr *A, mgx_vec *b, mgx_vec *x,
                               const mgx_solve_params *p, int64_t *ticket) {
  return guard([&] {
    MGX_REQUIRE(s && A && b && p && ticket, "mgx_server_submit_resident: NULL argument");
    MGX_REQUIRE(method == MGX_METHOD_CG || method == MGX_METHOD_BICGSTAB,
                "mgx_server_submit_resident: unknown method %d", method);
    MGX_REQUIRE(A->ctx == s->ctx && b->ctx == s->ctx && (!x || x->ctx == s->ctx),
                "mgx_server_submit_resident: context mismatch");
    MGX_REQUIRE(A->dtype == MGX_F64 && b->dtype == MGX_F64 && (!x || x->dtype == MGX_F64),
                "mgx_server_submit_resident: FP64 system required");
    MGX_REQUIRE(A->rows == A->cols && A->rows > 0 && b->n == A->rows && (!x || x->n == A->rows),
                "mgx_server_submit_resident: size mismatch");
    MGX_REQUIRE(p->precond >= MGX_PRECOND_DEFAULT && p->precond <= MGX_PRECOND_JACOBI,
                "mgx_server_submit_resident: preconditioner must be default, none or Jacobi");
    auto r = std::make_shared<Request>();
    r->method = method;
    r->n = A->rows;
    r->nnz = A->nnz;
    r->by_handle = true;
    r->A = A;
    r->bd = static_cast<double *>(b->d);
    r->own_x = x == nullptr;
    r->x = x ? static_cast<double *>(x->d) : nullptr;
    r->cfg = cfg_of(p, method);
    r->ir = p->mode == MGX_MODE_MIXED_IR;
    MGX_REQUIRE(!r->ir || method == MGX_METHOD_CG,
                "mgx_server_submit_resident: FP32-refined requests are CG requests");
    r->bytes = s->request_bytes(*r);
    std::lock_guard<std::mutex> lk(s->mu);
    MGX_REQUIRE(!s->stop, "mgx_server_submit_resident: server stopped (%s)", s->fatal.c_str());
    r->id = s->next_id++;
    r->t_submit = seconds_now();
    if (s->first_submit < 0) s->first_submit = r->t_submit;
    s->submitted++;
    s->all[r->id] = r;
    s->queue.push_back(r);
    s->cv_work.notify_one();
    *ticket = r->id;
    return MGX_OK;
  });
}

int mgx_server_wait(mgx_server *s, int64_t ticket, double timeout_s, mgx_solve_result *res,
                    double *x) {
  return guard([&] {
    MGX_REQUIRE(s && res, "mgx_server_wait: NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    MGX_REQUIRE(it != s->all.end(), "mgx_server_wait: unknown ticket %lld", (long long)ticket);
    auto r = it->second;
    auto ready = [&] { return r->done || (s->stop && !s->fatal.empty()); };
    if (timeout_s < 0)
      s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
      return MGX_ERR_NOT_FOUND;
    if (!r->done) {
      set_error("mgx_server_wait: server failed: %s", s->fatal.c_str());
      return MGX_ERR_HIP;
    }
    *res = r->res;
    if (x && !r->x_out.empty()) std::memcpy(x, r->x_out.data(), (size_t)r->n * 8);
    s->all.erase(it);
    return MGX_OK;
  });
}

int mgx_server_cancel(mgx_server *s, int64_t ticket) {
  return guard([&] {
    MGX_REQUIRE(s, "mgx_server_cancel: NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) {
      set_error("mgx_server_cancel: unknown ticket %lld", (long long)ticket);
      return MGX_ERR_NOT_FOUND;
    }
    auto r = it->second;
    if (r->done) return MGX_OK;  // finished first: the result stands
    r->cancel = true;
    if (!r->active) s->drop_queued_locked(seconds_now());
    s->cv_work.notify_one();
    return MGX_OK;
  });
}

int mgx_server_set_deadline(mgx_server *s, int64_t ticket, double seconds) {
  return guard([&] {
    MGX_REQUIRE(s, "mgx_server_set_deadline: NULL server");
    MGX_REQUIRE(seconds >= 0, "mgx_server_set_deadline: negative time");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) {
      set_error("mgx_server_set_deadline: unknown ticket %lld", (long long)ticket);
      return MGX_ERR_NOT_FOUND;
    }
    it->second->deadline = seconds > 0 ? it->second->t_submit + seconds : 0;
    s->cv_work.notify_one();
    return MGX_OK;
  });
}

int mgx_server_poll(mgx_server *s, int64_t ticket, int32_t *state) {
  return guard([&] {
    MGX_REQUIRE(s && state, "mgx_server_poll: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) {
      set_error("mgx_server_poll: unknown ticket %lld", (long long)ticket);
      return MGX_ERR_NOT_FOUND;
    }
    const auto &r = it->second;
    *state = r->done ? MGX_REQ_DONE : (r->active ? MGX_REQ_ACTIVE : MGX_REQ_QUEUED);
    return MGX_OK;
  });
}

int mgx_server_stats_get(mgx_server *s, mgx_server_stats *out) {
  return guard([&] {
    MGX_REQUIRE(s && out, "mgx_server_stats_get: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->queued = (int64_t)s->queue.size();
    out->active = s->submitted - s->completed - s->cancelled - s->expired - out->queued;
    out->p50_latency_s = percentile(s->lat, 0.50);
    out->p99_latency_s = percentile(s->lat, 0.99);
    double sum = 0, mx = 0, qs = 0;
    for (double v : s->lat) {
      sum += v;
      mx = std::max(mx, v);
    }
    for (double v : s->qwait) qs += v;
    out->mean_latency_s = s->lat.empty() ? 0 : sum / (double)s->lat.size();
    out->max_latency_s = mx;
    out->mean_queue_s = s->qwait.empty() ? 0 : qs / (double)s->qwait.size();
    const double span = s->last_done - s->first_submit;
    out->solves_per_s = span > 0 ? (double)s->completed / span : 0;
    out->peak_bytes = s->peak;
    out->cancelled = s->cancelled;
    out->expired = s->expired;
    return MGX_OK;
  });
}

int mgx_server_destroy(mgx_server *s) {
  return guard([&] {
    if (!s) return MGX_OK;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
      s->cv_work.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    for (auto &r : s->by_slot)
      if (r) s->free_resident(*r);
    for (auto &r : s->by_slot32)
      if (r) s->free_resident(*r);
    for (auto &r : s->ir_wait) s->free_resident(*r);
    s->eng.reset();
    s->eres.reset();
    s->e32.reset();
    dev_free(s->d_ir);
    dev_free(s->d_irtl);
    s->pool.reset();
    delete s;
    return MGX_OK;
  });
}

}  // extern "C"
