// batch_ir.hip -- the FP32-refined form of the batched CG: iterative
// refinement of many FP64 systems at once.  Per refinement step, for every
// system still running: r = b - A x and ||r|| in FP64 (the set-up pass of an
// FP64 engine computes both for all systems together), the stop tests on
// ||r||, then the corrections A32 d = r / ||r|| of all systems by one
// batched FP32 Jacobi/none-PCG to the accuracy inner_tol, then x += ||r|| d
// in FP64.  The arithmetic of each system is that of mgx_ir_cg_solve.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "mgx_batch.hpp"
#include "mgx_device.hpp"

namespace mgx {

namespace {

// r32 = (float)(r / ||r||) and d32 = 0 over one 512-row tile of a system
__global__ __launch_bounds__(kBlock) void k_ir_scale(const IrSys *sys, const TileRef *tiles,
                                                     int64_t nlaunch) {
  const int64_t b = blockIdx.x;
  if (b >= nlaunch) return;
  const TileRef tr = tiles[b];
  const IrSys &s = sys[tr.slot];
  const int64_t end = min(s.n, ((int64_t)tr.tile + 1) * kTile);
  for (int64_t i = (int64_t)tr.tile * kTile + threadIdx.x; i < end; i += kBlock) {
    s.r32[i] = (float)(s.r[i] * s.scale);
    s.d32[i] = 0.0f;
  }
}

// x += ||r|| d over one tile of a system
__global__ __launch_bounds__(kBlock) void k_ir_update(const IrSys *sys, const TileRef *tiles,
                                                      int64_t nlaunch) {
  const int64_t b = blockIdx.x;
  if (b >= nlaunch) return;
  const TileRef tr = tiles[b];
  const IrSys &s = sys[tr.slot];
  const int64_t end = min(s.n, ((int64_t)tr.tile + 1) * kTile);
  for (int64_t i = (int64_t)tr.tile * kTile + threadIdx.x; i < end; i += kBlock)
    s.x[i] = s.x[i] + s.rn * (double)s.d32[i];
}

}  // namespace

void launch_ir_scale(Ctx *c, const IrSys *sys, const TileRef *tiles, int64_t ntiles) {
  if (ntiles <= 0) return;
  MGX_REQUIRE(sys && tiles, "refinement: system list not on the device");
  k_ir_scale<<<(unsigned)ntiles, kBlock, 0, c->stream>>>(sys, tiles, ntiles);
  MGX_HIP(hipGetLastError());
}

void launch_ir_update(Ctx *c, const IrSys *sys, const TileRef *tiles, int64_t ntiles) {
  if (ntiles <= 0) return;
  MGX_REQUIRE(sys && tiles, "refinement: system list not on the device");
  k_ir_update<<<(unsigned)ntiles, kBlock, 0, c->stream>>>(sys, tiles, ntiles);
  MGX_HIP(hipGetLastError());
}

void batch_ir_cg(Ctx *c, int64_t nsys, Csr *const *A, double *const *x, const double *const *b,
                 const mgx_solve_params *p, double inner_tol, int64_t max_outer,
                 mgx_solve_result *res) {
  const double t0 = seconds_now();
  DevicePool &pool = ctx_pool(c);
  const int64_t max_inner = p->maxiter <= 0 ? INT64_MAX / 4 : p->maxiter;
  for (int64_t i = 0; i < nsys; i++) ensure_val32(A[i]);
  // per-system FP32 work vectors
  std::vector<float *> r32(nsys, nullptr), d32(nsys, nullptr);
  std::vector<void *> owned;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(c->stream);
    for (void *q : owned) pool.put(q);
    owned.clear();
  };
  try {
    {
      // the FP32 vectors and both engines' work vectors of every system from
      // one slab when the pool lacks them
      SolveCfg r0, i0;
      r0.method = i0.method = kMethodCG;
      r0.precond = MGX_PRECOND_NONE;
      i0.precond = p->precond;
      std::vector<size_t> sizes;
      for (int64_t i = 0; i < nsys; i++) {
        sizes.push_back((size_t)A[i]->rows * sizeof(float));
        sizes.push_back((size_t)A[i]->rows * sizeof(float));
        Engine<double>::sizes_of(A[i], r0, sizes);
        Engine<float>::sizes_of(A[i], i0, sizes);
      }
      (void)pool.reserve(sizes);
    }
    for (int64_t i = 0; i < nsys; i++) {
      const size_t bytes = (size_t)A[i]->rows * sizeof(float);
      r32[i] = (float *)pool.get(bytes);
      if (r32[i]) owned.push_back(r32[i]);
      d32[i] = (float *)pool.get(bytes);
      if (d32[i]) owned.push_back(d32[i]);
      MGX_REQUIRE(r32[i] && d32[i], "mgx_batch_ir_cg_solve: no device memory for system %lld",
                  (long long)i);
    }
    std::vector<double> thr(nsys, 0.0), rn(nsys, 0.0);
    std::vector<int64_t> inner(nsys, 0), outer(nsys, 0);
    std::vector<char> running(nsys, 1);
    // FP64 residual pass: CG set-up without preconditioner gives r = b - A x
    // and <r, r>; on the first pass it also gives the thresholds
    SolveCfg rcfg;
    rcfg.method = kMethodCG;
    rcfg.precond = MGX_PRECOND_NONE;
    rcfg.tol = p->tol;
    rcfg.tol_relative = p->tol_is_relative;
    SolveCfg icfg;
    icfg.method = kMethodCG;
    icfg.precond = p->precond;
    icfg.tol = inner_tol;  // absolute on the unit-norm scaled residual
    icfg.tol_relative = 0;
    icfg.gather_once = 1;  // every correction of the batch in the same launches
    IrSys *d_ir = (IrSys *)dev_alloc(sizeof(IrSys) * std::max<int64_t>(nsys, 1));
    TileRef *d_tl = nullptr;
    int64_t tl_cap = 0;
    std::vector<IrSys> h_ir(nsys);
    bool first = true;
    try {
      for (;;) {
        std::vector<int64_t> act;
        for (int64_t i = 0; i < nsys; i++)
          if (running[i]) act.push_back(i);
        if (act.empty()) break;
        Engine<double> e64(c, &pool, (int)act.size());
        std::vector<int> s64(act.size());
        for (size_t k = 0; k < act.size(); k++) {
          const int64_t i = act[k];
          s64[k] = e64.add(A[i], A[i]->val, x[i], b[i], rcfg);
          MGX_REQUIRE(s64[k] >= 0, "mgx_batch_ir_cg_solve: no device memory for system %lld",
                      (long long)i);
        }
        e64.prepare();
        e64.fetch_states();
        std::vector<int64_t> go;
        for (size_t k = 0; k < act.size(); k++) {
          const int64_t i = act[k];
          const SolveState st = e64.state(s64[k]);
          if (first) thr[i] = st.tol;
          rn[i] = std::sqrt(st.zr);
          mgx_solve_result &r = res[i];
          int status = kStatRunning;
          if (rn[i] < thr[i])
            status = MGX_OK;
          else if (std::isnan(rn[i]))
            status = MGX_SOLVER_RESIDUAL_NAN;
          else if (outer[i] >= max_outer || inner[i] >= max_inner)
            status = MGX_SOLVER_MAXITER;
          if (status != kStatRunning) {
            running[i] = 0;
            r.status = status;
            r.final_iter = inner[i];
            r.final_resid = rn[i];
            r.inner_iter = inner[i];
            r.outer_iter = outer[i];
            r.seconds = seconds_now() - t0;
            continue;
          }
          go.push_back(i);
          IrSys &s = h_ir[i];
          s.r = e64.desc(s64[k]).r;
          s.r32 = r32[i];
          s.d32 = d32[i];
          s.x = x[i];
          s.scale = 1.0 / rn[i];
          s.rn = rn[i];
          s.n = A[i]->rows;
        }
        first = false;
        if (go.empty()) break;
        // scaled FP32 residuals of the systems that go on
        std::vector<TileRef> tl;
        for (int64_t i : go)
          for (int64_t k = 0; k < ntiles_of(A[i]->rows); k++)
            tl.push_back(TileRef{(int32_t)i, (int32_t)k});
        if ((int64_t)tl.size() > tl_cap) {
          dev_free(d_tl);
          d_tl = nullptr;
          tl_cap = std::max<int64_t>((int64_t)tl.size(), 2 * tl_cap);
          d_tl = (TileRef *)dev_alloc(sizeof(TileRef) * tl_cap);
        }
        MGX_HIP(hipMemcpy(d_ir, h_ir.data(), sizeof(IrSys) * nsys, hipMemcpyHostToDevice));
        MGX_HIP(hipMemcpy(d_tl, tl.data(), sizeof(TileRef) * tl.size(), hipMemcpyHostToDevice));
        launch_ir_scale(c, d_ir, d_tl, (int64_t)tl.size());
        // all corrections in one batched FP32 solve
        {
          Engine<float> e32(c, &pool, (int)go.size());
          std::vector<int> s32(go.size());
          for (size_t k = 0; k < go.size(); k++) {
            const int64_t i = go[k];
            SolveCfg cfg = icfg;
            cfg.maxiter = max_inner - inner[i];
            s32[k] = e32.add(A[i], A[i]->val32, d32[i], r32[i], cfg);
            MGX_REQUIRE(s32[k] >= 0, "mgx_batch_ir_cg_solve: no device memory for system %lld",
                        (long long)i);
          }
          const int every = p->check_every > 0 ? p->check_every : 8;
          while (e32.running() > 0) {
            e32.step(every);
            e32.poll();
          }
          for (size_t k = 0; k < go.size(); k++) {
            inner[go[k]] += e32.state(s32[k]).iter;
            e32.release(s32[k]);
          }
        }
        launch_ir_update(c, d_ir, d_tl, (int64_t)tl.size());
        for (int64_t i : go) outer[i]++;
        for (size_t k = 0; k < act.size(); k++) e64.release(s64[k]);
      }
    } catch (...) {
      (void)hipStreamSynchronize(c->stream);
      dev_free(d_ir);
      dev_free(d_tl);
      throw;
    }
    MGX_HIP(hipStreamSynchronize(c->stream));
    dev_free(d_ir);
    dev_free(d_tl);
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
}

}  // namespace mgx

using namespace mgx;

extern "C" int mgx_batch_ir_cg_solve(mgx_ctx *ctx, int64_t nsys, mgx_csr *const *A,
                                     mgx_vec *const *x, mgx_vec *const *b,
                                     const mgx_solve_params *p, double inner_tol,
                                     int64_t max_outer, mgx_solve_result *results) {
  return guard([&] {
    MGX_REQUIRE(ctx && p && results && (nsys == 0 || (A && x && b)),
                "mgx_batch_ir_cg_solve: NULL argument");
    MGX_REQUIRE(nsys >= 0 && nsys < (1 << 30), "mgx_batch_ir_cg_solve: bad system count");
    MGX_REQUIRE(inner_tol > 0 && inner_tol < 1, "mgx_batch_ir_cg_solve: inner_tol in (0,1)");
    MGX_REQUIRE(p->precond >= MGX_PRECOND_DEFAULT && p->precond <= MGX_PRECOND_JACOBI,
                "mgx_batch_ir_cg_solve: unknown preconditioner");
    if (nsys == 0) return MGX_OK;
    std::vector<Csr *> As(nsys);
    std::vector<double *> xs(nsys);
    std::vector<const double *> bs(nsys);
    for (int64_t i = 0; i < nsys; i++) {
      MGX_REQUIRE(A[i] && x[i] && b[i], "mgx_batch_ir_cg_solve: NULL system %lld",
                  (long long)i);
      MGX_REQUIRE(A[i]->ctx == ctx && x[i]->ctx == ctx && b[i]->ctx == ctx,
                  "mgx_batch_ir_cg_solve: context mismatch in system %lld", (long long)i);
      MGX_REQUIRE(A[i]->rows == A[i]->cols && A[i]->rows > 0,
                  "mgx_batch_ir_cg_solve: system %lld is not square and non-empty",
                  (long long)i);
      MGX_REQUIRE(A[i]->dtype == MGX_F64 && x[i]->dtype == MGX_F64 && b[i]->dtype == MGX_F64,
                  "mgx_batch_ir_cg_solve: system %lld is not FP64", (long long)i);
      MGX_REQUIRE(x[i]->n == A[i]->rows && b[i]->n == A[i]->rows,
                  "mgx_batch_ir_cg_solve: size mismatch in system %lld", (long long)i);
      As[i] = A[i];
      xs[i] = (double *)x[i]->d;
      bs[i] = (const double *)b[i]->d;
    }
    activate(ctx);
    batch_ir_cg(ctx, nsys, As.data(), xs.data(), bs.data(), p, inner_tol,
                max_outer > 0 ? max_outer : 100, results);
    return MGX_OK;
  });
}
