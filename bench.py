"""Benchmark of the CRS SpMV + Krylov path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]

The timed step is one SpMV (blas::matvec) over the 3D 7-point Laplacian on a
192^3 grid in FP64 (n = 7,077,888, nnz = 49,324,032), operands resident in HBM.
``value`` is the SpMV's achieved algorithmic bandwidth in GB/s, measured with
HIP events on the library's stream around exactly K launches.  The line also
carries the roofline object, the CPU baseline (the oracle's CPU restatement of
monolish's OpenMP row loop, timed on this host), the FP32 / random-matrix SpMV
sweep, and the Krylov side of the metric: Jacobi-PCG solves/s on the 2D Poisson
256x256 problem to a 1e-10 relative residual, plain FP64 GPU CG (one kernel per
BLAS call, scalars returned to the host) against the fused solver, with p50/p99
latency, plus mixed-precision refinement against FP64 CG on the 27-point
stencil.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import statistics
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


class HipEvents:
    """hipEvent timing on a given stream (the library's own stream)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipEventDestroy.argtypes = [C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, C.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = C.c_float(0)
        assert self.hip.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return float(ms.value)


def spmv_bytes(nnz: int, rows: int, cols: int, sv: int) -> int:
    # val + col_ind + row_ptr + x (read once) + y (written)
    return nnz * (sv + 4) + (rows + 1) * 4 + cols * sv + rows * sv


def time_spmv(m, ev, ctx, A, X, Y, steps, warmup):
    for _ in range(warmup):
        m.matvec(A, X, Y)
    ctx.sync()
    a, b = ev.event(), ev.event()
    t0 = time.perf_counter()
    ev.record(a, ctx.stream)
    for _ in range(steps):
        m.matvec(A, X, Y)
    ev.record(b, ctx.stream)
    ms = ev.elapsed_ms(a, b)
    ctx.sync()
    wall = time.perf_counter() - t0
    return ms / steps, wall


def cpu_info():
    model = "unknown"
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except OSError:
        pass
    return model


def bench_spmv_matrix(m, ev, ctx, name, rp, ci, val, x, steps, warmup, plan_spec=None):
    rows = len(rp) - 1
    A = m.CRS(ctx, rows, len(x), rp, ci, val)
    X = m.Vector(ctx, x, dtype=val.dtype)
    Y = m.Vector(ctx, n=rows, dtype=val.dtype)
    if plan_spec:
        # a fixed plan (profiling runs: every dispatch of the kernel is timed work)
        plan = dict(parse_plan(plan_spec), long_row=256, candidates=0)
        A.set_plan(**{k: plan[k] for k in ("rows_per_block", "nnz_per_block", "xcd_remap",
                                           "row_walk", "column_panels")})
    else:
        # the library's autotuner picks the plan (outside the timed region)
        plan = A.autotune(reps=20)
    ms, wall = time_spmv(m, ev, ctx, A, X, Y, steps, warmup)
    nbytes = spmv_bytes(len(val), rows, len(x), val.dtype.itemsize)
    gbs = nbytes / (ms * 1e-3) / 1e9
    log(f"{name}: {ms * 1e3:.1f} us/SpMV, {gbs:.0f} GB/s ({gbs / HBM_PEAK_GBS:.3f} of 8 TB/s)")
    A.free()
    X.free()
    Y.free()
    out = {"us_per_spmv": round(ms * 1e3, 2), "GB_s": round(gbs, 1),
           "frac_hbm": round(gbs / HBM_PEAK_GBS, 4), "bytes": nbytes, "nnz": int(len(val)),
           "rows": rows, "kernel": spmv_kernel_name(plan, val.dtype)}
    out["plan"] = {k: plan[k] for k in ("rows_per_block", "nnz_per_block", "long_row",
                                        "xcd_remap", "row_walk", "column_panels", "candidates")
                   if k in plan}
    return out, ms, wall


def c1_problem(m):
    """C1: laplacian_matrix_2D_5p(256, 256) as CRS, b = vector(n, -1, 1, 1001)."""
    rp, ci, val = m.laplacian2d(256, 256)
    return rp, ci, val, m.random_vector(len(rp) - 1, -1.0, 1.0, 1001)


def bench_krylov(m, ctx, reps_fused, reps_plain):
    rp, ci, val, b = c1_problem(m)
    n = len(rp) - 1
    A = m.CRS(ctx, n, n, rp, ci, val)
    plan = A.autotune(reps=10)
    B = m.Vector(ctx, b)
    X = m.Vector(ctx, n=n)
    bn = m.nrm2(B)
    tol = 1e-10 * bn
    out = {"workload": "Jacobi-PCG, 2D 5-point Poisson 256x256 (n=65536, nnz=326656), "
                       "b=U[-1,1) seed 1001, x0=0, ||r||<1e-10||b||",
           "spmv_plan": {k: plan[k] for k in ("row_walk", "rows_per_block", "nnz_per_block",
                                              "xcd_remap", "column_panels")}}

    def run(mode, reps):
        lat = []
        res = None
        for i in range(reps + 1):
            X.fill(0.0)
            ctx.sync()
            t0 = time.perf_counter()
            res = m.cg(A, X, B, tol=tol, precond=m.PRECOND_JACOBI, mode=mode)
            dt = time.perf_counter() - t0
            if i > 0:
                lat.append(dt)
        lat.sort()
        p99 = lat[min(len(lat) - 1, int(np.ceil(0.99 * len(lat))) - 1)]
        return res, {"solves_per_s": round(1.0 / statistics.mean(lat), 2),
                     "p50_ms": round(1e3 * statistics.median(lat), 3), "p99_ms": round(1e3 * p99, 3),
                     "iterations": res.final_iter, "final_resid": res.final_resid, "solves": len(lat)}

    res_p, plain = run(m.MODE_REFERENCE, reps_plain)
    xp = X.download()
    res_f, fused = run(m.MODE_FUSED, reps_fused)
    xf = X.download()
    out["plain_fp64_gpu_cg"] = plain
    out["fused_fp64_gpu_cg"] = fused
    out["fused_vs_plain"] = round(fused["solves_per_s"] / plain["solves_per_s"], 2)
    out["fused_bit_identical_to_plain"] = bool(np.array_equal(xp.view(np.uint64), xf.view(np.uint64)))
    out["true_rel_residual"] = m.residual_l2(A, X, B) / bn
    log(f"krylov C1: plain {plain['solves_per_s']}/s fused {fused['solves_per_s']}/s "
        f"(x{out['fused_vs_plain']}), {fused['iterations']} it, p99 {fused['p99_ms']} ms")
    A.free()
    B.free()
    X.free()
    return out, xf, tol


def bench_sor(m, ctx, reps, with_cpu):
    """SOR-, IC(0)- and ILU(0)-preconditioned CG and the SOR solver on the C1
    Poisson system: the sweeps and the factorisation run level by level on
    the device (511 levels of at most 256 rows for the 256x256 grid)."""
    rp, ci, val, b = c1_problem(m)
    n = len(rp) - 1
    A = m.CRS(ctx, n, n, rp, ci, val)
    B = m.Vector(ctx, b)
    X = m.Vector(ctx, n=n)
    bn = m.nrm2(B)
    tol = 1e-10 * bn
    out = {"workload": "SOR-PCG (omega 1.9), IC(0)-PCG, ILU(0)-PCG and SOR(1.9) iteration, "
                       "2D 5-point Poisson 256x256, b=U[-1,1) seed 1001, x0=0"}
    runs = {"sor_pcg": lambda: m.cg(A, X, B, tol=tol, precond=m.PRECOND_SOR, omega=1.9),
            "sor_solver": lambda: m.sor(A, X, B, tol=1e-6, maxiter=2000, omega=1.9),
            "ic_pcg": lambda: m.cg(A, X, B, tol=tol, precond=m.PRECOND_IC),
            "ilu_pcg": lambda: m.cg(A, X, B, tol=tol, precond=m.PRECOND_ILU)}
    xs = {}
    for name, fn in runs.items():
        lat = []
        for i in range(reps + 1):
            X.fill(0.0)
            ctx.sync()
            t0 = time.perf_counter()
            res = fn()
            dt = time.perf_counter() - t0
            if i > 0:
                lat.append(dt)
        xs[name] = X.download()
        med = statistics.median(lat)
        out[name] = {"solves_per_s": round(1.0 / med, 3), "p50_ms": round(1e3 * med, 2),
                     "iterations": res.final_iter, "status": res.status,
                     "ms_per_iteration": round(1e3 * med / max(res.final_iter, 1), 4)}
        log(f"sor {name}: {out[name]}")
    if with_cpu:
        import oracle
        oracle.set_sor_omega(1.9)
        t0 = time.perf_counter()
        ref = oracle.solve("cg", rp, ci, val, np.zeros(n), b, precond=oracle.PRECOND_SOR, tol=tol)
        t1 = time.perf_counter()
        ref2 = oracle.solve("sor", rp, ci, val, np.zeros(n), b, precond=oracle.PRECOND_NONE,
                            tol=1e-6, maxiter=2000)
        t2 = time.perf_counter()
        ref3 = oracle.solve("cg", rp, ci, val, np.zeros(n), b, precond=oracle.PRECOND_IC, tol=tol)
        t3 = time.perf_counter()
        out["cpu"] = {"kind": "port", "threads": 1,
                      "sor_pcg_solves_per_s": round(1.0 / (t1 - t0), 3),
                      "sor_solver_solves_per_s": round(1.0 / (t2 - t1), 3),
                      "ic_pcg_solves_per_s": round(1.0 / (t3 - t2), 3),
                      "ic_pcg_gpu_bit_identical": bool(np.array_equal(
                          ref3[1].view(np.uint64), xs["ic_pcg"].view(np.uint64))),
                      "sor_pcg_gpu_bit_identical": bool(np.array_equal(
                          ref[1].view(np.uint64), xs["sor_pcg"].view(np.uint64))),
                      "sor_solver_gpu_bit_identical": bool(np.array_equal(
                          ref2[1].view(np.uint64), xs["sor_solver"].view(np.uint64)))}
        log(f"sor cpu: {out['cpu']}")
    A.free()
    B.free()
    X.free()
    out["wide_levels"] = bench_sor_wide(m, ctx)
    return out


def bench_sor_wide(m, ctx, iters=20):
    """The same preconditioners on the 27-point 128^3 system (2.1 M rows in
    890 levels of about 2400 rows): the sweeps run dependency-driven across
    the chip.  A fixed budget of iterations (tolerance out of reach) after
    one untimed solve that builds the plans; every IC / ILU solve factors A
    first (one workgroup), which the per-iteration figure includes."""
    rp, ci, val = m.stencil3d(128, 128, 128, 27, 0.01)
    n = len(rp) - 1
    A = m.CRS(ctx, n, n, rp, ci, val)
    B = m.Vector(ctx, m.random_vector(n, -1.0, 1.0, 3001))
    X = m.Vector(ctx, n=n)
    tol = 1e-30 * m.nrm2(B)
    out = {"workload": f"{iters} PCG iterations, 27-point 128^3 (diagonal shift 0.01), "
                       "b=U[-1,1) seed 3001, x0=0; IC and ILU include the factorisation "
                       "each solve starts with"}
    def timed(prec, k):
        X.fill(0.0)
        ctx.sync()
        t0 = time.perf_counter()
        r = m.cg(A, X, B, tol=tol, precond=prec, mode=m.MODE_REFERENCE, maxiter=k)
        ctx.sync()
        return time.perf_counter() - t0, r

    for name, prec in (("jacobi", m.PRECOND_JACOBI), ("sor", m.PRECOND_SOR),
                       ("ic", m.PRECOND_IC), ("ilu", m.PRECOND_ILU)):
        m.cg(A, X, B, tol=tol, precond=prec, mode=m.MODE_REFERENCE, maxiter=1)
        t1, _ = timed(prec, 1)
        dt, res = timed(prec, iters)
        # solve set-up (the IC / ILU factorisation, the preconditioner's
        # vectors) = the one-iteration solve minus one iteration
        per_it = (dt - t1) / max(res.final_iter - 1, 1)
        out[name] = {"ms_per_iteration": round(1e3 * dt / max(res.final_iter, 1), 3),
                     "ms_per_iteration_after_setup": round(1e3 * per_it, 3),
                     "setup_ms": round(1e3 * (t1 - per_it), 2),
                     "iterations": res.final_iter}
        log(f"sor wide {name}: {out[name]}")
    A.free()
    B.free()
    X.free()
    return out


def bench_eigen(m, ctx, with_cpu):
    """LOBPCG, 10 smallest eigenpairs of the 2D 5-point Laplacian 128x128
    (the reference's SEVP benchmark shape: 10 pairs, residual tol 1e-4,
    x(i, 7i mod n) = 1), without and with the Jacobi preconditioner."""
    rp, ci, val = m.laplacian2d(128, 128)
    n = len(rp) - 1
    k = 10
    X0 = np.zeros((k, n))
    for i in range(k):
        X0[i, (i * 7) % n] = 1.0
    exact = np.sort([4 - 2 * np.cos(i * np.pi / 129) - 2 * np.cos(j * np.pi / 129)
                     for i in range(1, 20) for j in range(1, 20)])[:k]
    A = m.CRS(ctx, n, n, rp, ci, val)
    X = m.Vector(ctx, n=k * n)
    out = {"workload": "LOBPCG 10 smallest eigenpairs, 2D 5-point Laplacian 128x128 "
                       "(n=16384), tol 1e-4, x0(i, 7i mod n) = 1"}
    for name, prec in (("none", m.PRECOND_DEFAULT), ("jacobi", m.PRECOND_JACOBI)):
        X.upload(X0.reshape(-1))
        ctx.sync()
        t0 = time.perf_counter()
        res, lam = m.lobpcg(A, X, k, tol=1e-4, maxiter=n, precond=prec)
        dt = time.perf_counter() - t0
        out[name] = {"seconds": round(dt, 3), "iterations": res.final_iter, "status": res.status,
                     "ms_per_iteration": round(1e3 * dt / max(res.final_iter, 1), 3),
                     "max_rel_eig_error": float(np.max(np.abs(lam - exact) / exact))}
        log(f"eigen {name}: {out[name]}")
    if with_cpu:
        import oracle
        it = 40
        t0 = time.perf_counter()
        oracle.lobpcg(rp, ci, val, X0, k, tol=0.0, maxiter=it)
        dt = time.perf_counter() - t0
        out["cpu"] = {"kind": "port", "threads": oracle.num_threads(),
                      "ms_per_iteration": round(1e3 * dt / it, 3),
                      "sample": f"first {it} iterations of the unpreconditioned solve "
                                "(numpy restatement; SpMV in the oracle's C loop)"}
        log(f"eigen cpu: {out['cpu']}")
    A.free()
    X.free()
    return out


def bench_mixed(m, ctx, side, reps):
    rp, ci, val = m.stencil3d(side, side, side, 27, 1e-2)
    n = len(rp) - 1
    b = m.random_vector(n, -1.0, 1.0, 3001)
    A = m.CRS(ctx, n, n, rp, ci, val)
    A.autotune(reps=5)
    B = m.Vector(ctx, b)
    X = m.Vector(ctx, n=n)
    bn = m.nrm2(B)
    res = {}
    # restarted refinement: the inner FP32 solves restart from zero at every
    # refinement step, so the inner tolerance trades restarts against FP32
    # stagnation; residual replacement keeps the search direction and
    # replaces the FP32 residual by the FP64 one after each drop by delta
    inner_tols = (1e-4, 1e-5, 1e-6)
    deltas = (1e-1, 1e-2, 1e-3, 1e-4)
    legs = [("fp64_fused_cg", lambda: m.cg(A, X, B, tol=1e-10 * bn, precond=m.PRECOND_JACOBI)),
            ("fp64_plain_cg", lambda: m.cg(A, X, B, tol=1e-10 * bn, precond=m.PRECOND_JACOBI,
                                           mode=m.MODE_REFERENCE))]
    for it in inner_tols:
        legs.append((f"mixed_ir_cg_inner{it:g}",
                     lambda it=it: m.ir_cg(A, X, B, tol=1e-10, tol_is_relative=True, inner_tol=it)))
    for dl in deltas:
        legs.append((f"mixed_rr_cg_delta{dl:g}",
                     lambda dl=dl: m.mixed_cg(A, X, B, tol=1e-10, tol_is_relative=True, delta=dl)))
    for name, fn in legs:
        ts = []
        r = None
        for i in range(reps + 1):
            X.fill(0.0)
            ctx.sync()
            t0 = time.perf_counter()
            r = fn()
            ts.append(time.perf_counter() - t0)
        x = X.download()
        true_rel = m.residual_l2(A, X, B) / bn
        res[name] = {"ms": round(1e3 * statistics.median(ts[1:]), 3), "iterations": r.final_iter,
                     "outer": r.outer_iter, "true_rel_residual": true_rel, "status": r.status}
        res[name + "_x"] = x
    ref = res["fp64_fused_cg_x"]
    for name, _ in legs[1:]:
        xx = res.pop(name + "_x")
        res[name]["rel_err_vs_fp64"] = float(np.linalg.norm(xx - ref) / np.linalg.norm(ref))
    res.pop("fp64_fused_cg_x")

    def fastest(prefix):
        ok = [k for k, _ in legs if k.startswith(prefix) and res[k]["status"] == 0
              and res[k]["true_rel_residual"] <= 1e-10]
        return min(ok, key=lambda k: res[k]["ms"]) if ok else None

    # the fastest converged setting of each form; the faster form is reported
    for form, prefix, key in (("mixed_ir_cg", "mixed_ir_cg_inner", "inner_tol"),
                              ("mixed_rr_cg", "mixed_rr_cg_delta", "delta")):
        best = fastest(prefix)
        if best:
            res[form] = dict(res[best], **{key: float(best[len(prefix):])})
    forms = [f for f in ("mixed_ir_cg", "mixed_rr_cg") if f in res]
    if forms:
        best = min(forms, key=lambda f: res[f]["ms"])
        res["mixed_best"] = dict(res[best], form=best)
        res["mixed_vs_plain_fp64"] = round(res["fp64_plain_cg"]["ms"] / res[best]["ms"], 2)
        res["mixed_vs_fused_fp64"] = round(res["fp64_fused_cg"]["ms"] / res[best]["ms"], 2)
    res["fused_vs_plain_fp64"] = round(res["fp64_plain_cg"]["ms"] / res["fp64_fused_cg"]["ms"], 2)
    res["workload"] = (f"27-point stencil {side}^3 (n={n}, nnz={len(val)}), diag 26+1e-2, "
                       f"b=U[-1,1) seed 3001, ||b-Ax|| < 1e-10||b||")
    log(f"mixed C3 {side}^3: {res}")
    return res


def bench_batched(m, ctx, nsys, seed0=4000):
    """4096 SPD M-matrices, sizes log-uniform in [500, 20000], 5-15 random
    off-diagonals per row U[-1,0) symmetrised, diag = row sum |offdiag| + 0.5;
    b = U[-1,1) per system; Jacobi-PCG to 1e-10 relative, all systems in the
    same fused launches."""
    import concurrent.futures as cf

    rng = np.random.default_rng(seed0)
    sizes = np.floor(np.exp(rng.uniform(np.log(500), np.log(20000), nsys))).astype(int)
    t0 = time.perf_counter()

    def gen(s):
        rp, ci, v = m.mmatrix(int(sizes[s]), 5, 15, 0.5, seed=seed0 + s)
        b = m.random_vector(int(sizes[s]), -1.0, 1.0, seed0 + s)
        return rp, ci, v, b

    with cf.ThreadPoolExecutor(16) as ex:
        mats = list(ex.map(gen, range(nsys)))
    t_gen = time.perf_counter() - t0
    As = [m.CRS(ctx, len(b), len(b), rp, ci, v) for rp, ci, v, b in mats]
    Bs = [m.Vector(ctx, b) for _, _, _, b in mats]
    Xs = [m.Vector(ctx, n=len(b)) for _, _, _, b in mats]
    nnz = int(sum(len(v) for _, _, v, _ in mats))
    del mats
    ctx.sync()
    t_setup = time.perf_counter() - t0 - t_gen
    # fused batch: every system in shared launches.  One untimed warm-up
    # batch sizes the library's device pool (its first call allocates the
    # work-vector slab); the timed batch then allocates nothing.
    t1 = time.perf_counter()
    m.batch_solve(ctx, m.METHOD_CG, As, Xs, Bs, tol=1e-10, tol_is_relative=True,
                  precond=m.PRECOND_JACOBI)
    t_cold = time.perf_counter() - t1
    for X in Xs:
        X.fill(0.0)
    ctx.sync()
    t1 = time.perf_counter()
    res = m.batch_solve(ctx, m.METHOD_CG, As, Xs, Bs, tol=1e-10, tol_is_relative=True,
                        precond=m.PRECOND_JACOBI)
    t_batch = time.perf_counter() - t1
    ok = sum(r.status == 0 for r in res)
    iters = [r.final_iter for r in res]
    # FP32-refined form: FP64 residuals and stopping test, batched FP32 PCG
    # corrections; error against the FP64 batch on a sample of systems
    x64 = {s: Xs[s].download() for s in range(0, nsys, max(1, nsys // 64))}
    for X in Xs:
        X.fill(0.0)
    ctx.sync()
    t1 = time.perf_counter()
    res_ir = m.batch_ir_cg(ctx, As, Xs, Bs, tol=1e-10, tol_is_relative=True, inner_tol=1e-5,
                           precond=m.PRECOND_JACOBI)
    t_ir = time.perf_counter() - t1
    ok_ir = sum(r.status == 0 for r in res_ir)
    err_ir = max(float(np.linalg.norm(Xs[s].download() - x) / np.linalg.norm(x))
                 for s, x in x64.items())
    # the same systems one at a time: fused single-system and plain FP64 CG
    # (a sample, extrapolated to all systems)
    sample = list(range(0, nsys, max(1, nsys // 32)))
    for X in Xs:
        X.fill(0.0)

    def one(mode):
        t = time.perf_counter()
        for s in sample:
            Xs[s].fill(0.0)
            m.cg(As[s], Xs[s], Bs[s], tol=1e-10, tol_is_relative=True, precond=m.PRECOND_JACOBI,
                 mode=mode)
        dt = time.perf_counter() - t
        return dt / len(sample) * sum(1 for _ in range(nsys))

    t_single = one(m.MODE_FUSED)
    t_plain = one(m.MODE_REFERENCE)
    out = {"workload": f"{nsys} SPD M-matrices, n log-uniform [500,20000] (sum n = {int(sizes.sum())}, "
                       f"nnz = {nnz}), Jacobi-PCG FP64 to 1e-10 relative",
           "batched_solves_per_s": round(nsys / t_batch, 1), "batched_seconds": round(t_batch, 3),
           "first_call_seconds": round(t_cold, 3),
           "converged": ok, "iterations_min_max": [int(min(iters)), int(max(iters))],
           "one_at_a_time_fused_solves_per_s_est": round(nsys / t_single, 1),
           "plain_fp64_gpu_cg_solves_per_s_est": round(nsys / t_plain, 1),
           "batched_vs_plain": round(t_plain / t_batch, 2),
           "fp32_refined": {"solves_per_s": round(nsys / t_ir, 1), "seconds": round(t_ir, 3),
                            "converged": ok_ir,
                            "outer_min_max": [int(min(r.outer_iter for r in res_ir)),
                                              int(max(r.outer_iter for r in res_ir))],
                            "inner_iterations_min_max": [int(min(r.inner_iter for r in res_ir)),
                                                         int(max(r.inner_iter for r in res_ir))],
                            "max_rel_err_vs_fp64_batch": err_ir,
                            "vs_plain": round(t_plain / t_ir, 2)},
           "estimate_sample": f"{len(sample)} systems solved one at a time, scaled to {nsys}",
           "generation_s": round(t_gen, 2), "upload_s": round(t_setup, 2)}
    log(f"batched C4: {out}")
    for A in As:
        A.free()
    return out


def bench_served(m, ctx, rate, seconds, nmin, nmax, seed=5001, pool=48, drain=60.0,
                 cg_mode=None, one_at_a_time=False):
    """Poisson arrivals; 50% 2D/3D Poisson (CG), 50% 2D upwind convection-
    diffusion with Peclet U[1,50] (BiCGSTAB); Jacobi, 1e-10 relative; sizes
    log-uniform in [nmin, nmax].  Requests draw from a pool of systems that
    are uploaded and SpMV-autotuned once (the served matrices and right-hand
    sides stay resident in HBM); each request is submitted by handle, so the client keeps up with
    the arrival rate and queued requests hold no host copies.  Arrivals stop
    after `seconds`; requests still unfinished `drain` seconds later are
    cancelled (a bounded run when the offered load exceeds capacity) and
    reported as such; latency percentiles are over completed requests.
    cg_mode MODE_MIXED_IR serves the CG requests by FP32-refined iterations
    (the BiCGSTAB requests stay FP64)."""
    if cg_mode is None:
        cg_mode = m.MODE_FUSED
    rng = np.random.default_rng(seed)
    systems = []
    nnzs = []
    sliced = 0
    for k in range(pool):
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        if k % 2 == 0:
            if k % 4 == 0:
                side = max(int(round(np.sqrt(n))), 4)
                rp, ci, v = m.laplacian2d(side, side)
            else:
                side = max(int(round(n ** (1 / 3))), 3)
                rp, ci, v = m.stencil3d(side, side, side, 7)
            meth = m.METHOD_CG
        else:
            side = max(int(round(np.sqrt(n))), 4)
            rp, ci, v = m.convdiff2d(side, float(rng.uniform(1, 50)))
            meth = m.METHOD_BICGSTAB
        nn = len(rp) - 1
        b = m.random_vector(nn, -1.0, 1.0, seed + k)
        A = m.CRS(ctx, nn, nn, rp, ci, v)
        # resident systems are tuned once at upload: with the sliced layout
        # the engine's SpMV stages walk it one lane per row
        sliced += A.autotune(reps=3)["row_walk"] == 2
        systems.append((meth, A, m.Vector(ctx, b), nn))
        nnzs.append(len(v))
    # the arrival schedule (times, systems) first; the first request of a
    # few systems keeps its solution, for the check that a result in the mix
    # equals the same system solved alone: those vectors are made before the
    # server exists (its context takes no other calls meanwhile)
    sched = []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > seconds:
            break
        sched.append((t, int(rng.integers(pool))))
    keep = {}
    for i, (_, k) in enumerate(sched):
        if k not in keep and len(keep) < 4:
            keep[k] = (m.Vector(ctx, n=systems[k][3]), i)
    ctx.sync()
    srv = m.Server(ctx, max_active=1024, iters_per_poll=8)
    sys_of = []
    t0 = time.perf_counter()
    tickets = []
    for i, (t, k) in enumerate(sched):
        now = time.perf_counter() - t0
        if t > now:
            time.sleep(t - now)
        meth, A, B, nn = systems[k]
        X = keep[k][0] if k in keep and keep[k][1] == i else None
        mode = cg_mode if meth == m.METHOD_CG else m.MODE_FUSED
        tickets.append(srv.submit_resident(meth, A, B, X, tol=1e-10, tol_is_relative=True,
                                           mode=mode))
        sys_of.append(k)
    t_submitted = time.perf_counter() - t0
    statuses = [None] * len(tickets)
    iters = [0] * len(tickets)
    cutoff = t0 + seconds + drain
    late = []
    for i, tk in enumerate(tickets):
        left = cutoff - time.perf_counter()
        got = srv.wait(tk, timeout=left) if left > 0 else None
        if got is None:
            late.append(i)
        else:
            statuses[i], iters[i] = got[0].status, got[0].final_iter
    # past the cutoff: cancel every unfinished request at once, then collect
    for i in late:
        srv.cancel(tickets[i])
    for i in late:
        got = srv.wait(tickets[i], timeout=60.0)[0]
        statuses[i], iters[i] = got.status, got.final_iter
    wall = time.perf_counter() - t0
    # a floor on the HBM traffic of the work done: the SpMV bytes of every
    # iteration run (one SpMV per CG iteration, two per BiCGSTAB iteration;
    # the inner iterations of FP32-refined CG requests at FP32 sizes),
    # cancelled requests included
    spmv_bytes_done = 0
    for i, k in enumerate(sys_of):
        meth, _, _, nn = systems[k]
        sv = 4 if meth == m.METHOD_CG and cg_mode == m.MODE_MIXED_IR else 8
        per = spmv_bytes(nnzs[k], nn, nn, sv) * (1 if meth == m.METHOD_CG else 2)
        spmv_bytes_done += per * max(iters[i], 0)
    st = srv.stats()
    srv.close()
    # kept requests that converged: the result in the mix equals the lone
    # solve bit for bit (a request cancelled at the cutoff holds a partial
    # iterate and is not compared)
    same = True
    compared = 0
    for k, (X, ti) in keep.items():
        if statuses[ti] != 0:
            X.free()
            continue
        compared += 1
        meth, A, B, nn = systems[k]
        X1 = m.Vector(ctx, n=nn)
        if meth == m.METHOD_CG and cg_mode == m.MODE_MIXED_IR:
            m.ir_cg(A, X1, B, tol=1e-10, tol_is_relative=True, inner_tol=1e-5)
        else:
            fn = m.cg if meth == m.METHOD_CG else m.bicgstab
            fn(A, X1, B, tol=1e-10, tol_is_relative=True, precond=m.PRECOND_JACOBI,
               mode=m.MODE_FUSED)
        same = same and bool(np.array_equal(X1.download().view(np.uint64),
                                            X.download().view(np.uint64)))
        X1.free()
        X.free()
    out = {"workload": f"Poisson arrivals {rate}/s for {seconds}s, n log-uniform [{nmin},{nmax}], "
                       f"50% 2D/3D Poisson (CG) / 50% upwind convection-diffusion Pe U[1,50] "
                       f"(BiCGSTAB), Jacobi, 1e-10 relative; {pool} distinct resident systems; "
                       f"requests unfinished {drain}s after the last arrival are cancelled",
           "requests": len(tickets), "converged": int(sum(s == 0 for s in statuses)),
           "cancelled_at_cutoff": int(sum(s == m.ERR_CANCELLED for s in statuses)),
           "cg_requests": "fp32-refined (FP64 residual, FP32 corrections)"
                          if cg_mode == m.MODE_MIXED_IR else "fp64",
           "solves_per_s": round(st["solves_per_s"], 1),
           "offered_per_s": round(len(tickets) / t_submitted, 1),
           "p50_latency_ms": round(1e3 * st["p50_latency_s"], 2),
           "p99_latency_ms": round(1e3 * st["p99_latency_s"], 2),
           "mean_queue_ms": round(1e3 * st["mean_queue_s"], 2), "wall_s": round(wall, 2),
           "spmv_GB_s_floor": round(spmv_bytes_done / wall / 1e9, 1),
           "pool_sliced_layout": int(sliced),
           "peak_resident_MB": round(st["peak_bytes"] / 2 ** 20, 1),
           "result_in_mix_equals_alone": same if compared else None,
           "results_compared": compared}
    if one_at_a_time:
        # the same request mix (uniform over the pool) served one request at a
        # time: each pool system solved alone by the plain FP64 GPU CG /
        # BiCGSTAB (the reference's call sequence) and by the fused solver
        per = {}
        for mode, key in ((m.MODE_REFERENCE, "plain_fp64_gpu"), (m.MODE_FUSED, "fused")):
            ts = []
            for meth, A, B, nn in systems:
                X1 = m.Vector(ctx, n=nn)
                fn = m.cg if meth == m.METHOD_CG else m.bicgstab
                ctx.sync()
                t1 = time.perf_counter()
                fn(A, X1, B, tol=1e-10, tol_is_relative=True, precond=m.PRECOND_JACOBI, mode=mode)
                ctx.sync()
                ts.append(time.perf_counter() - t1)
                X1.free()
            per[key] = len(ts) / sum(ts)
        out["plain_fp64_gpu_one_at_a_time_solves_per_s"] = round(per["plain_fp64_gpu"], 2)
        out["fused_one_at_a_time_solves_per_s"] = round(per["fused"], 2)
        out["served_vs_plain_one_at_a_time"] = round(out["solves_per_s"] / per["plain_fp64_gpu"], 2)
    log(f"served C5: {out}")
    for _, A, B, _ in systems:
        A.free()
        B.free()
    return out


def bench_ingest(m, ctx, side, with_cpu):
    """MatrixMarket -> device CRS (COO::input_mm + CRS(COO&), optionally
    COO::sort(true)) on the 3D 7-point side^3 matrix: the file is written by
    the library's output_mm, read back with sort = 0 and sort = 2, and the
    device conversion is timed alone from shuffled host triplets."""
    import shutil
    import tempfile
    rp, ci, val = m.stencil3d(side, side, side, 7)
    n, nnz = len(rp) - 1, len(val)
    A = m.CRS(ctx, n, n, rp, ci, val)
    d = tempfile.mkdtemp(prefix="mgx_mm_")
    out = {"workload": f"3D 7-point {side}^3 (n={n}, nnz={nnz}) as a MatrixMarket file written "
                       f"by output_mm (%.17e values)"}
    try:
        path = os.path.join(d, "a.mtx")
        t0 = time.perf_counter()
        A.write_mm(path)
        t_write = time.perf_counter() - t0
        size = os.path.getsize(path)

        def timed(fn):
            ctx.sync()
            t = time.perf_counter()
            B = fn()
            ctx.sync()
            return B, time.perf_counter() - t

        B0, t0s = timed(lambda: m.CRS.read_mm(ctx, path, sort=0))
        B2, t2s = timed(lambda: m.CRS.read_mm(ctx, path, sort=2))
        rows_coo = np.repeat(np.arange(n, dtype=np.int32), np.diff(rp))
        perm = np.random.default_rng(6001).permutation(nnz)
        r_s, c_s, v_s = rows_coo[perm], ci[perm], val[perm]
        Bc, tcs = timed(lambda: m.CRS.from_coo(ctx, n, n, r_s, c_s, v_s, sort=2))
        same = True
        for B in (B0, B2, Bc):
            brp, bci, bv = B.download()
            same = same and np.array_equal(brp, rp) and np.array_equal(bci, ci) and \
                np.array_equal(bv.view(np.uint64), val.view(np.uint64))
            B.free()
        out.update({
            "file_MB": round(size / 1e6, 1), "nnz": nnz,
            "write_mm_s": round(t_write, 3),
            "read_mm_sort0_s": round(t0s, 3),
            "read_mm_sort0_Mentries_per_s": round(nnz / t0s / 1e6, 1),
            "read_mm_sort0_MB_per_s": round(size / t0s / 1e6, 1),
            "read_mm_sort2_s": round(t2s, 3),
            "from_coo_shuffled_sort2_s": round(tcs, 3),
            "from_coo_shuffled_sort2_Mentries_per_s": round(nnz / tcs / 1e6, 1),
            "identical_to_source": same})
        if with_cpu:
            out["cpu"] = cpu_baseline_ingest(path, r_s, c_s, v_s, n)
    finally:
        A.free()
        shutil.rmtree(d, ignore_errors=True)
    log(f"ingest: {out}")
    return out


def cpu_baseline_ingest(path, r_s, c_s, v_s, n):
    """The oracle's COO::input_mm + CRS(COO&) on the same file, and its
    COO::sort(true) on a 1M-entry sample of the shuffled triplets (the
    reference's quicksort takes the last element as pivot)."""
    import oracle
    t = time.perf_counter()
    rows, cols, r, c, v = oracle.input_mm(path)
    t_read = time.perf_counter() - t
    t = time.perf_counter()
    oracle.crs_from_coo(rows, cols, r, c, v)
    t_conv = time.perf_counter() - t
    k = min(len(r_s), 1_000_000)
    t = time.perf_counter()
    oracle.coo_sort(r_s[:k], c_s[:k], v_s[:k], merge=True)
    t_sort = time.perf_counter() - t
    return {"input_mm_s": round(t_read, 3), "crs_from_coo_s": round(t_conv, 3),
            "Mentries_per_s": round(len(v) / (t_read + t_conv) / 1e6, 2),
            "coo_sort_merge_1M_s": round(t_sort, 3), "threads": 1, "kind": "port",
            "sample": "full file for input_mm + CRS(COO&); 1M shuffled entries for the sort"}


def cpu_baseline_cg(rp, ci, val, b, tol, x_gpu):
    """The oracle's restatement of the reference CG loop on the C1 inputs,
    timed on this host and compared with the GPU's fused solution."""
    import oracle
    n = len(rp) - 1
    t0 = time.perf_counter()
    ref = oracle.solve("cg", rp, ci, val, np.zeros(n), b, precond=oracle.PRECOND_JACOBI, tol=tol)
    t_cpu = time.perf_counter() - t0
    return {"solves_per_s": round(1.0 / t_cpu, 2), "iterations": ref[2],
            "threads": oracle.num_threads(), "kind": "port",
            "gpu_bit_identical": bool(np.array_equal(ref[1].view(np.uint64),
                                                      x_gpu.view(np.uint64)))}


def cpu_baseline_spmv(budget_s):
    """The oracle's OpenMP SpMV (the reference's CRS row loop) on the full
    192^3 7-point FP64 matrix, in a child process that pins its threads
    (physical cores, OMP_PROC_BIND=close, OMP_PLACES=cores) before libgomp
    loads; median of >= 5 SpMVs after one warm-up."""
    r = subprocess.run([sys.executable, "-m", "oracle.cpu_baseline", "--matrix", "7pt192",
                        "--seconds", str(budget_s)], capture_output=True, text=True, cwd=ROOT,
                       timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline failed: {r.stderr[-2000:]}")
    out = json.loads(r.stdout.strip().splitlines()[-1])
    log(f"cpu baseline: {out}")
    return out


def spmv_kernel_name(plan, dtype) -> str:
    """The kernel that runs the SpMV under an autotuned plan, as rocprofv3
    names it (without the argument list)."""
    f64 = np.dtype(dtype) == np.float64
    t = "double" if f64 else "float"
    w = 16 // np.dtype(dtype).itemsize
    if plan["row_walk"] == 2:
        return f"k_spmv_sell<{t}, false, {w}>"
    if plan["column_panels"]:
        return f"k_spmv_panel<{t}"
    u = (plan["nnz_per_block"] + w - 1 + 256 * w - 1) // (256 * w)
    u = next(k for k in (1, 2, 3, 4, 5, 6, 8, 9, 12, 17) if u <= k)
    return f"k_spmv<{t}, false, {u}, {plan['row_walk']}>"


def parse_plan(spec: str):
    """'row_walk,rows_per_block,nnz_per_block,xcd_remap,column_panels'."""
    w, r, z, x, p = (int(v) for v in spec.split(","))
    return dict(row_walk=w, rows_per_block=r, nnz_per_block=z, xcd_remap=x, column_panels=p)


def traffic_for(workload: str, kernel: str, plan: dict):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 counter summary of the same kernel under the same XCD order and
    column-panel count, or None when that plan was not profiled."""
    path = os.path.join(ROOT, "profiles", "spmv_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for e in json.load(f).get("entries", []):
            ep = e.get("plan", {})
            if (e.get("workload") == workload and e.get("kernel") == kernel
                    and ep.get("xcd_remap") == plan.get("xcd_remap")
                    and ep.get("column_panels") == plan.get("column_panels")
                    and ep.get("row_walk") == plan.get("row_walk")):
                return e.get("hbm_bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--quick", action="store_true", help="skip the sweep, Krylov and CPU legs")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--only", default="", help="comma list of legs: cpu,sweep,krylov,mixed,"
                                                 "batched,served,ingest,sor,eigen (default: all)")
    ap.add_argument("--batch-systems", type=int, default=4096)
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--serve-nmax", type=int, default=1000000,
                    help="largest served system (rows)")
    ap.add_argument("--serve-drain", type=float, default=60.0,
                    help="seconds after the last arrival before unfinished requests are "
                         "cancelled (the default lets the overloaded stream finish, so "
                         "its solves/s is the capacity for the whole mix)")
    ap.add_argument("--plan", default="",
                    help="fix the headline SpMV plan instead of autotuning: "
                         "'row_walk,rows_per_block,nnz_per_block,xcd_remap,column_panels' "
                         "(row_walk 2 = sliced layout); for profiling runs")
    ap.add_argument("--traffic-probe", default="",
                    help="c2a|c2b[:f32]: only run 5 vadd launches of known size (counter "
                         "calibration), then 5 SpMVs on that matrix under --plan (else the "
                         "autotuned plan), for rocprofv3 --pmc runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "runs on one GPU only: the row-sharded multi-GPU form is "
                                   "designed but not launched", "n_gpus": args.gpus}))
        sys.exit(2)

    import monolish_amd as m

    ev = HipEvents()
    ctx = m.Context(0)
    info = ctx.info()
    log("device", info)

    if args.traffic_probe:
        n = 1 << 25  # vadd: reads 2*n*8 bytes, writes n*8 bytes per launch
        a = m.Vector(ctx, np.ones(n))
        b = m.Vector(ctx, np.ones(n))
        y = m.Vector(ctx, n=n)
        for _ in range(5):
            m.vadd(a, b, y)
        ctx.sync()
        for v in (a, b, y):
            v.free()
        mat, _, prec = args.traffic_probe.partition(":")
        dt = np.float32 if prec == "f32" else np.float64
        if mat == "c2a":
            rp, ci, val = m.stencil3d(192, 192, 192, 7)
        else:
            rp, ci, val = m.random_lognormal(4_000_000, 4_000_000, 32.0, 1.0, 2048, seed=2002)
        val = val.astype(dt)
        rows = len(rp) - 1
        A = m.CRS(ctx, rows, rows, rp, ci, val)
        X = m.Vector(ctx, m.random_vector(rows, -1.0, 1.0, 2001).astype(dt), dtype=dt)
        Y = m.Vector(ctx, n=rows, dtype=dt)
        if args.plan:
            plan = parse_plan(args.plan)
            A.set_plan(**plan)
        else:
            plan = A.autotune(reps=20)
        ctx.sync()
        for _ in range(5):
            m.matvec(A, X, Y)
        ctx.sync()
        print(json.dumps({"vadd_read_bytes": 2 * n * 8, "vadd_write_bytes": n * 8,
                          "spmv_alg_bytes": spmv_bytes(len(val), rows, rows, dt().itemsize),
                          "kernel": spmv_kernel_name(plan, dt), "plan": plan}))
        ctx.close()
        return

    # ---- headline: SpMV, 3D 7-point 192^3, FP64 ---------------------------
    rp, ci, val = m.stencil3d(192, 192, 192, 7)
    x = m.random_vector(len(rp) - 1, -1.0, 1.0, 2001)
    head, ms, wall = bench_spmv_matrix(m, ev, ctx, "C2a fp64", rp, ci, val, x, args.steps,
                                       args.warmup, plan_spec=args.plan or None)
    sweep = {"c2a_7pt_192_fp64": head}
    # the kernel the chosen plan runs, and its HBM traffic from the committed
    # rocprofv3 FETCH_SIZE / WRITE_SIZE summary of that same kernel
    traffic = traffic_for("c2a_fp64", head["kernel"], head["plan"])
    roofline = {"bound": "hbm", "achieved": head["GB_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(head["GB_s"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": head["kernel"], "plan": head["plan"],
                "bytes_per_launch": head["bytes"]}
    line = {
        "metric": "CG solves/s @1e-10 + p99 latency vs plain FP64 GPU CG; SpMV HBM GB/s",
        "value": head["GB_s"], "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (generated matrices and seeded U[-1,1) vectors)",
        "config": {"workload": "SpMV (blas::matvec) on the 3D 7-point Laplacian 192^3, FP64, "
                               "CRS int32 indices, n=7077888, nnz=49324032",
                   "timed_region_s": round(wall, 3)},
        "roofline": roofline,
    }
    legs = set(args.only.split(",")) if args.only else {"cpu", "sweep", "krylov", "mixed", "batched",
                                                         "served", "ingest", "sor", "eigen"}
    if args.quick:
        legs = set()
    if "cpu" in legs:
        line["cpu_baseline"] = cpu_baseline_spmv(args.cpu_seconds)
    if "sweep" in legs:
        r32, _, _ = bench_spmv_matrix(m, ev, ctx, "C2a fp32", rp, ci, val.astype(np.float32),
                                      x.astype(np.float32), args.steps, args.warmup)
        sweep["c2a_7pt_192_fp32"] = r32
        del rp, ci, val, x
        n2 = 4_000_000
        rp2, ci2, v2 = m.random_lognormal(n2, n2, 32.0, 1.0, 2048, seed=2002)
        x2 = m.random_vector(n2, -1.0, 1.0, 2002)
        s64, _, _ = bench_spmv_matrix(m, ev, ctx, "C2b fp64", rp2, ci2, v2, x2, max(args.steps // 4, 10),
                                      args.warmup)
        s32, _, _ = bench_spmv_matrix(m, ev, ctx, "C2b fp32", rp2, ci2, v2.astype(np.float32),
                                      x2.astype(np.float32), max(args.steps // 4, 10), args.warmup)
        sweep["c2b_lognormal_4M_fp64"] = s64
        sweep["c2b_lognormal_4M_fp32"] = s32
        for key, wl in (("c2a_7pt_192_fp32", "c2a_fp32"), ("c2b_lognormal_4M_fp64", "c2b_fp64"),
                        ("c2b_lognormal_4M_fp32", "c2b_fp32")):
            sweep[key]["traffic"] = traffic_for(wl, sweep[key]["kernel"], sweep[key]["plan"])
        del rp2, ci2, v2, x2
        line["spmv_sweep"] = sweep
    if "krylov" in legs:
        line["krylov"], xf, tol = bench_krylov(m, ctx, reps_fused=30, reps_plain=5)
        if "cpu" in legs:
            rp1, ci1, val1, b1 = c1_problem(m)
            line["krylov"]["cpu_cg"] = cpu_baseline_cg(rp1, ci1, val1, b1, tol, xf)
    if "mixed" in legs:
        line["mixed_precision"] = bench_mixed(m, ctx, 128, reps=5)
    if "batched" in legs:
        line["batched"] = bench_batched(m, ctx, args.batch_systems)
    if "served" in legs:
        line["served"] = bench_served(m, ctx, 2000.0, args.serve_seconds, 1000, args.serve_nmax,
                                      drain=args.serve_drain)
        # the same mix at about 70 % of the measured capacity: latency without
        # the queue that an overloaded arrival rate builds
        cap = line["served"]["solves_per_s"]
        line["served"]["below_capacity"] = bench_served(m, ctx, round(0.7 * cap, 1),
                                                        args.serve_seconds, 1000,
                                                        args.serve_nmax, seed=5002,
                                                        drain=args.serve_drain)
        # a half-second window of the same stream, drained: every request
        # completes, so its solves/s is the capacity for the uniform mix,
        # set against the same mix solved one request at a time
        line["served"]["drained_window"] = bench_served(m, ctx, 2000.0,
                                                        min(args.serve_seconds, 0.5), 1000,
                                                        args.serve_nmax, drain=args.serve_drain,
                                                        one_at_a_time=True)
        # the same stream with the CG requests refined in FP32 (a half-second
        # window keeps the default run short)
        line["served"]["fp32_refined_cg"] = bench_served(m, ctx, 2000.0,
                                                         min(args.serve_seconds, 0.5), 1000,
                                                         args.serve_nmax, drain=args.serve_drain,
                                                         cg_mode=m.MODE_MIXED_IR)
    if "ingest" in legs:
        line["ingest"] = bench_ingest(m, ctx, 128, "cpu" in legs)
    if "sor" in legs:
        line["sor"] = bench_sor(m, ctx, 3, "cpu" in legs)
    if "eigen" in legs:
        line["eigen"] = bench_eigen(m, ctx, "cpu" in legs)
    line["device"] = info["name"]
    print(json.dumps(line))
    ctx.close()


if __name__ == "__main__":
    main()
