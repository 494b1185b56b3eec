"""GPU parity of the fused BiCGSTAB, the batched solver and the solve server.
Every system carries its own tile-order reductions, so its iterates are the
same bits alone, in a batch, in the server's mix, and in the oracle."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def m():
    import monolish_amd
    return monolish_amd


@pytest.fixture(scope="module")
def ctx(m):
    c = m.Context(0)
    yield c
    c.close()


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def systems(m, oracle):
    out = []
    for k, n in enumerate((500, 1333, 4000, 9000)):
        rp, ci, v = m.mmatrix(n, 5, 15, 0.5, seed=4000 + k)
        out.append(("cg", rp, ci, v, oracle.random_vector(n, -1.0, 1.0, 4000 + k)))
    for side in (30, 64):
        rp, ci, v = m.laplacian2d(side, side)
        out.append(("cg", rp, ci, v, oracle.random_vector(side * side, -1.0, 1.0, 1001)))
    rp, ci, v = m.stencil3d(14, 15, 16, 27, 1e-2)
    out.append(("cg", rp, ci, v, oracle.random_vector(len(rp) - 1, -1.0, 1.0, 3001)))
    for side, pe in ((40, 5.0), (70, 45.0)):
        rp, ci, v = m.convdiff2d(side, pe)
        out.append(("bicgstab", rp, ci, v, oracle.random_vector(side * side, -1.0, 1.0, 5001)))
    return out


def oracle_solve(oracle, method, rp, ci, v, b):
    tol = 1e-10 * oracle.nrm2(b)
    return oracle.solve(method, rp, ci, v, np.zeros(len(b)), b, precond=2, tol=tol)


def test_mmatrix_generator_is_spd_mmatrix(m):
    say("mmatrix generator")
    import scipy.sparse as sp
    rp, ci, v = m.mmatrix(800, 5, 15, 0.5, seed=7)
    A = sp.csr_matrix((v, ci, rp), shape=(800, 800))
    assert abs(A - A.T).max() == 0
    d = A.diagonal()
    off = A - sp.diags(d)
    assert off.max() <= 0
    assert np.allclose(d, -np.asarray(off.sum(axis=1)).ravel() + 0.5)
    for i in range(800):
        assert np.all(np.diff(ci[rp[i]:rp[i + 1]]) > 0)


@pytest.mark.parametrize("side,pe", [(50, 20.0), (33, 1.0)])
def test_bicgstab_fused_bit_exact(m, ctx, oracle, side, pe):
    say("bicgstab fused", side, pe)
    rp, ci, v = m.convdiff2d(side, pe)
    n = side * side
    b = oracle.random_vector(n, -1.0, 1.0, 5001)
    ref = oracle.solve("bicgstab", rp, ci, v, np.zeros(n), b, precond=2,
                       tol=1e-10 * oracle.nrm2(b), history=True)
    for mode in (0, 1):
        A = m.CRS(ctx, n, n, rp, ci, v)
        X = m.Vector(ctx, n=n)
        B = m.Vector(ctx, b)
        res = m.bicgstab(A, X, B, tol=1e-10, tol_is_relative=True, precond=2, mode=mode, history=5000)
        assert res.status == ref[0] == 0 and res.final_iter == ref[2], mode
        assert np.array_equal(bits(X.download()), bits(ref[1])), mode
        assert np.array_equal(bits(res.rhistory), bits(ref[4])), mode


def test_bicgstab_fused_breakdown_and_maxiter(m, ctx, oracle):
    say("bicgstab fused error paths")
    rp, ci, v = m.convdiff2d(20, 10.0)
    n = 400
    xe = oracle.random_vector(n, -1.0, 1.0, 2)
    be = oracle.spmv(rp, ci, v, xe)
    A = m.CRS(ctx, n, n, rp, ci, v)
    X = m.Vector(ctx, xe)
    res = m.bicgstab(A, X, m.Vector(ctx, be), tol=1e-8, precond=2, mode=1)
    assert res.status == m.SOLVER_BREAKDOWN and res.final_iter == 1
    b = oracle.random_vector(n, -1.0, 1.0, 3)
    ref = oracle.solve("bicgstab", rp, ci, v, np.zeros(n), b, precond=2, tol=1e-30, maxiter=6)
    X = m.Vector(ctx, n=n)
    res = m.bicgstab(A, X, m.Vector(ctx, b), tol=1e-30, maxiter=6, precond=2, mode=1)
    assert res.status == m.SOLVER_MAXITER and res.final_iter == 6 == ref[2]
    assert res.final_resid == ref[3] and np.array_equal(bits(X.download()), bits(ref[1]))


def test_batch_identical_to_alone_and_oracle(m, ctx, oracle):
    """One batch mixes scattered M-matrices (p formed first, gathered once
    per entry) and stencils (z and p gathered, p formed at each gather);
    every system reproduces the oracle's CG / BiCGSTAB bit for bit."""
    say("batch vs alone")
    syss = systems(m, oracle)
    scattered = [m.CRS(ctx, len(b), len(b), rp, ci, v).stats()["gather_lines"] > 24
                 for _, rp, ci, v, b in syss]
    assert scattered == [True] * 4 + [False] * 5
    for method in ("cg", "bicgstab"):
        group = [s for s in syss if s[0] == method]
        As, Xs, Bs = [], [], []
        for _, rp, ci, v, b in group:
            n = len(b)
            As.append(m.CRS(ctx, n, n, rp, ci, v))
            Xs.append(m.Vector(ctx, n=n))
            Bs.append(m.Vector(ctx, b))
        code = m.METHOD_CG if method == "cg" else m.METHOD_BICGSTAB
        res = m.batch_solve(ctx, code, As, Xs, Bs, tol=1e-10, tol_is_relative=True, precond=2)
        for (meth, rp, ci, v, b), r, X in zip(group, res, Xs):
            ref = oracle_solve(oracle, meth, rp, ci, v, b)
            say(meth, len(b), r.final_iter, ref[2])
            assert r.status == ref[0] == 0 and r.final_iter == ref[2]
            assert np.array_equal(bits(X.download()), bits(ref[1]))


def test_server_mix_equals_alone(m, ctx, oracle):
    """Requests of different sizes and methods share fused launches, join and
    leave at different iterations, and queue behind a small HBM budget; each
    result is bit-identical to the oracle's solve of that system alone."""
    say("server mix")
    syss = systems(m, oracle) * 3
    c2 = m.Context(0)
    srv = m.Server(c2, hbm_budget=24 << 20, max_active=6, iters_per_poll=4)
    tickets = []
    for meth, rp, ci, v, b in syss:
        code = m.METHOD_CG if meth == "cg" else m.METHOD_BICGSTAB
        tickets.append(srv.submit(code, rp, ci, v, b, tol=1e-10, tol_is_relative=True))
    for t, (meth, rp, ci, v, b) in zip(tickets, syss):
        r, x = srv.wait(t, len(b), timeout=120.0)
        ref = oracle_solve(oracle, meth, rp, ci, v, b)
        assert r.status == 0 and r.final_iter == ref[2]
        assert np.array_equal(bits(x), bits(ref[1]))
    st = srv.stats()
    say("stats", st)
    assert st["completed"] == len(syss) and st["queued"] == 0
    assert 0 < st["p50_latency_s"] <= st["p99_latency_s"] <= st["max_latency_s"]
    assert st["peak_bytes"] <= 24 << 20
    # a request that can never fit the budget is refused, not queued forever
    rp, ci, v = m.laplacian2d(1200, 1200)
    t = srv.submit(m.METHOD_CG, rp, ci, v, np.ones(1200 * 1200))
    r, _ = srv.wait(t, None, timeout=60.0)
    assert r.status == m.ERR_BUDGET
    srv.close()
    c2.close()


def test_batch_ir_cg_equals_single_ir(m, ctx, oracle):
    """FP32-refined batched CG: every system follows exactly the arithmetic
    of the single-system mixed-precision solve (same outer and inner
    iteration counts, same x bit for bit) and reaches the FP64 tolerance."""
    say("batch ir")
    syss = [m.mmatrix(n, 5, 15, 0.5, seed=70 + n) for n in (700, 1500, 3100, 5000)]
    bs = [oracle.random_vector(len(rp) - 1, -1.0, 1.0, 71 + i) for i, (rp, _, _) in enumerate(syss)]
    As = [m.CRS(ctx, len(b), len(b), rp, ci, v) for (rp, ci, v), b in zip(syss, bs)]
    Bs = [m.Vector(ctx, b) for b in bs]
    Xs = [m.Vector(ctx, n=len(b)) for b in bs]
    res = m.batch_ir_cg(ctx, As, Xs, Bs, tol=1e-10, tol_is_relative=True, inner_tol=1e-5)
    for A, X, B, (rp, ci, v), b, r in zip(As, Xs, Bs, syss, bs, res):
        assert r.status == 0 and r.outer_iter >= 1
        x = X.download()
        assert oracle.residual_l2(rp, ci, v, x, b) < 1e-10 * oracle.nrm2(b)
        X1 = m.Vector(ctx, n=len(b))
        r1 = m.ir_cg(A, X1, B, tol=1e-10, tol_is_relative=True, inner_tol=1e-5)
        say(len(b), r.outer_iter, r.inner_iter, r1.outer_iter, r1.inner_iter)
        assert (r.outer_iter, r.inner_iter) == (r1.outer_iter, r1.inner_iter)
        assert r.final_resid == r1.final_resid
        assert np.array_equal(bits(x), bits(X1.download()))


def test_server_resident_requests(m, ctx, oracle):
    """Requests by handle: several requests share one resident matrix and
    right-hand side; those given an x receive the solution, bit-identical to
    the oracle's solve of the system alone; requests without x only count."""
    say("server resident")
    syss = systems(m, oracle)
    c2 = m.Context(0)
    res_sys = []
    for meth, rp, ci, v, b in syss:
        n = len(b)
        res_sys.append((meth, m.CRS(c2, n, n, rp, ci, v), m.Vector(c2, b), rp, ci, v, b))
    srv = m.Server(c2, max_active=5, iters_per_poll=3)
    tickets = []
    for rep in range(3):
        for i, (meth, A, B, rp, ci, v, b) in enumerate(res_sys):
            code = m.METHOD_CG if meth == "cg" else m.METHOD_BICGSTAB
            X = m.Vector(c2, n=len(b)) if rep == 1 else None
            tickets.append((srv.submit_resident(code, A, B, X, tol=1e-10, tol_is_relative=True),
                            i, X))
    for t, i, X in tickets:
        r, _ = srv.wait(t, timeout=120.0)
        meth, A, B, rp, ci, v, b = res_sys[i]
        ref = oracle_solve(oracle, meth, rp, ci, v, b)
        assert r.status == 0 and r.final_iter == ref[2]
        if X is not None:
            assert np.array_equal(bits(X.download()), bits(ref[1]))
    st = srv.stats()
    assert st["completed"] == len(tickets) and st["queued"] == 0
    srv.close()
    for t, i, X in tickets:
        if X is not None:
            X.free()
    for _, A, B, *_ in res_sys:
        A.free()
        B.free()
    c2.close()
