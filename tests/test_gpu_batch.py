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


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batch_sliced_layout_systems(m, ctx, oracle, dtype):
    """Gather-once batch systems walk a sliced layout one lane per row in
    their SpMV stages (CG and both BiCGSTAB stages): the plan's sliced copy
    (system 0, 2), or the engine's own when the plan has none (system 1);
    M-matrices, whose 64-row slices would pad by about 50 %, stage products
    in LDS (the last system of each group).  Every system reproduces the
    oracle's solve bit for bit, sizes with a partial last 128-row quarter
    included."""
    say("batch sliced", dtype.__name__)
    cases = []
    for nx, ny, nz in ((33, 29, 17), (40, 40, 40)):
        rp, ci, v = m.stencil3d(nx, ny, nz, 7, 1e-2)
        cases.append(("cg", rp, ci, v))
    rp, ci, v = m.laplacian2d(70, 61)
    cases.append(("cg", rp, ci, v))
    cases.append(("cg",) + tuple(m.mmatrix(3001, 5, 15, 0.5, seed=77)))
    for side, pe in ((45, 7.0), (90, 30.0)):
        rp, ci, v = m.convdiff2d(side, pe)
        cases.append(("bicgstab", rp, ci, v))
    cases.append(("bicgstab",) + tuple(m.mmatrix(2500, 5, 15, 0.5, seed=78)))
    for method in ("cg", "bicgstab"):
        group = [c for c in cases if c[0] == method]
        As, Xs, Bs, refs = [], [], [], []
        for k, (_, rp, ci, v) in enumerate(group):
            n = len(rp) - 1
            v = v.astype(dtype)
            b = oracle.random_vector(n, -1.0, 1.0, 700 + k, dtype=dtype)
            A = m.CRS(ctx, n, n, rp, ci, v)
            if k != 1 and k != len(group) - 1:
                A.set_plan(row_walk=2)  # system 1 keeps a block plan
            As.append(A)
            Xs.append(m.Vector(ctx, n=n, dtype=dtype))
            Bs.append(m.Vector(ctx, b, dtype=dtype))
            # FP64 to 1e-10 relative; FP32 60 iterations (no tolerance reached)
            kw = dict(tol=1e-10 * oracle.nrm2(b), maxiter=400) if dtype == np.float64 else \
                dict(tol=1e-30, maxiter=60)
            refs.append(oracle.solve(method, rp, ci, v, np.zeros(n, dtype), b, precond=2, **kw))
        code = m.METHOD_CG if method == "cg" else m.METHOD_BICGSTAB
        if dtype == np.float64:
            res = m.batch_solve(ctx, code, As, Xs, Bs, tol=1e-10, tol_is_relative=True, precond=2,
                                maxiter=400)
        else:
            res = m.batch_solve(ctx, code, As, Xs, Bs, tol=1e-30, precond=2, maxiter=60)
        for r, X, ref in zip(res, Xs, refs):
            say(method, r.status, r.final_iter, ref[2])
            assert r.final_iter == ref[2]
            if dtype == np.float64:
                assert r.status == ref[0] == 0
            assert np.array_equal(bits(X.download()), bits(ref[1]))
        for o in As + Xs + Bs:
            o.free()


def test_batch_with_split_form_systems(m, ctx, oracle):
    """Two CG systems of at least 2^19 rows (the split form, launched one
    system at a time) in one batch with small systems: one with the sliced
    SpMV forming the <p, q> partials, one with the block-plan SpMV and a
    separate <p, q> pass, a partial last tile in both; 30 iterations of
    every system reproduce the oracle's bits."""
    say("batch split form")
    syss = []
    for (nx, ny, nz, shift, seed) in ((96, 96, 64, 0.0, 21), (100, 90, 60, 1e-2, 22)):
        rp, ci, v = m.stencil3d(nx, ny, nz, 7, shift)
        syss.append((rp, ci, v, oracle.random_vector(len(rp) - 1, -1.0, 1.0, seed)))
    assert all(len(b) >= 2 ** 19 for *_, b in syss) and len(syss[1][3]) % 512 != 0
    for _, rp, ci, v, b in systems(m, oracle)[3:5]:
        syss.append((rp, ci, v, b))
    As, Xs, Bs = [], [], []
    for k, (rp, ci, v, b) in enumerate(syss):
        n = len(b)
        A = m.CRS(ctx, n, n, rp, ci, v)
        if k == 0:
            A.set_plan(row_walk=2)
        else:
            A.set_plan(rows_per_block=256, nnz_per_block=2048)
        As.append(A)
        Xs.append(m.Vector(ctx, n=n))
        Bs.append(m.Vector(ctx, b))
    res = m.batch_solve(ctx, m.METHOD_CG, As, Xs, Bs, tol=1e-30, maxiter=30, precond=2)
    for (rp, ci, v, b), r, X in zip(syss, res, Xs):
        ref = oracle.solve("cg", rp, ci, v, np.zeros(len(b)), b, precond=2, tol=1e-30,
                           maxiter=30)
        say(len(b), r.status, r.final_iter, ref[2])
        assert r.final_iter == ref[2] == 30
        assert np.array_equal(bits(X.download()), bits(ref[1]))
    for v in As + Xs + Bs:
        v.free()


def test_c4_full_size_batch(m, ctx, oracle):
    """The full C4 batch: 4096 SPD M-matrices, sizes log-uniform in
    [500, 20000] (seed 4000), 5-15 random off-diagonals per row symmetrised,
    diagonal = row sum |offdiag| + 0.5, b = U[-1,1) per system; all solved in
    one batched Jacobi-PCG to 1e-10 relative.  Every system converges, and
    64 systems spread over the batch reproduce the oracle's lone solve bit
    for bit (iterations and x)."""
    import concurrent.futures as cf
    say("c4 full size")
    nsys, seed0 = 4096, 4000
    rng = np.random.default_rng(seed0)
    sizes = np.floor(np.exp(rng.uniform(np.log(500), np.log(20000), nsys))).astype(int)

    def gen(s):
        rp, ci, v = m.mmatrix(int(sizes[s]), 5, 15, 0.5, seed=seed0 + s)
        return rp, ci, v, m.random_vector(int(sizes[s]), -1.0, 1.0, seed0 + s)

    with cf.ThreadPoolExecutor(8) as ex:
        mats = list(ex.map(gen, range(nsys)))
    As = [m.CRS(ctx, len(b), len(b), rp, ci, v) for rp, ci, v, b in mats]
    Bs = [m.Vector(ctx, b) for *_, b in mats]
    Xs = [m.Vector(ctx, n=len(b)) for *_, b in mats]
    res = m.batch_solve(ctx, m.METHOD_CG, As, Xs, Bs, tol=1e-10, tol_is_relative=True,
                        precond=m.PRECOND_JACOBI)
    assert all(r.status == 0 for r in res)
    for s in range(0, nsys, nsys // 64):
        rp, ci, v, b = mats[s]
        ref = oracle_solve(oracle, "cg", rp, ci, v, b)
        assert res[s].final_iter == ref[2], s
        assert np.array_equal(bits(Xs[s].download()), bits(ref[1])), s
    for o in As + Bs + Xs:
        o.free()


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
    # the solution vectors exist before the server and are read after it
    # closes: while a server runs, its context takes no other calls
    xs = [m.Vector(c2, n=len(b)) for *_, b in res_sys]
    srv = m.Server(c2, max_active=5, iters_per_poll=3)
    tickets = []
    for rep in range(3):
        for i, (meth, A, B, rp, ci, v, b) in enumerate(res_sys):
            code = m.METHOD_CG if meth == "cg" else m.METHOD_BICGSTAB
            X = xs[i] if rep == 1 else None
            tickets.append((srv.submit_resident(code, A, B, X, tol=1e-10, tol_is_relative=True),
                            i, X))
    results = [(srv.wait(t, timeout=120.0)[0], i, X) for t, i, X in tickets]
    st = srv.stats()
    srv.close()
    for r, i, X in results:
        meth, A, B, rp, ci, v, b = res_sys[i]
        ref = oracle_solve(oracle, meth, rp, ci, v, b)
        assert r.status == 0 and r.final_iter == ref[2]
        if X is not None:
            assert np.array_equal(bits(X.download()), bits(ref[1]))
    assert st["completed"] == len(tickets) and st["queued"] == 0
    for X in xs:
        X.free()
    for _, A, B, *_ in res_sys:
        A.free()
        B.free()
    c2.close()


def test_graph_replay_bit_identical(m, oracle):
    """mgx_set_graphs(ctx, 1): the enqueued iterations are captured once per
    tile-list change and replayed, with host polls (eager work) between the
    replays.  A lone fused CG (two-kernel form), a lone split-form CG
    (n >= 2^19, plan-based SpMV in the graph), a CG batch and a BiCGSTAB
    batch give the same bits as with graphs off and as the oracle, on two
    solves in a row."""
    say("graphs")
    c = m.Context(0)
    try:
        syss = systems(m, oracle)

        def run_batch(method):
            group = [s for s in syss if s[0] == method]
            As = [m.CRS(c, len(b), len(b), rp, ci, v) for _, rp, ci, v, b in group]
            Xs = [m.Vector(c, n=len(b)) for *_, b in group]
            Bs = [m.Vector(c, b) for *_, b in group]
            code = m.METHOD_CG if method == "cg" else m.METHOD_BICGSTAB
            res = m.batch_solve(c, code, As, Xs, Bs, tol=1e-10, tol_is_relative=True, precond=2)
            return [(r.status, r.final_iter, bits(X.download()).copy()) for r, X in zip(res, Xs)]

        def run_single(rp, ci, v, b, every, tol, maxiter):
            n = len(b)
            A = m.CRS(c, n, n, rp, ci, v)
            X = m.Vector(c, n=n, dtype=v.dtype)
            r = m.cg(A, X, m.Vector(c, b, dtype=v.dtype), tol=tol, maxiter=maxiter, precond=2,
                     mode=m.MODE_FUSED, check_every=every)
            return r.status, r.final_iter, bits(X.download()).copy()

        rp1, ci1, v1 = m.laplacian2d(128, 128)
        rp2, ci2, v2 = m.stencil3d(96, 96, 64, 7, 0.0)
        singles = []
        for k in range(2):
            b = oracle.random_vector(128 * 128, -1.0, 1.0, 1001 + k)
            singles.append((rp1, ci1, v1, b, 8 + 9 * k, 1e-10 * oracle.nrm2(b), 0))
        for k in range(2):  # 40 iterations of the split form, replayed 16 at a time
            b = oracle.random_vector(len(rp2) - 1, -1.0, 1.0, 9 + k, dtype=np.float32)
            singles.append((rp2, ci2, v2.astype(np.float32), b, 16, 1e-30, 40))
        out = {}
        for g in (0, 1):
            m.set_graphs(c, bool(g))
            out[g] = ([run_single(*s) for s in singles], run_batch("cg"), run_batch("bicgstab"))
        m.set_graphs(c, False)
        for a, b_ in zip(out[0], out[1]):
            for (s0, i0, x0), (s1, i1, x1) in zip(a, b_):
                assert s0 == s1 and i0 == i1 and np.array_equal(x0, x1)
        for (rp, ci, v, b, _, tol, maxiter), (st, it, x) in zip(singles, out[1][0]):
            kw = {"maxiter": maxiter} if maxiter else {}
            ref = oracle.solve("cg", rp, ci, v, np.zeros(len(b), v.dtype), b, precond=2, tol=tol,
                               **kw)
            assert st == ref[0] and it == ref[2] and np.array_equal(x, bits(ref[1]))
        for method, got in (("cg", out[1][1]), ("bicgstab", out[1][2])):
            for (meth, rp, ci, v, b), (st, it, x) in zip([s for s in syss if s[0] == method], got):
                ref = oracle_solve(oracle, meth, rp, ci, v, b)
                assert st == ref[0] and it == ref[2] and np.array_equal(x, bits(ref[1]))
    finally:
        c.close()


def test_server_cancel_deadline_poll(m, oracle):
    """Per-request control: a long request (one slot) runs while others
    queue; poll reports ACTIVE / QUEUED; a queued request cancelled ends at
    once with ERR_CANCELLED; the long request ends at its deadline with
    ERR_TIMEOUT after some iterations; the remaining request completes,
    bit-identical to the oracle; statistics count each outcome."""
    say("server control")
    import time
    c = m.Context(0)
    try:
        rp_l, ci_l, v_l = m.laplacian2d(400, 400)
        b_l = oracle.random_vector(400 * 400, -1.0, 1.0, 9)
        rp, ci, v = m.laplacian2d(40, 40)
        b = oracle.random_vector(1600, -1.0, 1.0, 10)
        srv = m.Server(c, max_active=1, iters_per_poll=4)
        t_long = srv.submit(m.METHOD_CG, rp_l, ci_l, v_l, b_l, tol=1e-300, tol_is_relative=False,
                            maxiter=10 ** 8)
        t_cancel = srv.submit(m.METHOD_CG, rp, ci, v, b, tol=1e-10, tol_is_relative=True)
        t_keep = srv.submit(m.METHOD_CG, rp, ci, v, b, tol=1e-10, tol_is_relative=True)
        t0 = time.time()
        while srv.poll(t_long) != m.REQ_ACTIVE and time.time() - t0 < 30:
            time.sleep(0.01)
        assert srv.poll(t_long) == m.REQ_ACTIVE
        assert srv.poll(t_cancel) == m.REQ_QUEUED and srv.poll(t_keep) == m.REQ_QUEUED
        srv.cancel(t_cancel)
        r, _ = srv.wait(t_cancel, 1600, timeout=10.0)
        assert r.status == m.ERR_CANCELLED
        srv.set_deadline(t_long, 0.5)
        r, x_long = srv.wait(t_long, 400 * 400, timeout=60.0)
        assert r.status == m.ERR_TIMEOUT and r.final_iter > 0
        assert np.all(np.isfinite(x_long)) and np.any(x_long != 0)
        r, x = srv.wait(t_keep, 1600, timeout=60.0)
        ref = oracle_solve(oracle, "cg", rp, ci, v, b)
        assert r.status == 0 and r.final_iter == ref[2]
        assert np.array_equal(bits(x), bits(ref[1]))
        st = srv.stats()
        assert (st["completed"], st["cancelled"], st["expired"]) == (1, 1, 1)
        with pytest.raises(m.MgxError):
            srv.poll(t_keep)  # collected by wait: unknown now
        srv.close()
    finally:
        c.close()


def test_batch_large_system_shares_launches(m, ctx, oracle):
    """In a batch a CG system of n >= 2^19 rows joins the shared tile
    launches (p formed first, gathered once) instead of running its own
    SpMV per iteration; 40 iterations of it and a small system are
    bit-identical to the oracle."""
    say("batch large system")
    rp1, ci1, v1 = m.stencil3d(96, 96, 64, 7, 0.0)
    rp2, ci2, v2 = m.laplacian2d(50, 50)
    syss = [(rp1, ci1, v1), (rp2, ci2, v2)]
    bs = [oracle.random_vector(len(rp) - 1, -1.0, 1.0, 31 + k) for k, (rp, _, _) in
          enumerate(syss)]
    As = [m.CRS(ctx, len(b), len(b), rp, ci, v) for (rp, ci, v), b in zip(syss, bs)]
    Xs = [m.Vector(ctx, n=len(b)) for b in bs]
    Bs = [m.Vector(ctx, b) for b in bs]
    res = m.batch_solve(ctx, m.METHOD_CG, As, Xs, Bs, tol=1e-30, maxiter=40, precond=2)
    for (rp, ci, v), b, r, X in zip(syss, bs, res, Xs):
        ref = oracle.solve("cg", rp, ci, v, np.zeros(len(b)), b, precond=2, tol=1e-30, maxiter=40)
        assert r.status == ref[0] == m.SOLVER_MAXITER and r.final_iter == 40
        assert np.array_equal(bits(X.download()), bits(ref[1]))


def test_server_fp32_refined_requests(m, oracle):
    """FP32-refined served requests (FP64 residuals and updates, FP32
    corrections sharing the FP32 engine's launches) mixed with FP64 requests
    in one server: each refined result equals mgx_ir_cg_solve of that system
    alone bit for bit (same outer and inner iteration counts) and reaches the
    FP64 tolerance; each FP64 result equals the oracle's."""
    say("server fp32-refined")
    c = m.Context(0)
    try:
        mats = [m.mmatrix(n, 5, 15, 0.5, seed=90 + n) for n in (900, 2500, 6000)]
        mats.append(m.laplacian2d(70, 70))
        bs = [oracle.random_vector(len(rp) - 1, -1.0, 1.0, 91 + k)
              for k, (rp, _, _) in enumerate(mats)]
        srv = m.Server(c, max_active=5, iters_per_poll=4, ir_inner_tol=1e-5)
        tickets = []
        for rep in range(2):
            for k, ((rp, ci, v), b) in enumerate(zip(mats, bs)):
                mode = m.MODE_MIXED_IR if (k + rep) % 2 == 0 else m.MODE_FUSED
                tickets.append((srv.submit(m.METHOD_CG, rp, ci, v, b, tol=1e-10,
                                           tol_is_relative=True, mode=mode), k, mode))
        results = [(srv.wait(t, len(bs[k]), timeout=120.0), k, mode) for t, k, mode in tickets]
        srv.close()
        for (r, x), k, mode in results:
            rp, ci, v = mats[k]
            b = bs[k]
            assert r.status == 0
            if mode == m.MODE_MIXED_IR:
                n = len(b)
                A, X1 = m.CRS(c, n, n, rp, ci, v), m.Vector(c, n=n)
                r1 = m.ir_cg(A, X1, m.Vector(c, b), tol=1e-10, tol_is_relative=True,
                             inner_tol=1e-5)
                say(n, r.outer_iter, r.inner_iter, r1.outer_iter, r1.inner_iter)
                assert (r.outer_iter, r.inner_iter) == (r1.outer_iter, r1.inner_iter)
                assert np.array_equal(bits(x), bits(X1.download()))
                assert oracle.residual_l2(rp, ci, v, x, b) < 1e-10 * oracle.nrm2(b)
            else:
                ref = oracle_solve(oracle, "cg", rp, ci, v, b)
                assert r.final_iter == ref[2] and np.array_equal(bits(x), bits(ref[1]))
    finally:
        c.close()
