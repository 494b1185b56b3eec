"""GPU parity of the Krylov solvers against the oracle.  The GPU loops use the
reference's arithmetic and the oracle's tile-order dots, so FP64 and FP32 runs
must reproduce the oracle's iterates bit for bit: same x, same iteration count,
same residual history -- in the reference-shaped mode and in the fused mode."""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, ans_check

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


def bits_equal(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.dtype != b.dtype or a.shape != b.shape:
        return False
    u = np.uint64 if a.dtype == np.float64 else np.uint32
    return np.array_equal(a.view(u), b.view(u))


PREC = {"default": 0, "none": 1, "jacobi": 2, "ic": 4, "ilu": 5}


def gpu_solve(m, ctx, method, rp, ci, val, x0, b, precond, tol, maxiter=0, miniter=0,
              mode=None, history=0, omega=None, plan=None):
    n = len(rp) - 1
    A = m.CRS(ctx, n, n, rp, ci, val)
    if plan:
        A.set_plan(**plan)
    X = m.Vector(ctx, x0, dtype=val.dtype)
    B = m.Vector(ctx, b, dtype=val.dtype)
    kw = dict(tol=tol, maxiter=maxiter, miniter=miniter, precond=precond, history=history)
    if mode is not None:
        kw["mode"] = mode
    if omega is not None:
        kw["omega"] = omega
    res = getattr(m, method)(A, X, B, **kw)
    return res, X.download()


def _mtx(oracle, name, dtype):
    rows, cols, r, c, v = oracle.input_mm(os.path.join(GOLDEN, name), dtype == np.float32)
    return oracle.crs_from_coo(rows, cols, r, c, v, dtype=dtype)


@pytest.mark.parametrize("method,mtx", [("cg", "test.mtx"), ("bicgstab", "test.mtx"),
                                        ("bicgstab", "test2.mtx"), ("jacobi", "test.mtx")])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-8), (np.float32, 1e-4)])
@pytest.mark.parametrize("precond", ["none", "jacobi", "ic", "ilu"])
def test_reference_solver_harness(m, ctx, oracle, method, mtx, dtype, tol, precond):
    # test/equation/iterative/equation_kernel.hpp: b = A*1, x0 = vector(n,0,1,
    # mt19937(1234)()), maxiter 10000, answer within 10*tol of 1
    if mtx == "test2.mtx" and precond in ("ic", "ilu"):
        pytest.skip("the reference's harness runs IC / ILU on test.mtx; on test2.mtx "
                    "ILU(0) is exact and FP32 BiCGSTAB's s is exactly 0 (omega = 0/0)")
    say("harness", method, mtx, dtype.__name__, precond)
    rp, ci, val = _mtx(oracle, mtx, dtype)
    n = len(rp) - 1
    b = oracle.spmv(rp, ci, val, np.ones(n, dtype))
    x0 = oracle.random_vector(n, 0.0, 1.0, 822569775, dtype=dtype)
    modes = [0, 1] if method == "cg" else [0]
    ref = oracle.solve(method, rp, ci, val, x0, b, precond=PREC[precond], tol=tol,
                       maxiter=10000, history=True)
    for mode in modes:
        res, x = gpu_solve(m, ctx, method, rp, ci, val, x0, b, PREC[precond], tol, 10000,
                           mode=mode, history=10000)
        assert res.status == 0
        assert ans_check(x, np.ones(n), tol * 10)
        assert res.final_iter == ref[2]
        assert bits_equal(x, ref[1])
        assert res.final_resid == ref[3]
        assert bits_equal(res.rhistory, ref[4])


def _poisson(oracle, side, dtype=np.float64):
    rows, cols, r, c, v = oracle.laplacian_2d5p(side, side)
    return oracle.crs_from_coo(rows, cols, r, c, v, dtype=dtype)


@pytest.mark.parametrize("mode,sliced", [(0, False), (1, False), (1, True)])
def test_cg_poisson256_bit_exact(m, ctx, oracle, mode, sliced):
    """Jacobi-PCG on the 2D 5-point Poisson 256x256, b = U[-1,1) seed 1001,
    x0 = 0, stop at ||r|| < 1e-10 ||b||; the fused form also with the
    matrix in the sliced layout (its SpMV stage then walks the slices one
    lane per row)."""
    say("cg poisson256 mode", mode, "sliced", sliced)
    rp, ci, val = _poisson(oracle, 256)
    n = len(rp) - 1
    b = oracle.random_vector(n, -1.0, 1.0, 1001)
    tol = 1e-10 * oracle.nrm2(b)
    ref = oracle.solve("cg", rp, ci, val, np.zeros(n), b, precond=2, tol=tol, history=True)
    assert ref[0] == 0
    res, x = gpu_solve(m, ctx, "cg", rp, ci, val, np.zeros(n), b, 2, tol, mode=mode, history=5000,
                       plan=dict(row_walk=2) if sliced else None)
    say("iters", res.final_iter, "resid", res.final_resid, "sec", res.seconds)
    assert res.status == 0 and res.final_iter == ref[2]
    assert bits_equal(x, ref[1])
    assert bits_equal(res.rhistory, ref[4])
    assert oracle.residual_l2(rp, ci, val, x, b) < 2 * tol
    # util::get_residual_l2 of the converged iterate, on the device
    A, X, B = m.CRS(ctx, n, n, rp, ci, val), m.Vector(ctx, x), m.Vector(ctx, b)
    assert m.residual_l2(A, X, B) == oracle.residual_l2(rp, ci, val, x, b)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("case", ["poisson256", "stencil27_48", "rect_lognormal"])
def test_residual_l2_bit_exact(m, ctx, oracle, case, dtype):
    """util::get_residual_l2 (equation_utils.cpp:45-56): ||b - A x||_2 through
    mgx_residual_l2 equals the oracle's tmp = A x; tmp = b - tmp; nrm2(tmp)
    bit for bit: the C1 matrix at full size, the 27-point stencil, and a
    rectangular random matrix (x sized by the columns, b by the rows)."""
    say("residual_l2", case, dtype.__name__)
    if case == "poisson256":
        rp, ci, val = _poisson(oracle, 256, dtype)
        cols = len(rp) - 1
    elif case == "stencil27_48":
        rp, ci, val = m.stencil3d(48, 48, 48, 27, 1e-2)
        val = val.astype(dtype)
        cols = len(rp) - 1
    else:
        cols = 7001
        rp, ci, val = m.random_lognormal(5003, cols, 12.0, 1.0, 300, seed=77)
        val = val.astype(dtype)
    rows = len(rp) - 1
    x = oracle.random_vector(cols, -1.0, 1.0, 11, dtype=dtype)
    b = oracle.random_vector(rows, -1.0, 1.0, 12, dtype=dtype)
    A = m.CRS(ctx, rows, cols, rp, ci, val)
    X, B = m.Vector(ctx, x, dtype=dtype), m.Vector(ctx, b, dtype=dtype)
    got = m.residual_l2(A, X, B)
    want = float(oracle.residual_l2(rp, ci, val, x, b))
    assert got == want, (got, want)
    # b := A x from the device SpMV (bit-exact to the row loop) gives exactly 0
    Y = m.Vector(ctx, n=rows, dtype=dtype)
    m.matvec(A, X, Y)
    assert m.residual_l2(A, X, Y) == 0.0
    for h in (A, X, B, Y):
        h.free()


def _banded_irregular(side=64):
    """SPD: the 2D 5-point Laplacian plus symmetric couplings (i, i +- k),
    k = 2..6, from every 7th row, diagonal = row sum |offdiag| + 0.5; row
    lengths 5..15 within 64-row slices, so no sliced copy qualifies, and
    columns near the diagonal, so the fused form gathers z and p."""
    import scipy.sparse as sp
    n = side * side
    A = sp.lil_matrix((n, n))
    for i in range(n):
        for j in (i - side, i - 1, i + 1, i + side):
            if 0 <= j < n:
                A[i, j] = -1.0
    for i in range(0, n, 7):
        for k in range(2, 7):
            for j in (i - k, i + k):
                if 0 <= j < n:
                    A[i, j] = A[j, i] = -0.1
    A = A.tocsr()
    A.setdiag(np.asarray(abs(A).sum(axis=1)).ravel() + 0.5)
    A = A.tocsr()
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("precond", [0, 1, 2])
@pytest.mark.parametrize("matrix", ["stencil27", "banded_irregular"])
def test_cg_fused_equals_reference_mode(m, ctx, oracle, dtype, precond, matrix):
    """The fused CG (two kernels, reductions finished in the next kernel)
    against the reference call sequence and the oracle: on a 27-point
    stencil (the SpMV stage walks the engine's sliced copy) and on a banded
    matrix whose rows vary too much for one (products staged in LDS)."""
    say("cg fused==reference", dtype.__name__, precond, matrix)
    if matrix == "stencil27":
        rp, ci, val = m.stencil3d(19, 17, 23, 27, 1e-2)
    else:
        rp, ci, val = _banded_irregular()
    val = val.astype(dtype)
    n = len(rp) - 1
    b = oracle.random_vector(n, -1.0, 1.0, 3001, dtype=dtype)
    tol = (1e-10 if dtype == np.float64 else 1e-5) * float(oracle.nrm2(b))
    ref = oracle.solve("cg", rp, ci, val, np.zeros(n, dtype), b, precond=precond, tol=tol,
                       history=True)
    for mode in (0, 1):
        res, x = gpu_solve(m, ctx, "cg", rp, ci, val, np.zeros(n, dtype), b, precond, tol,
                           mode=mode, history=4000)
        assert res.status == ref[0] and res.final_iter == ref[2], mode
        assert bits_equal(x, ref[1]), mode
        assert bits_equal(res.rhistory, ref[4]), mode


@pytest.mark.parametrize("blocks", [1, 0, 7, 64])
def test_cg_persistent_kernel(m, ctx, oracle, blocks):
    """The fused CG of one small system in one cooperative launch (blocks=1:
    a grid of one block per 128-row quarter; 7, 64: capped grids, blocks own
    several quarters; 0: one launch per stage) is bit-identical to the
    oracle, including the maxiter and NaN exits."""
    say("cg persistent", blocks)
    rp, ci, val = _poisson(oracle, 60)  # n = 3600: 8 tiles, 32 quarters
    n = len(rp) - 1
    b = oracle.random_vector(n, -1.0, 1.0, 1001)
    tol = 1e-10 * float(oracle.nrm2(b))
    m.set_persistent(ctx, blocks != 0, max_blocks=blocks)
    try:
        ref = oracle.solve("cg", rp, ci, val, np.zeros(n), b, precond=2, tol=tol, history=True)
        res, x = gpu_solve(m, ctx, "cg", rp, ci, val, np.zeros(n), b, 2, tol, mode=1, history=4000)
        assert res.status == ref[0] == 0 and res.final_iter == ref[2]
        assert bits_equal(x, ref[1]) and bits_equal(res.rhistory, ref[4])
        ref = oracle.solve("cg", rp, ci, val, np.zeros(n), b, precond=1, tol=1e-30, maxiter=9)
        res, x = gpu_solve(m, ctx, "cg", rp, ci, val, np.zeros(n), b, 1, 1e-30, maxiter=9, mode=1)
        assert res.status == m.SOLVER_MAXITER and res.final_iter == 9
        assert bits_equal(x, ref[1]) and res.final_resid == ref[3]
        bn = b.copy()
        bn[77] = np.nan
        res, x = gpu_solve(m, ctx, "cg", rp, ci, val, np.zeros(n), bn, 2, 1e-8, mode=1)
        assert res.status == m.SOLVER_RESIDUAL_NAN and res.final_iter == 1
    finally:
        m.set_persistent(ctx, False)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_cg_fused_large_split_iteration(m, ctx, oracle, dtype):
    """n >= 2^19 rows: the fused solver runs the split iteration (p update,
    plan-based SpMV, <p, q>, update); 40 iterations stay bit-identical to the
    oracle and to the reference-shaped GPU loop."""
    say("cg fused split", dtype.__name__)
    rp, ci, val = m.stencil3d(96, 96, 64, 7, 0.0)
    val = val.astype(dtype)
    n = len(rp) - 1
    assert n >= 1 << 19
    b = oracle.random_vector(n, -1.0, 1.0, 3001, dtype=dtype)
    ref = oracle.solve("cg", rp, ci, val, np.zeros(n, dtype), b, precond=2, tol=1e-30,
                       maxiter=40, history=True)
    assert ref[0] == m.SOLVER_MAXITER
    for mode in (0, 1):
        res, x = gpu_solve(m, ctx, "cg", rp, ci, val, np.zeros(n, dtype), b, 2, 1e-30,
                           maxiter=40, mode=mode, history=40)
        assert res.status == ref[0] and res.final_iter == 40, mode
        assert bits_equal(x, ref[1]), mode
        assert bits_equal(res.rhistory, ref[4]), mode


def test_cg_error_paths(m, ctx, oracle):
    say("cg error paths")
    rp, ci, val = _poisson(oracle, 16)
    n = len(rp) - 1
    b = oracle.random_vector(n, -1.0, 1.0, 1)
    for mode in (0, 1):
        res, x = gpu_solve(m, ctx, "cg", rp, ci, val, np.zeros(n), b, 2, 1e-30, maxiter=7, mode=mode)
        ref = oracle.solve("cg", rp, ci, val, np.zeros(n), b, precond=2, tol=1e-30, maxiter=7)
        assert res.status == m.SOLVER_MAXITER and res.final_iter == 7
        assert res.final_resid == ref[3] and bits_equal(x, ref[1])
        bn = b.copy()
        bn[5] = np.nan
        res, x = gpu_solve(m, ctx, "cg", rp, ci, val, np.zeros(n), bn, 2, 1e-8, mode=mode)
        assert res.status == m.SOLVER_RESIDUAL_NAN and res.final_iter == 1
        xe = oracle.random_vector(n, -1.0, 1.0, 2)
        be = oracle.spmv(rp, ci, val, xe)
        res, x = gpu_solve(m, ctx, "cg", rp, ci, val, xe, be, 2, 1e-8, mode=mode)
        assert res.status == m.SOLVER_RESIDUAL_NAN and res.final_iter == 1 and np.all(np.isnan(x))
        # a tolerance met at once: miniter alone decides the iteration count
        res, x = gpu_solve(m, ctx, "cg", rp, ci, val, np.zeros(n), b, 2, 1e6, miniter=9, mode=mode)
        ref = oracle.solve("cg", rp, ci, val, np.zeros(n), b, precond=2, tol=1e6, miniter=9)
        assert res.status == 0 and res.final_iter == 9 == ref[2]
        assert bits_equal(x, ref[1])
    res, x = gpu_solve(m, ctx, "bicgstab", rp, ci, val, xe, be, 2, 1e-8)
    assert res.status == m.SOLVER_BREAKDOWN and res.final_iter == 1


def test_bicgstab_convection_diffusion_bit_exact(m, ctx, oracle):
    say("bicgstab convdiff")
    import scipy.sparse as sp
    side = 60
    n = side * side
    pe = 20.0
    h = 1.0 / (side + 1)
    c = pe * h
    T1 = sp.diags([-1.0 - c, 2.0 + c, -1.0], [-1, 0, 1], shape=(side, side))
    I = sp.identity(side)
    A = (sp.kron(I, T1) + sp.kron(T1, I)).tocsr()
    A.sort_indices()
    rp, ci, val = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()
    b = oracle.random_vector(n, -1.0, 1.0, 5001)
    tol = 1e-10 * oracle.nrm2(b)
    ref = oracle.solve("bicgstab", rp, ci, val, np.zeros(n), b, precond=2, tol=tol, history=True)
    assert ref[0] == 0
    res, x = gpu_solve(m, ctx, "bicgstab", rp, ci, val, np.zeros(n), b, 2, tol, history=5000)
    assert res.status == 0 and res.final_iter == ref[2]
    assert bits_equal(x, ref[1]) and bits_equal(res.rhistory, ref[4])


def test_jacobi_solver_bit_exact(m, ctx, oracle):
    say("jacobi solver")
    rp, ci, val = m.stencil3d(12, 13, 14, 7, 2.0)
    n = len(rp) - 1
    b = oracle.random_vector(n, -1.0, 1.0, 8)
    ref = oracle.solve("jacobi", rp, ci, val, np.zeros(n), b, precond=0, tol=1e-9, history=True)
    res, x = gpu_solve(m, ctx, "jacobi", rp, ci, val, np.zeros(n), b, 0, 1e-9, history=10000)
    assert res.status == ref[0] == 0 and res.final_iter == ref[2]
    assert bits_equal(x, ref[1]) and res.final_resid == ref[3]


def test_mixed_precision_refinement(m, ctx, oracle):
    """FP32 inner / FP64 outer PCG on a 27-point shifted stencil: the FP64
    true residual reaches 1e-10 ||b|| and the answer agrees with the FP64
    oracle solve to the conditioning of the problem."""
    say("ir-cg 27pt")
    rp, ci, val = m.stencil3d(32, 32, 32, 27, 1e-2)
    n = len(rp) - 1
    b = oracle.random_vector(n, -1.0, 1.0, 3001)
    A = m.CRS(ctx, n, n, rp, ci, val)
    X = m.Vector(ctx, np.zeros(n))
    B = m.Vector(ctx, b)
    res = m.ir_cg(A, X, B, tol=1e-10, tol_is_relative=True, inner_tol=1e-4)
    say("ir", res)
    x = X.download()
    bn = oracle.nrm2(b)
    assert res.status == 0 and res.outer_iter >= 2
    assert oracle.residual_l2(rp, ci, val, x, b) < 1e-10 * bn
    ref = oracle.solve("cg", rp, ci, val, np.zeros(n), b, precond=2, tol=1e-10 * bn)
    # both meet ||r|| < 1e-10 ||b||; kappa(27-pt, shift 1e-2) < 3e3
    assert np.linalg.norm(x - ref[1]) <= 3e3 * 2e-10 * np.linalg.norm(ref[1])


@pytest.mark.parametrize("grid,delta", [(32, 0.1), (32, 1e-3), (81, 0.05)])
def test_mixed_cg_residual_replacement(m, ctx, oracle, grid, delta):
    """FP32 PCG with FP64 residual replacement on the 27-point shifted
    stencil (32^3: the two-kernel iteration; 81^3 = 531,441 rows: the split
    iteration).  The FP64 true residual meets 1e-10 ||b||, and because the
    search direction survives every replacement the FP32 iteration count
    stays within 10 % of the FP64 CG's (a wrong beta after a replacement
    restarts or breaks the recurrence and shows up here first)."""
    say("mixed-cg 27pt", grid, delta)
    rp, ci, val = m.stencil3d(grid, grid, grid, 27, 1e-2)
    n = len(rp) - 1
    b = oracle.random_vector(n, -1.0, 1.0, 3001)
    bn = oracle.nrm2(b)
    A = m.CRS(ctx, n, n, rp, ci, val)
    B = m.Vector(ctx, b)
    X64 = m.Vector(ctx, np.zeros(n))
    ref = m.cg(A, X64, B, tol=1e-10 * bn, precond=m.PRECOND_JACOBI)
    assert ref.status == 0
    X = m.Vector(ctx, np.zeros(n))
    res = m.mixed_cg(A, X, B, tol=1e-10, tol_is_relative=True, delta=delta)
    say("mixed", res, "fp64 iterations", ref.final_iter)
    x = X.download()
    assert res.status == 0 and res.outer_iter >= 2
    assert oracle.residual_l2(rp, ci, val, x, b) < 1e-10 * bn
    assert res.inner_iter <= 1.10 * ref.final_iter + 5
    x64 = X64.download()
    # both meet ||r|| < 1e-10 ||b||; kappa(27-pt, shift 1e-2) < 3e3
    assert np.linalg.norm(x - x64) <= 3e3 * 2e-10 * np.linalg.norm(x64)
    for v in (A, B, X64, X):
        v.free()


def test_mixed_cg_split_form_keeps_fp64_iteration_count(m, ctx, oracle):
    """The split-form FP32 iteration of the mixed solver reduces in FP64 and
    corrects the FP32-rounded diagonal, so on the 27-point stencil (81^3,
    531,441 rows) it needs no more iterations than the FP64 CG: with FP32
    dot products or the rounded diagonal (every eigenvalue shifted by the
    same rounding) it needs several percent more.  After new values (another
    diagonal shift) the correction follows them."""
    say("mixed-cg split iteration count")
    grid = 81
    for shift in (1e-2, 3e-2):
        rp, ci, val = m.stencil3d(grid, grid, grid, 27, shift)
        n = len(rp) - 1
        if shift == 1e-2:
            A = m.CRS(ctx, n, n, rp, ci, val)
            b = oracle.random_vector(n, -1.0, 1.0, 3001)
            bn = oracle.nrm2(b)
            B = m.Vector(ctx, b)
        else:
            A.update_values(val)
        X64 = m.Vector(ctx, np.zeros(n))
        ref = m.cg(A, X64, B, tol=1e-10 * bn, precond=m.PRECOND_JACOBI)
        X = m.Vector(ctx, np.zeros(n))
        res = m.mixed_cg(A, X, B, tol=1e-10, tol_is_relative=True, delta=1e-3)
        say("shift", shift, "mixed", res, "fp64 iterations", ref.final_iter)
        assert ref.status == 0 and res.status == 0
        assert oracle.residual_l2(rp, ci, val, X.download(), b) < 1e-10 * bn
        assert res.inner_iter <= ref.final_iter + 2
        X64.free()
        X.free()
    A.free()
    B.free()


def test_mixed_cg_edge_cases(m, ctx, oracle):
    """x0 already a solution: no FP32 iteration; maxiter reached: status
    MAXITER with the FP64 residual reported; delta outside (0, 1) rejected."""
    say("mixed-cg edges")
    rp, ci, val = m.stencil3d(8, 8, 8, 7)
    n = len(rp) - 1
    b = oracle.random_vector(n, -1.0, 1.0, 5)
    A = m.CRS(ctx, n, n, rp, ci, val)
    B = m.Vector(ctx, np.zeros(n))
    X = m.Vector(ctx, np.zeros(n))
    res = m.mixed_cg(A, X, B, tol=1e-10, tol_is_relative=False)
    assert res.status == 0 and res.inner_iter == 0 and res.outer_iter == 0
    B = m.Vector(ctx, b)
    X = m.Vector(ctx, np.zeros(n))
    res = m.mixed_cg(A, X, B, tol=1e-12, tol_is_relative=True, maxiter=3)
    assert res.status == m.SOLVER_MAXITER and res.inner_iter == 3
    assert abs(res.final_resid - oracle.residual_l2(rp, ci, val, X.download(), b)) <= 1e-12 * oracle.nrm2(b)
    with pytest.raises(m.MgxError):
        m.mixed_cg(A, X, B, delta=1.0)


def _block_diag(nblk, bs, seed):
    """nblk independent symmetric, diagonally dominant bs x bs blocks: a few
    SOR levels of nblk rows each (the one-launch-per-level path)."""
    rng = np.random.default_rng(seed)
    n = nblk * bs
    rp = np.arange(0, n * bs + 1, bs, dtype=np.int32)
    ci = np.concatenate([np.tile(np.arange(k * bs, (k + 1) * bs), bs) for k in range(nblk)])
    blk = rng.uniform(-1, 1, (nblk, bs, bs))
    blk = blk + blk.transpose(0, 2, 1)
    blk[:, np.arange(bs), np.arange(bs)] = 2.0 * bs
    return rp, ci.astype(np.int32), blk.reshape(-1)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("omega", [1.9, 1.0])
def test_sor_preconditioned_solvers_bit_exact(m, ctx, oracle, dtype, omega):
    """CG and BiCGSTAB with the SOR preconditioner (forward and backward
    sweeps, level-scheduled on the device) reproduce the oracle's serial
    sweeps bit for bit: x, iteration count and residual history; the fused
    mode takes the reference call sequence for it and gives the same bits."""
    say("sor precond", dtype.__name__, omega)
    cases = [("cg", m.laplacian2d(48, 40)), ("cg", m.stencil3d(12, 10, 9, 7)),
             ("bicgstab", m.convdiff2d(40, 20.0)), ("cg", _block_diag(6000, 3, 1)),
             ("cg", _block_diag(2000, 3, 8)), ("cg", _block_diag(600, 3, 10)),
             ("bicgstab", _arrow(10000, sym=False)),
             ("jacobi", m.laplacian2d(16, 16))]
    oracle.set_sor_omega(omega)
    for method, (rp, ci, val) in cases:
        val = val.astype(dtype)
        n = len(rp) - 1
        b = oracle.random_vector(n, -1.0, 1.0, 77, dtype=dtype)
        tol = float((1e-10 if dtype == np.float64 else 1e-5) * oracle.nrm2(b))
        ref = oracle.solve(method, rp, ci, val, np.zeros(n, dtype), b, precond=3, tol=tol,
                           maxiter=5000, history=True)
        for mode in (m.MODE_REFERENCE, m.MODE_FUSED):
            res, x = gpu_solve(m, ctx, method, rp, ci, val, np.zeros(n, dtype), b,
                               m.PRECOND_SOR, tol, maxiter=5000, mode=mode, history=5000,
                               omega=omega)
            say(method, n, mode, res.final_iter, ref[2])
            assert res.status == ref[0] == 0
            assert res.final_iter == ref[2]
            assert bits_equal(x, ref[1])
            assert bits_equal(res.rhistory, ref[4])
    oracle.set_sor_omega(1.9)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sor_solver_bit_exact(m, ctx, oracle, dtype):
    """equation::SOR on the device (lower sweep level by level) against the
    oracle's serial loop: Poisson, a tridiagonal chain (one row per level),
    independent blocks (levels wider than one workgroup's sweep), an unsorted row
    whose sweep stops at its first upper entry, and the reference harness
    (test.mtx, b = A 1, x0 random, x -> 1 within 10 tol)."""
    say("sor solver", dtype.__name__)
    n3 = 3000
    rows, cols, vals = [], [], []
    for i in range(n3):
        for j in (i - 1, i, i + 1):
            if 0 <= j < n3:
                rows.append(i)
                cols.append(j)
                vals.append(4.0 if i == j else -1.0)
    rp = np.zeros(n3 + 1, np.int32)
    np.add.at(rp, np.array(rows) + 1, 1)
    trid = (np.cumsum(rp).astype(np.int32), np.array(cols, np.int32), np.array(vals))
    unsorted = (np.array([0, 2, 4, 7], np.int32), np.array([0, 1, 1, 0, 2, 0, 1], np.int32),
                np.array([3.0, -1.0, 3.0, -1.0, 3.0, -1.0, -1.0]))
    cases = [m.laplacian2d(30, 30), trid, _block_diag(5000, 3, 2), _block_diag(1500, 3, 5),
             _arrow(12000, sym=False), unsorted]
    for omega in (1.9, 1.2):
        oracle.set_sor_omega(omega)
        # the unsorted system diverges at 1.9 (the lower sweep skips its rows'
        # off-diagonal entries): run it where the iteration converges
        for rp, ci, val in cases if omega < 1.5 else cases[:-1]:
            val = val.astype(dtype)
            n = len(rp) - 1
            b = oracle.random_vector(n, -1.0, 1.0, 5, dtype=dtype)
            x0 = oracle.random_vector(n, 0.0, 1.0, 6, dtype=dtype)
            tol = 1e-9 if dtype == np.float64 else 1e-4
            ref = oracle.solve("sor", rp, ci, val, x0, b, precond=1, tol=tol, maxiter=3000,
                               history=True)
            res, x = gpu_solve(m, ctx, "sor", rp, ci, val, x0, b, m.PRECOND_NONE, tol,
                               maxiter=3000, history=3000, omega=omega)
            say("sor", n, omega, res.final_iter, ref[2], res.status)
            assert res.status == ref[0] and res.final_iter == ref[2]
            assert bits_equal(x, ref[1])
            assert bits_equal(res.rhistory, ref[4])
    oracle.set_sor_omega(1.9)
    rp, ci, val = _mtx(oracle, "test.mtx", dtype)
    n = len(rp) - 1
    ans = np.ones(n, dtype)
    b = oracle.spmv(rp, ci, val, ans)
    x0 = oracle.random_vector(n, 0.0, 1.0, 822569775, dtype=dtype)
    tol = 1e-8 if dtype == np.float64 else 1e-4
    res, x = gpu_solve(m, ctx, "sor", rp, ci, val, x0, b, m.PRECOND_NONE, tol, maxiter=10000)
    assert res.status == 0 and ans_check(x, ans, tol * 10)


def test_sor_rejected_where_unsupported(m, ctx, oracle):
    say("sor rejects")
    rp, ci, val = m.laplacian2d(8, 8)
    n = len(rp) - 1
    A = m.CRS(ctx, n, n, rp, ci, val)
    X, B = m.Vector(ctx, n=n), m.Vector(ctx, np.ones(n))
    with pytest.raises(m.MgxError):
        m.mixed_cg(A, X, B, precond=m.PRECOND_SOR)
    with pytest.raises(m.MgxError):
        m.batch_solve(ctx, m.METHOD_CG, [A], [X], [B], precond=m.PRECOND_SOR)
    with pytest.raises(m.MgxError):
        m.cg(A, X, B, precond=m.PRECOND_SOR, omega=2.0)
    A.free()
    X.free()
    B.free()


def _random_sorted(n, density, seed, diag=6.0, sym=False):
    import scipy.sparse as sp
    M = sp.random(n, n, density=density, random_state=seed, format="csr")
    if sym:
        M = M + M.T
    M = sp.csr_matrix(M + sp.eye(n) * diag)
    M.sort_indices()
    return M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.copy()


def _arrow(n, eps=1e-4, sym=True):
    """tridiag(-1, 4, -1) with a coupling eps to row / column 0 (sym) or, with
    rows of at most four entries, to column 0 below the diagonal and column
    n-1 above it: rows far down the level order read a row more than 8192
    positions back (the sweeps' reads outside their LDS copy), in the
    forward and in the backward sweeps; n levels of one row each."""
    import scipy.sparse as sp
    M = sp.diags([-np.ones(n - 1), 4.0 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).tolil()
    M[1:, 0] = eps
    if sym:
        M[0, 1:] = eps
    else:
        M[:-1, n - 1] = eps
    M = sp.csr_matrix(M)
    M.sort_indices()
    return M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.copy()


def _factor_cases(m, kind):
    cases = [("poisson", m.laplacian2d(48, 40)), ("stencil27", m.stencil3d(14, 12, 10, 27)),
             ("blocks", _block_diag(6000, 3, 3)), ("blocks2k", _block_diag(2000, 3, 7)),
             ("blocks600", _block_diag(600, 3, 9)),
             ("arrow", _arrow(20000)), ("arrow_short_rows", _arrow(20000, 1e-3, sym=False)),
             ("random_spd", _random_sorted(3000, 0.002, 5, diag=12.0, sym=True))]
    if kind == 5:
        cases += [("convdiff", m.convdiff2d(40, 20.0)),
                  ("random", _random_sorted(4000, 0.002, 6))]
    return cases


@pytest.mark.parametrize("kind", [4, 5])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_ic_ilu_factor_and_apply_bit_exact(m, ctx, oracle, kind, dtype):
    """IC::solve / ILU::solve on the device (level-scheduled factorisation
    and triangular solves in one workgroup, or across the chip for the
    independent-blocks matrix, whose levels are 6000 rows wide) reproduce the
    oracle's serial factor-and-apply bit for bit."""
    for name, (rp, ci, val) in _factor_cases(m, kind):
        say("factor", kind, dtype.__name__, name)
        val = val.astype(dtype)
        n = len(rp) - 1
        b = oracle.random_vector(n, -1.0, 1.0, 31, dtype=dtype)
        rc, ref = oracle.ilu_precond(rp, ci, val, kind, b)
        assert rc == 0
        A = m.CRS(ctx, n, n, rp, ci, val)
        X = m.Vector(ctx, n=n, dtype=dtype)
        B = m.Vector(ctx, b, dtype=dtype)
        res = (m.ic_solve if kind == 4 else m.ilu_solve)(A, X, B)
        assert res.status == 0
        assert bits_equal(X.download(), ref)
        for v in (A, X, B):
            v.free()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_ic_ilu_preconditioned_solvers_bit_exact(m, ctx, oracle, dtype):
    """CG with IC(0) and ILU(0), BiCGSTAB with ILU(0): x, iteration count and
    residual history bit-identical to the oracle, in both modes (the fused
    mode takes the reference call sequence for these preconditioners)."""
    cases = [("cg", 4, m.laplacian2d(48, 40)), ("cg", 5, m.stencil3d(12, 10, 9, 27)),
             ("bicgstab", 5, m.convdiff2d(40, 20.0)), ("cg", 4, _block_diag(6000, 3, 4))]
    for method, kind, (rp, ci, val) in cases:
        val = val.astype(dtype)
        n = len(rp) - 1
        b = oracle.random_vector(n, -1.0, 1.0, 41, dtype=dtype)
        tol = float((1e-10 if dtype == np.float64 else 1e-5) * oracle.nrm2(b))
        ref = oracle.solve(method, rp, ci, val, np.zeros(n, dtype), b, precond=kind, tol=tol,
                           maxiter=5000, history=True)
        assert ref[0] == 0
        for mode in (m.MODE_REFERENCE, m.MODE_FUSED):
            res, x = gpu_solve(m, ctx, method, rp, ci, val, np.zeros(n, dtype), b, kind, tol,
                               maxiter=5000, mode=mode, history=5000)
            say(method, kind, n, mode, res.final_iter, ref[2])
            assert res.status == 0 and res.final_iter == ref[2]
            assert bits_equal(x, ref[1])
            assert bits_equal(res.rhistory, ref[4])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_large_wide_level_sweeps_bit_exact(m, ctx, oracle, dtype):
    """A 27-point 64x64x48 system (196608 rows, 378 levels of about 520
    rows) takes the dependency-driven sweep across the whole chip: IC(0) and
    ILU(0) factor-and-apply, and 12 iterations of SOR-, IC- and
    ILU-preconditioned CG, bit-identical to the oracle's serial sweeps."""
    rp, ci, val = m.stencil3d(64, 64, 48, 27, 1e-2)
    val = val.astype(dtype)
    n = len(rp) - 1
    b = oracle.random_vector(n, -1.0, 1.0, 57, dtype=dtype)
    for kind in (4, 5):
        say("wide factor", kind, dtype.__name__)
        rc, ref = oracle.ilu_precond(rp, ci, val, kind, b)
        assert rc == 0
        A = m.CRS(ctx, n, n, rp, ci, val)
        X = m.Vector(ctx, n=n, dtype=dtype)
        B = m.Vector(ctx, b, dtype=dtype)
        res = (m.ic_solve if kind == 4 else m.ilu_solve)(A, X, B)
        assert res.status == 0
        assert bits_equal(X.download(), ref)
        for v in (A, X, B):
            v.free()
    tol = float(1e-30 * oracle.nrm2(b))
    for kind in (m.PRECOND_SOR, 4, 5):
        ref = oracle.solve("cg", rp, ci, val, np.zeros(n, dtype), b, precond=kind, tol=tol,
                           maxiter=12, history=True)
        res, x = gpu_solve(m, ctx, "cg", rp, ci, val, np.zeros(n, dtype), b, kind, tol,
                           maxiter=12, mode=m.MODE_REFERENCE, history=12)
        say("wide cg", kind, dtype.__name__, res.status, res.final_iter, ref[0], ref[2])
        assert res.status == ref[0] and res.final_iter == ref[2]
        assert bits_equal(x, ref[1])
        assert bits_equal(res.rhistory, ref[4])


def _wide_then_chain(n_free, n_chain):
    """n_free independent rows (diagonal 4), then a tridiagonal chain
    (4 on the diagonal, -1 beside it): level 0 is wider than a workgroup's
    block and the chain adds n_chain - 1 levels of one row."""
    n = n_free + n_chain
    rp = [0]
    ci, val = [], []
    for i in range(n):
        if i >= n_free + 1:
            ci.append(i - 1)
            val.append(-1.0)
        ci.append(i)
        val.append(4.0)
        if n_free <= i < n - 1:
            ci.append(i + 1)
            val.append(-1.0)
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(val)


def test_deep_plan_with_a_wide_level_bit_exact(m, ctx, oracle):
    """5000 independent rows plus a 100000-row chain: one level wider than a
    workgroup's block, then 99999 levels of one row.  Such deep plans sweep
    in one workgroup (every wait there is a barrier); IC(0) / ILU(0) apply and
    SOR-, IC-preconditioned CG stay bit-identical to the oracle, with no
    give-up of the dependency-driven form."""
    rp, ci, val = _wide_then_chain(5000, 100000)
    n = len(rp) - 1
    b = oracle.random_vector(n, -1.0, 1.0, 58)
    for kind in (4, 5):
        say("deep plan factor", kind)
        rc, ref = oracle.ilu_precond(rp, ci, val, kind, b)
        assert rc == 0
        A = m.CRS(ctx, n, n, rp, ci, val)
        X = m.Vector(ctx, n=n)
        res = (m.ic_solve if kind == 4 else m.ilu_solve)(A, X, m.Vector(ctx, b))
        assert res.status == 0
        assert bits_equal(X.download(), ref)
    tol = float(1e-30 * oracle.nrm2(b))
    for kind in (m.PRECOND_SOR, 4):
        ref = oracle.solve("cg", rp, ci, val, np.zeros(n), b, precond=kind, tol=tol, maxiter=3,
                           history=True)
        res, x = gpu_solve(m, ctx, "cg", rp, ci, val, np.zeros(n), b, kind, tol, maxiter=3,
                           mode=m.MODE_REFERENCE, history=3)
        say("deep plan cg", kind, res.status, res.final_iter)
        # SOR runs out its 3 iterations; IC(0) of this tridiagonal chain is
        # exact, so its PCG converges first
        assert res.status == ref[0] and res.final_iter == ref[2]
        assert ref[0] == (m.SOLVER_MAXITER if kind == m.PRECOND_SOR else 0)
        assert bits_equal(x, ref[1]) and bits_equal(res.rhistory[:ref[2]], ref[4][:ref[2]])


def test_ic_ilu_errors(m, ctx, oracle):
    """A missing diagonal entry, unsorted columns, an indefinite matrix under
    IC and a zero ILU pivot are refused with an error naming the row; the
    fused-engine forms refuse IC / ILU."""
    say("ic/ilu errors")

    def crs(rows):
        rp = np.cumsum([0] + [len(r) for r in rows]).astype(np.int32)
        ci = np.array([c for r in rows for c, _ in r], np.int32)
        val = np.array([v for r in rows for _, v in r])
        return rp, ci, val
    bad = [(crs([[(0, 2.0), (1, 1.0)], [(0, 1.0)]]), (4, 5), "missing"),
           (crs([[(0, 2.0)], [(1, 2.0), (0, 1.0)]]), (4, 5), "increasing"),
           (crs([[(0, 1.0), (1, 2.0)], [(0, 2.0), (1, 1.0)]]), (4,), "pivot at row 1"),
           (crs([[(0, 1.0), (1, 1.0)], [(0, 1.0), (1, 1.0)]]), (5,), "pivot at row 1")]
    for (rp, ci, val), kinds, msg in bad:
        n = len(rp) - 1
        A = m.CRS(ctx, n, n, rp, ci, val)
        X, B = m.Vector(ctx, n=n), m.Vector(ctx, np.ones(n))
        for kind in kinds:
            with pytest.raises(m.MgxError, match=msg):
                m.cg(A, X, B, precond=kind)
            with pytest.raises(m.MgxError, match=msg):
                (m.ic_solve if kind == 4 else m.ilu_solve)(A, X, B)
        for v in (A, X, B):
            v.free()
    rp, ci, val = m.laplacian2d(8, 8)
    n = len(rp) - 1
    A = m.CRS(ctx, n, n, rp, ci, val)
    X, B = m.Vector(ctx, n=n), m.Vector(ctx, np.ones(n))
    for kind in (4, 5):
        with pytest.raises(m.MgxError):
            m.mixed_cg(A, X, B, precond=kind)
        with pytest.raises(m.MgxError):
            m.batch_solve(ctx, m.METHOD_CG, [A], [X], [B], precond=kind)
    for v in (A, X, B):
        v.free()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_triangular_preconditioners_tiny_systems(m, ctx, oracle, dtype):
    """1x1, 2x2 and 3x3 systems (one level, one block, no dependency
    entries in some sweeps) through SOR, IC and ILU: bit-identical to the
    oracle, and IC/ILU solves equal the direct solution to rounding."""
    mats = [(np.array([0, 1], np.int32), np.array([0], np.int32), np.array([4.0])),
            (np.array([0, 2, 4], np.int32), np.array([0, 1, 0, 1], np.int32),
             np.array([4.0, -1.0, -1.0, 3.0])),
            (np.array([0, 2, 5, 7], np.int32), np.array([0, 1, 0, 1, 2, 1, 2], np.int32),
             np.array([2.0, -1.0, -1.0, 2.0, -1.0, -1.0, 2.0]))]
    for rp, ci, val in mats:
        val = val.astype(dtype)
        n = len(rp) - 1
        say("tiny", n, dtype.__name__)
        b = oracle.random_vector(n, -1.0, 1.0, 3, dtype=dtype)
        for kind in (4, 5):
            rc, ref = oracle.ilu_precond(rp, ci, val, kind, b)
            A = m.CRS(ctx, n, n, rp, ci, val)
            X, B = m.Vector(ctx, n=n, dtype=dtype), m.Vector(ctx, b, dtype=dtype)
            res = (m.ic_solve if kind == 4 else m.ilu_solve)(A, X, B)
            assert res.status == 0 and bits_equal(X.download(), ref)
            for v in (A, X, B):
                v.free()
        for prec in (3, 4, 5):
            ref = oracle.solve("cg", rp, ci, val, np.zeros(n, dtype), b, precond=prec, tol=1e-30,
                               maxiter=3, history=True)
            res, x = gpu_solve(m, ctx, "cg", rp, ci, val, np.zeros(n, dtype), b, prec, 1e-30,
                               maxiter=3, history=3)
            assert res.status == ref[0] and res.final_iter == ref[2]
            assert bits_equal(x, ref[1])


def test_triangular_plans_follow_updated_values(m, ctx, oracle):
    """The sweep plans and IC/ILU structure are built once per matrix
    structure; the coefficients are gathered per solve.  After
    CRS::operator= (new values, same pattern) SOR-, IC- and ILU-PCG give
    the oracle's results for the new values, bit for bit."""
    rp, ci, val = m.laplacian2d(40, 30)
    n = len(rp) - 1
    A = m.CRS(ctx, n, n, rp, ci, val)
    b = oracle.random_vector(n, -1.0, 1.0, 12)
    B = m.Vector(ctx, b)
    X = m.Vector(ctx, n=n)
    val2 = val * oracle.random_vector(len(val), 0.5, 1.5, 13)
    val2 = 0.5 * (val2 + _transpose_values(rp, ci, val2))  # keep it symmetric
    val2[_diag_pos(rp, ci)] = val[_diag_pos(rp, ci)] * 1.3
    for vals in (val, val2, val):
        A.update_values(vals)
        for prec in (3, 4, 5):
            say("updated values", prec)
            ref = oracle.solve("cg", rp, ci, vals, np.zeros(n), b, precond=prec, tol=1e-10,
                               maxiter=2000)
            X.fill(0.0)
            res = m.cg(A, X, B, tol=1e-10, maxiter=2000, precond=prec)
            assert res.status == ref[0] == 0 and res.final_iter == ref[2]
            assert bits_equal(X.download(), ref[1])
    for v in (A, B, X):
        v.free()


def _diag_pos(rp, ci):
    return np.array([rp[i] + int(np.nonzero(ci[rp[i]:rp[i + 1]] == i)[0][0])
                     for i in range(len(rp) - 1)])


def _transpose_values(rp, ci, val):
    """values of A^T at A's positions (A's pattern is symmetric)"""
    n = len(rp) - 1
    pos = {}
    for i in range(n):
        for e in range(rp[i], rp[i + 1]):
            pos[(i, int(ci[e]))] = e
    out = np.empty_like(val)
    for i in range(n):
        for e in range(rp[i], rp[i + 1]):
            out[e] = val[pos[(int(ci[e]), i)]]
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_engine_sliced_copy_follows_values(m, ctx, oracle, dtype):
    """The fused solver's own sliced copy (a matrix whose SpMV plan is a
    block plan): built on the first fused solve, rewritten by update_values,
    so the next fused solve walks the new values; both solves reproduce the
    oracle bit for bit, and so does a mixed-precision solve (its FP32 copy
    gets a sliced copy of its own) for FP64."""
    say("engine sliced copy", dtype.__name__)
    rp, ci, val = m.stencil3d(30, 27, 21, 7, 0.5)
    n = len(rp) - 1
    b = oracle.random_vector(n, -1.0, 1.0, 4242, dtype=dtype)
    A = m.CRS(ctx, n, n, rp, ci, val.astype(dtype))
    A.set_plan(rows_per_block=256, nnz_per_block=2048)
    X = m.Vector(ctx, n=n, dtype=dtype)
    B = m.Vector(ctx, b, dtype=dtype)
    for scale in (1.0, 3.0):
        v = (val * scale).astype(dtype)
        if scale != 1.0:
            A.update_values(v)
        ref = oracle.solve("cg", rp, ci, v, np.zeros(n, dtype), b, precond=2, tol=1e-30,
                           maxiter=40)
        X.fill(0.0)
        res = m.cg(A, X, B, tol=1e-30, maxiter=40, precond=2, mode=m.MODE_FUSED)
        assert res.final_iter == ref[2] == 40, scale
        assert bits_equal(X.download(), ref[1]), scale
    if dtype == np.float64:
        bn = float(np.linalg.norm(b))
        X.fill(0.0)
        r = m.mixed_cg(A, X, B, tol=1e-9 * bn, tol_is_relative=False, delta=1e-3)
        x1 = X.download()
        A.set_plan(row_walk=2)  # the plan's sliced copy from now on
        X.fill(0.0)
        r2 = m.mixed_cg(A, X, B, tol=1e-9 * bn, tol_is_relative=False, delta=1e-3)
        assert r.status == r2.status == 0 and r.final_iter == r2.final_iter
        assert bits_equal(X.download(), x1)
    for h in (A, X, B):
        h.free()


def test_c3_full_size_bit_exact(m, ctx, oracle):
    """The 27-point 128^3 system at full size (diagonal shift 1e-2, b =
    U[-1,1) seed 3001, Jacobi-PCG to 1e-10 ||b||): the fused FP64 solve
    (split form: per-system streaming kernels, sliced SpMV forming the <p,q>
    partials) and the plain one reproduce the oracle's 273 iterations bit
    for bit, residual history included; the mixed-precision solve (FP32
    iterations, FP64 residual replacement) ends below 1e-10 ||b|| in true
    residual, within 1e-7 of that solution."""
    say("c3 full size")
    rp, ci, val = m.stencil3d(128, 128, 128, 27, 1e-2)
    n = len(rp) - 1
    b = oracle.random_vector(n, -1.0, 1.0, 3001)
    bn = oracle.nrm2(b)
    tol = 1e-10 * bn
    ref = oracle.solve("cg", rp, ci, val, np.zeros(n), b, precond=2, tol=tol, history=True)
    assert ref[0] == 0 and ref[2] == 273
    A = m.CRS(ctx, n, n, rp, ci, val)
    A.autotune(reps=3)
    X = m.Vector(ctx, n=n)
    B = m.Vector(ctx, b)
    for mode in (m.MODE_FUSED, m.MODE_REFERENCE):
        X.fill(0.0)
        res = m.cg(A, X, B, tol=tol, precond=2, mode=mode, history=400)
        say("mode", mode, res.final_iter)
        assert res.status == 0 and res.final_iter == ref[2], mode
        assert bits_equal(X.download(), ref[1]), mode
        assert bits_equal(res.rhistory, ref[4]), mode
    X.fill(0.0)
    res = m.mixed_cg(A, X, B, tol=1e-10, tol_is_relative=True, delta=1e-3)
    x = X.download()
    say("mixed", res.status, res.final_iter, res.outer_iter)
    assert res.status == 0
    assert oracle.residual_l2(rp, ci, val, x, b) < 1e-10 * bn
    assert np.linalg.norm(x - ref[1]) < 1e-7 * np.linalg.norm(ref[1])
    for h in (A, X, B):
        h.free()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_split_iteration_with_sliced_spmv(m, ctx, oracle, dtype):
    """The split CG iteration with the sliced-layout SpMV, which also forms
    the <p, q> tile partials (the stage-3 kernel then skips the system):
    bit-identical to the oracle, as with the row-block plan, in every block
    order (launch order, a range per XCD, runs of 16 and of 3 blocks per
    XCD); the SpMV alone gives the same bits in each order; and the
    mixed-precision CG gives the same bits with either plan."""
    say("split sliced", dtype.__name__)
    rp, ci, val = m.stencil3d(96, 96, 64, 27, 0.01)
    val = val.astype(dtype)
    n = len(rp) - 1
    b = oracle.random_vector(n, -1.0, 1.0, 3003, dtype=dtype)
    ref = oracle.solve("cg", rp, ci, val, np.zeros(n, dtype), b, precond=2, tol=1e-30,
                       maxiter=30, history=True)
    A = m.CRS(ctx, n, n, rp, ci, val)
    X = m.Vector(ctx, n=n, dtype=dtype)
    B = m.Vector(ctx, b, dtype=dtype)
    Y = m.Vector(ctx, n=n, dtype=dtype)
    ys = []
    for remap in (1, 0, 16, 3):
        A.set_plan(row_walk=2, xcd_remap=remap)
        X.fill(0.0)
        res = m.cg(A, X, B, tol=1e-30, maxiter=30, precond=2, mode=m.MODE_FUSED, history=30)
        assert res.final_iter == 30 and bits_equal(X.download(), ref[1]), remap
        assert bits_equal(res.rhistory, ref[4]), remap
        m.matvec(A, B, Y)
        ys.append(Y.download())
    assert all(bits_equal(y, ys[0]) for y in ys[1:])
    assert bits_equal(ys[0], oracle.spmv(rp, ci, val, b))
    Y.free()
    if dtype == np.float64:
        bn = float(np.linalg.norm(b))
        xs = []
        for walk in (2, 0):
            A.set_plan(row_walk=walk)
            X.fill(0.0)
            r = m.mixed_cg(A, X, B, tol=1e-9 * bn, tol_is_relative=False, delta=1e-3)
            assert r.status == 0
            xs.append(X.download())
        assert bits_equal(xs[0], xs[1])
    for v in (A, X, B):
        v.free()
