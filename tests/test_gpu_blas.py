"""GPU parity of the sparse and vector BLAS (HIP kernels through the C-ABI)
against the oracle on the same seeded inputs.  Integer/index work and every
CRS-order row sum must match bit for bit; long rows (tree order) are held to
the forward-error bound of the CPU loop."""
import math
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

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


def gpu_spmv(m, ctx, rp, ci, val, x, cols=None):
    rows = len(rp) - 1
    cols = cols if cols is not None else len(x)
    A = m.CRS(ctx, rows, cols, rp, ci, val)
    X = m.Vector(ctx, x, dtype=val.dtype)
    Y = m.Vector(ctx, n=rows, dtype=val.dtype)
    m.matvec(A, X, Y)
    return Y.download()


def test_device_is_gfx950(m, ctx):
    info = ctx.info()
    say("device", info)
    assert "gfx950" in info["name"]
    assert info["compute_units"] >= 1


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_matvec_reference_known_answers(m, ctx, oracle, dtype):
    say("matvec KAT", dtype)
    row = np.array([0, 0, 0, 1, 1, 2, 2, 2], np.int32)
    col = np.array([0, 1, 2, 0, 2, 0, 1, 2], np.int32)
    val = np.array([1, 2, 3, 4, 5, 6, 7, 8], np.float64)
    rp, ci, v = oracle.crs_from_coo(3, 3, row, col, val, dtype=dtype)
    y = gpu_spmv(m, ctx, rp, ci, v, np.full(3, 10, dtype))
    assert y.tolist() == [60, 90, 210]
    rows, cols, r, c, vv = oracle.input_mm(os.path.join(GOLDEN, "test.mtx"), dtype == np.float32)
    rp, ci, v = oracle.crs_from_coo(rows, cols, r, c, vv, dtype=dtype)
    y = gpu_spmv(m, ctx, rp, ci, v, np.full(3, 10, dtype))
    assert y.tolist() == [10, 0, 10]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_matvec_poisson2d_bit_exact(m, ctx, oracle, dtype):
    say("matvec poisson 256^2", dtype)
    rows, cols, r, c, v = oracle.laplacian_2d5p(256, 256)
    rp, ci, val = oracle.crs_from_coo(rows, cols, r, c, v, dtype=dtype)
    x = oracle.random_vector(rows, -1.0, 1.0, 1001, dtype=dtype)
    y = gpu_spmv(m, ctx, rp, ci, val, x)
    assert bits_equal(y, oracle.spmv(rp, ci, val, x))


@pytest.mark.parametrize("points,shift,side", [(7, 0.0, 37), (27, 1e-2, 23)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_matvec_stencils_bit_exact(m, ctx, oracle, points, shift, side, dtype):
    say("matvec stencil", points, side, dtype)
    rp, ci, val = m.stencil3d(side, side + 1, side - 1, points, shift)
    val = val.astype(dtype)
    x = oracle.random_vector(len(rp) - 1, -1.0, 1.0, 2001, dtype=dtype)
    y = gpu_spmv(m, ctx, rp, ci, val, x)
    assert bits_equal(y, oracle.spmv(rp, ci, val, x))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_matvec_lognormal_rows(m, ctx, oracle, dtype):
    say("matvec lognormal", dtype)
    n = 120000
    rp, ci, val = m.random_lognormal(n, n, 32.0, 1.0, 2048, seed=77)
    val = val.astype(dtype)
    x = oracle.random_vector(n, -1.0, 1.0, 2002, dtype=dtype)
    y = gpu_spmv(m, ctx, rp, ci, val, x)
    ref = oracle.spmv(rp, ci, val, x)
    lens = np.diff(rp)
    short = lens <= 256  # CRS-order rows: bit-exact
    assert bits_equal(y[short], ref[short])
    # long rows: forward error bound of any summation order
    u = 2.0 ** -53 if dtype == np.float64 else 2.0 ** -24
    absrow = oracle.spmv(rp, ci, np.abs(val), np.abs(x)).astype(np.float64)
    k = lens.astype(np.float64)
    bound = 2 * k * u / (1 - 2 * k * u) * absrow
    assert np.all(np.abs(y.astype(np.float64) - ref) <= bound + 1e-300)
    if dtype == np.float64:
        assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)


def test_matvec_tuning_invariance(m, oracle):
    say("matvec tuning invariance")
    n = 50000
    rp, ci, val = m.random_lognormal(n, n, 20.0, 1.2, 600, seed=3)
    x = oracle.random_vector(n, -1.0, 1.0, 9)
    outs = []
    for tun in ((256, 2048, 256, True), (64, 512, 64, False), (256, 8192, 8192, True), (1, 64, 64, True)):
        c = m.Context(0)
        c.set_spmv_tuning(*tun)
        outs.append(gpu_spmv(m, c, rp, ci, val, x))
        c.close()
    ref = oracle.spmv(rp, ci, val, x)
    short = np.diff(rp) <= 64
    for y in outs:
        assert bits_equal(y[short], ref[short])
        assert np.linalg.norm(y - ref) <= 1e-13 * np.linalg.norm(ref)
    assert bits_equal(outs[2], ref)  # every row sequential when long_row >= max length


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_matvec_autotune(m, ctx, oracle, dtype):
    """The autotuner picks one of its candidate plans; the product is the same
    bit for bit before and after, and equals the CPU row loop."""
    say("matvec autotune", dtype.__name__)
    rp, ci, val = m.stencil3d(40, 37, 29, 27, 1e-2)
    val = val.astype(dtype)
    n = len(rp) - 1
    x = oracle.random_vector(n, -1.0, 1.0, 2001, dtype=dtype)
    A = m.CRS(ctx, n, n, rp, ci, val)
    X = m.Vector(ctx, x, dtype=dtype)
    Y = m.Vector(ctx, n=n, dtype=dtype)
    m.matvec(A, X, Y)
    before = Y.download()
    t = A.autotune(reps=3)
    assert t["rows_per_block"] in (64, 128, 256) and t["nnz_per_block"] in (1024, 2048, 4096)
    assert t["candidates"] >= 27 and t["us"] > 0 and t["long_row"] == 256
    Y.fill(0.0)
    m.matvec(A, X, Y)
    after = Y.download()
    assert bits_equal(before, after)
    assert bits_equal(after, oracle.spmv(rp, ci, val, x))


def test_matvec_edge_shapes(m, ctx, oracle):
    say("matvec edge shapes")
    # empty rows, trailing empty rows left at 0 by CRS(COO&), one long row
    row = np.array([1, 1, 3] + [5] * 5000, np.int32)
    col = np.concatenate([[0, 2, 1], np.arange(5000) % 7]).astype(np.int32)
    val = np.linspace(-1, 1, len(row))
    # distinct columns per row are not required by CRS; keep COO order
    rp, ci, v = oracle.crs_from_coo(8, 7, row, col, val)
    assert rp[-1] == 0  # trailing rows keep 0
    x = oracle.random_vector(7, -1.0, 1.0, 4)
    A = m.CRS(ctx, 8, 7, rp, ci, v)
    X = m.Vector(ctx, x)
    Y = m.Vector(ctx, n=8)
    Y.fill(123.0)
    m.matvec(A, X, Y)
    y = Y.download()
    ref = oracle.spmv(rp, ci, v, x)
    assert bits_equal(y[[0, 1, 2, 3, 4, 6, 7]], ref[[0, 1, 2, 3, 4, 6, 7]])
    assert abs(y[5] - ref[5]) <= 1e-12 * np.sum(np.abs(v[3:]))
    rp2, ci2, v2 = A.download()  # downloads the arrays as given
    assert np.array_equal(rp2, rp) and np.array_equal(ci2, ci) and bits_equal(v2, v)
    # zero-row and zero-nnz matrices
    Z = m.CRS(ctx, 0, 5, np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0))
    m.matvec(Z, m.Vector(ctx, np.ones(5)), m.Vector(ctx, n=0))
    E = m.CRS(ctx, 4, 4, np.zeros(5, np.int32), np.zeros(0, np.int32), np.zeros(0))
    Ye = m.Vector(ctx, np.full(4, 7.0))
    m.matvec(E, m.Vector(ctx, np.ones(4)), Ye)
    assert Ye.download().tolist() == [0, 0, 0, 0]


def test_csr_create_rejects_bad_indices(m, ctx):
    say("csr validation")
    with pytest.raises(m.MgxError):
        m.CRS(ctx, 2, 2, np.array([0, 1, 2], np.int32), np.array([0, 2], np.int32), np.ones(2))
    with pytest.raises(m.MgxError):
        m.CRS(ctx, 2, 2, np.array([0, 2, 1], np.int32), np.array([0, 1], np.int32), np.ones(2))
    A = m.CRS(ctx, 2, 3, np.array([0, 1, 2], np.int32), np.array([0, 2], np.int32), np.ones(2))
    with pytest.raises(m.MgxError):  # x has the wrong size
        m.matvec(A, m.Vector(ctx, np.ones(2)), m.Vector(ctx, n=2))


def test_matvec_alpha_beta(m, ctx, oracle):
    say("matvec alpha beta")
    rows, cols, r, c, v = oracle.laplacian_2d5p(40, 30)
    rp, ci, val = oracle.crs_from_coo(rows, cols, r, c, v)
    x = oracle.random_vector(rows, -1.0, 1.0, 5)
    y0 = oracle.random_vector(rows, -1.0, 1.0, 6)
    A = m.CRS(ctx, rows, cols, rp, ci, val)
    X = m.Vector(ctx, x)
    Y = m.Vector(ctx, y0)
    m.matvec_ab(2.5, A, X, -0.75, Y)
    ax = oracle.spmv(rp, ci, val, x)
    ref = 2.5 * ax + (-0.75) * y0  # numpy: each product rounded, then the sum
    assert bits_equal(Y.download(), ref)
    Y.upload(np.full(rows, np.nan))
    m.matvec_ab(1.0, A, X, 0.0, Y)  # beta = 0 does not read y
    assert bits_equal(Y.download(), ax)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("layout", ["sorted", "unsorted_dup"])
def test_matvec_transpose(m, ctx, oracle, dtype, layout):
    """A^T is built on the device by a stable sort of the entries by column;
    each column's products are added in ascending row and, within a row, in
    storage order (unsorted and repeated columns included), as the CPU
    scatter loop does."""
    say("matvec_T", dtype, layout)
    rng = np.random.default_rng(1)
    rows, cols = 3000, 1700
    lens = rng.integers(0, 30, rows)
    rp = np.zeros(rows + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    if layout == "sorted":
        ci = np.concatenate([np.sort(rng.choice(cols, l, replace=False)) for l in lens])
    else:
        ci = np.concatenate([rng.integers(0, cols, l) for l in lens])
    ci = ci.astype(np.int32)
    val = rng.uniform(-1, 1, rp[-1]).astype(dtype)
    x = oracle.random_vector(rows, -1.0, 1.0, 12, dtype=dtype)
    A = m.CRS(ctx, rows, cols, rp, ci, val)
    X = m.Vector(ctx, x, dtype=dtype)
    Y = m.Vector(ctx, n=cols, dtype=dtype)
    m.matvec_t(A, X, Y)
    assert bits_equal(Y.download(), oracle.spmv_t(rp, ci, val, x, cols))


def test_matvec_c2a_full_size(m, ctx, oracle):
    """3D 7-point 192^3 (n = 7,077,888, nnz = 49,324,032) at full size."""
    say("matvec 192^3 full size")
    rp, ci, val = m.stencil3d(192, 192, 192, 7)
    assert len(val) == 49324032
    x = oracle.random_vector(len(rp) - 1, -1.0, 1.0, 2001)
    y = gpu_spmv(m, ctx, rp, ci, val, x)
    assert bits_equal(y, oracle.spmv(rp, ci, val, x))
    # the all-ones vector: row sums of the stencil (6 - #neighbours)
    ones = gpu_spmv(m, ctx, rp, ci, val, np.ones(len(x)))
    assert np.array_equal(ones, 7.0 - np.diff(rp))


def test_matvec_c2b_full_size(m, ctx, oracle):
    """The random CRS of 4M rows at full size (row lengths lognormal, mean
    32, sigma 1, clipped to [1, 2048]; uniform columns without repeats;
    seed 2002), FP64 and FP32, under the autotuned plan (column panels):
    every row of at most 256 entries carries the CPU row loop's bits, the
    longer ones (the block tree) are within 1e-12 (FP64) / 1e-5 (FP32) of
    it; and the bits do not depend on the plan (the block plan gives the
    same y)."""
    say("matvec c2b full size")
    rp, ci, val = m.random_lognormal(4_000_000, 4_000_000, 32.0, 1.0, 2048, seed=2002)
    short = np.diff(rp) <= 256
    assert short.sum() < len(short)
    for dtype, tol in ((np.float64, 1e-12), (np.float32, 1e-5)):
        v = val.astype(dtype)
        x = oracle.random_vector(len(rp) - 1, -1.0, 1.0, 2003, dtype=dtype)
        A = m.CRS(ctx, len(rp) - 1, len(x), rp, ci, v)
        X = m.Vector(ctx, x, dtype=dtype)
        Y = m.Vector(ctx, n=len(rp) - 1, dtype=dtype)
        t = A.autotune(reps=3)
        say(dtype.__name__, t)
        m.matvec(A, X, Y)
        y = Y.download()
        ref = oracle.spmv(rp, ci, v, x)
        assert bits_equal(y[short], ref[short])
        d = (y - ref)[~short].astype(np.float64)
        assert np.linalg.norm(d) <= tol * np.linalg.norm(ref[~short].astype(np.float64))
        A.set_plan(rows_per_block=256, nnz_per_block=2048)
        Y.fill(0.0)
        m.matvec(A, X, Y)
        assert bits_equal(Y.download(), y)
        for h in (A, X, Y):
            h.free()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [0, 1, 2, 511, 512, 513, 4099, 300001, 3000000])
def test_dot_nrm2_bit_exact(m, ctx, oracle, dtype, n):
    say("dot/nrm2", dtype, n)
    x = oracle.random_vector(n, -1.0, 1.0, 100 + n % 97, dtype=dtype)
    y = oracle.random_vector(n, -1.0, 1.0, 200 + n % 89, dtype=dtype)
    X = m.Vector(ctx, x, dtype=dtype)
    Y = m.Vector(ctx, y, dtype=dtype)
    d = m.dot(X, Y)
    assert d == float(oracle.dot(x, y))
    assert m.nrm2(X) == float(oracle.nrm2(x))


def test_dot_repeatable_across_blocks(m, ctx, oracle):
    """2048+ blocks publish partials that one block sums: 40 repeats must all
    give the oracle's bits (a stale or lost partial would show here)."""
    say("dot repeat")
    n = 5_000_003
    x = oracle.random_vector(n, -1.0, 1.0, 41)
    y = oracle.random_vector(n, -1.0, 1.0, 42)
    X = m.Vector(ctx, x)
    Y = m.Vector(ctx, y)
    ref = oracle.dot(x, y)
    got = {m.dot(X, Y) for _ in range(40)}
    assert got == {ref}


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_vector_ops_bit_exact(m, ctx, oracle, dtype):
    say("vector ops", dtype)
    n = 100003
    x = oracle.random_vector(n, -1.0, 1.0, 1, dtype=dtype)
    y = oracle.random_vector(n, -1.0, 1.0, 2, dtype=dtype)
    a = dtype(0.37)

    def run(fn, *vecs):
        hs = [m.Vector(ctx, v, dtype=dtype) for v in vecs]
        fn(*hs)
        return [h.download() for h in hs]

    assert bits_equal(run(lambda X, Y: m.axpy(float(a), X, Y), x, y)[1], oracle.axpy(a, x, y))
    assert bits_equal(run(lambda X, Y: m.xpay(float(a), X, Y), x, y)[1], oracle.xpay(a, x, y))
    assert bits_equal(run(lambda X: m.scal(float(a), X), x)[0], oracle.scal(a, x))
    assert bits_equal(run(lambda X, Y: m.copy(X, Y), x, y)[1], x)
    z = np.zeros(n, dtype)
    assert bits_equal(run(lambda X, Y, Z: m.axpyz(float(a), X, Y, Z), x, y, z)[2], oracle.axpyz(a, x, y))
    assert bits_equal(run(lambda X, Y, Z: m.vadd(X, Y, Z), x, y, z)[2], oracle.vadd(x, y))
    assert bits_equal(run(lambda X, Y, Z: m.vsub(X, Y, Z), x, y, z)[2], oracle.vsub(x, y))
    assert bits_equal(run(lambda X, Y, Z: m.vmul(X, Y, Z), x, y, z)[2], oracle.vmul(x, y))
    assert bits_equal(run(lambda X, Y: m.vreciprocal(X, Y), x, z)[1], oracle.vreciprocal(x))


def test_views_and_residency(m, ctx, oracle):
    say("views")
    n = 10000
    x = oracle.random_vector(n, -1.0, 1.0, 31)
    y = oracle.random_vector(n, -1.0, 1.0, 32)
    X = m.Vector(ctx, x)
    Y = m.Vector(ctx, y)
    xv, yv = X.view(5, n // 2), Y.view(5, n // 2)  # view1D(vec, 5, size/2)
    m.axpy(123.0, xv, yv)
    ref = y.copy()
    ref[5:5 + n // 2] = oracle.axpy(123.0, x[5:5 + n // 2], y[5:5 + n // 2])
    assert bits_equal(Y.download(), ref)
    assert m.dot(xv, yv) == oracle.dot(x[5:5 + n // 2], ref[5:5 + n // 2])
    X.upload(y)  # send() on a resident vector updates it
    assert bits_equal(X.download(), y)


def test_crs_diag(m, ctx, oracle):
    say("crs diag")
    rp = np.array([0, 3, 4, 4], np.int32)
    ci = np.array([0, 0, 2, 0], np.int32)
    val = np.array([1.0, 9.0, 5.0, 7.0])
    A = m.CRS(ctx, 3, 3, rp, ci, val)
    d = m.Vector(ctx, n=3)
    A.diag(d)
    assert d.download().tolist() == oracle.crs_diag(3, 3, rp, ci, val).tolist() == [9.0, 0.0, 0.0]


@pytest.mark.parametrize("rows,cols", [(256, 256), (1024, 1024), (1300, 1297), (700, 1000)])
def test_crs_diag_blocks(m, ctx, oracle, rows, cols):
    """diag over several full 256-row blocks: empty rows, duplicated diagonal
    entries (the last stored one wins), a long row, rectangular shapes."""
    say("crs diag blocks", rows, cols)
    rng = np.random.default_rng(rows + cols)
    lens = rng.integers(0, 6, rows)
    lens[rng.random(rows) < 0.2] = 0
    lens[rows // 3] = 600
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    nnz = int(rp[-1])
    ci = rng.integers(0, cols, nnz).astype(np.int32)
    for i in range(min(rows, cols)):
        if lens[i] > 0 and rng.random() < 0.7:
            ci[rp[i] + rng.integers(0, lens[i])] = i
        if lens[i] > 1 and rng.random() < 0.2:
            ci[rp[i]] = i
            ci[rp[i + 1] - 1] = i
    for dt in (np.float64, np.float32):
        val = rng.standard_normal(nnz).astype(dt)
        A = m.CRS(ctx, rows, cols, rp, ci, val)
        d = m.Vector(ctx, n=min(rows, cols), dtype=dt)
        A.diag(d)
        assert bits_equal(d.download(), oracle.crs_diag(rows, cols, rp, ci, val))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_matvec_plans_bit_identical(m, ctx, oracle, dtype):
    """Every SpMV plan the autotuner can pick -- products staged in LDS, the
    row walk, column panels (entries stored a second time panel-major, rows
    continuing their sums across the passes), XCD order on and off, entry
    budgets up to 4096 --
    computes the same bits: the CPU row loop's for rows up to the long-row
    threshold, one block tree for longer rows."""
    say("matvec plans", dtype.__name__)
    n, cols = 40000, 700000  # x of 700k columns: 5.6 MB (FP64)
    rp, ci, val = m.random_lognormal(n, cols, 24.0, 1.0, 900, seed=11)
    val = val.astype(dtype)
    x = oracle.random_vector(cols, -1.0, 1.0, 12, dtype=dtype)
    A = m.CRS(ctx, n, cols, rp, ci, val)
    X = m.Vector(ctx, x, dtype=dtype)
    Y = m.Vector(ctx, n=n, dtype=dtype)
    m.matvec(A, X, Y)
    base = Y.download()
    ref = oracle.spmv(rp, ci, val, x)
    short = np.diff(rp) <= 256
    assert short.sum() < n  # some long rows
    assert bits_equal(base[short], ref[short])
    tol = 1e-13 if dtype == np.float64 else 1e-5
    assert np.linalg.norm(base - ref) <= tol * np.linalg.norm(ref)
    for plan in [dict(rows_per_block=128, nnz_per_block=1024, row_walk=1),
                 dict(rows_per_block=256, nnz_per_block=1024, row_walk=1, xcd_remap=False),
                 dict(rows_per_block=256, nnz_per_block=2048, row_walk=1),
                 dict(rows_per_block=256, nnz_per_block=4096, row_walk=1),
                 dict(rows_per_block=7, nnz_per_block=4096, row_walk=1, xcd_remap=False),
                 dict(rows_per_block=64, nnz_per_block=4096),
                 dict(rows_per_block=128, nnz_per_block=1024, row_walk=1, xcd_remap=16),
                 dict(rows_per_block=64, nnz_per_block=4096, xcd_remap=3),
                 dict(rows_per_block=256, nnz_per_block=2048, column_panels=4, xcd_remap=5),
                 dict(rows_per_block=256, nnz_per_block=2048, column_panels=4),
                 dict(rows_per_block=128, nnz_per_block=1024, column_panels=16, xcd_remap=False),
                 dict(row_walk=3, column_panels=1), dict(row_walk=3, column_panels=6),
                 dict(row_walk=3, column_panels=64)]:
        A.set_plan(**plan)
        Y.fill(0.0)
        m.matvec(A, X, Y)
        assert bits_equal(Y.download(), base), plan
    for bad in (-1, 1025):
        with pytest.raises(m.MgxError):
            A.set_plan(rows_per_block=128, nnz_per_block=1024, xcd_remap=bad)
    t = A.autotune(reps=2)
    assert t["candidates"] >= 26 + 3 * (dtype == np.float64)
    # this matrix's row lengths vary too much for the sliced layout
    with pytest.raises(m.MgxError):
        A.set_plan(row_walk=2)
    Y.fill(0.0)
    m.matvec(A, X, Y)
    assert bits_equal(Y.download(), base), t


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_matvec_l2_panels(m, ctx, oracle, dtype):
    """The L2-panel plan (row sums in LDS, column panels walked in order,
    entries grouped by rank within their row's run of a panel): the CPU row
    loop's bits for every short row -- with empty rows, rows whose entries
    fall in one panel (deep rank groups), rows above the long-row threshold
    (the block tree, as every plan) -- for y = A x and y = a A x + b y, for
    panel counts from 1 to a panel per 1000 columns; update_values drops the
    layout and the block plan gives the same bits."""
    say("matvec l2 panels", dtype.__name__)
    rng = np.random.default_rng(21)
    n, cols = 30000, 250000
    lens = np.clip(np.round(np.exp(rng.normal(np.log(24) - 0.5, 1.0, n))), 0, 1500).astype(int)
    lens[::97] = 0
    lens[5::1013] = 700
    rows_ci = []
    for r, ln in enumerate(lens):
        if r % 777 == 3:
            rows_ci.append(np.arange(200, dtype=np.int64) * 3)  # 200 entries, one panel
        else:
            rows_ci.append(np.sort(rng.choice(cols, ln, replace=False)))
    lens = np.array([len(c) for c in rows_ci])
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    ci = np.concatenate(rows_ci).astype(np.int32)
    val = rng.uniform(-1, 1, rp[-1]).astype(dtype)
    x = oracle.random_vector(cols, -1.0, 1.0, 22, dtype=dtype)
    y0 = oracle.random_vector(n, -1.0, 1.0, 23, dtype=dtype)
    A = m.CRS(ctx, n, cols, rp, ci, val)
    X = m.Vector(ctx, x, dtype=dtype)
    Y = m.Vector(ctx, n=n, dtype=dtype)
    A.set_plan(rows_per_block=128, nnz_per_block=2048)
    m.matvec(A, X, Y)
    base = Y.download()
    Y.upload(y0)
    m.matvec_ab(0.75, A, X, -1.5, Y)
    base_ab = Y.download()
    ref = oracle.spmv(rp, ci, val, x)
    short = np.diff(rp) <= 256
    assert (~short).sum() > 10 and bits_equal(base[short], ref[short])
    for panels in (1, 3, 17, 250):
        A.set_plan(row_walk=3, column_panels=panels)
        Y.fill(np.nan)
        m.matvec(A, X, Y)
        assert bits_equal(Y.download(), base), panels
        Y.upload(y0)
        m.matvec_ab(0.75, A, X, -1.5, Y)
        assert bits_equal(Y.download(), base_ab), panels
    A.update_values((val * 2).astype(dtype))
    m.matvec(A, X, Y)
    ref2 = oracle.spmv(rp, ci, (val * 2).astype(dtype), x)
    assert bits_equal(Y.download()[short], ref2[short])
    # unsorted columns within a row: not applicable
    ci_bad = ci.copy()
    r0 = int(np.argmax(lens > 3))
    ci_bad[rp[r0]], ci_bad[rp[r0] + 1] = ci[rp[r0] + 1], ci[rp[r0]]
    B = m.CRS(ctx, n, cols, rp, ci_bad, val)
    with pytest.raises(m.MgxError):
        B.set_plan(row_walk=3, column_panels=4)
    for h in (A, B, X, Y):
        h.free()


def test_split_cg_with_l2_panel_spmv(m, ctx, oracle):
    """The split CG iteration (n >= 2^19) runs the matrix's SpMV plan with
    the solver's stop flag; under the L2-panel plan 40 iterations are
    bit-identical to the oracle."""
    say("split cg l2 panels")
    rp, ci, val = m.stencil3d(96, 96, 64, 7, 0.0)
    n = len(rp) - 1
    b = oracle.random_vector(n, -1.0, 1.0, 3001)
    ref = oracle.solve("cg", rp, ci, val, np.zeros(n), b, precond=2, tol=1e-30, maxiter=40,
                       history=True)
    A = m.CRS(ctx, n, n, rp, ci, val)
    A.set_plan(row_walk=3, column_panels=5)
    X = m.Vector(ctx, n=n)
    res = m.cg(A, X, m.Vector(ctx, b), tol=1e-30, maxiter=40, precond=2, mode=m.MODE_FUSED,
               history=40)
    assert res.status == ref[0] and res.final_iter == 40
    assert bits_equal(X.download(), ref[1]) and bits_equal(res.rhistory, ref[4])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_matvec_sliced_layout(m, ctx, oracle, dtype):
    """The sliced layout (64-row slices, one lane per row) computes the bits
    of every other plan: 27-point and 7-point stencils, a matrix with rows
    above the long-row threshold and short rows of 255 and 256 entries (the
    layout's byte lengths send those to row_ptr), y = a A x + b y, after
    update_values, and inside the mixed-precision solver (its FP32 copy of
    the values)."""
    say("matvec sliced", dtype.__name__)
    rng = np.random.default_rng(5)
    mats = [m.stencil3d(24, 20, 17, 27, 1e-2), m.stencil3d(40, 33, 9, 7)]
    lens = rng.integers(10, 13, 9000)  # 64-row slices padded by under 10 %
    lens[[17, 4000, 8999]] = [600, 300, 1000]  # long rows: the block-tree path
    lens[[5000, 5001]] = [255, 256]  # short rows whose length row_ptr gives
    rp = np.zeros(len(lens) + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    cols = 12000
    ci = np.concatenate([np.sort(rng.choice(cols, l, replace=False)) for l in lens])
    mats.append((rp, ci.astype(np.int32), rng.uniform(-1, 1, rp[-1])))
    for k, (rp, ci, val) in enumerate(mats):
        val = val.astype(dtype)
        rows = len(rp) - 1
        ncol = cols if k == 2 else rows
        x = oracle.random_vector(ncol, -1.0, 1.0, 40 + k, dtype=dtype)
        A = m.CRS(ctx, rows, ncol, rp, ci, val)
        X = m.Vector(ctx, x, dtype=dtype)
        Y = m.Vector(ctx, n=rows, dtype=dtype)
        m.matvec(A, X, Y)
        base = Y.download()
        ref = oracle.spmv(rp, ci, val, x)
        short = np.diff(rp) <= 256
        assert bits_equal(base[short], ref[short])
        y0 = oracle.random_vector(rows, -1.0, 1.0, 60 + k, dtype=dtype)
        Y.upload(y0)
        m.matvec_ab(0.5, A, X, -2.0, Y)
        base_ab = Y.download()
        A.set_plan(rows_per_block=128, nnz_per_block=1024, row_walk=2)
        Y.fill(0.0)
        m.matvec(A, X, Y)
        assert bits_equal(Y.download(), base), k
        Y.upload(y0)
        m.matvec_ab(0.5, A, X, -2.0, Y)
        assert bits_equal(Y.download(), base_ab), k
        v2 = (val * 3).astype(dtype)
        A.update_values(v2)
        B = m.CRS(ctx, rows, ncol, rp, ci, v2)
        Y2 = m.Vector(ctx, n=rows, dtype=dtype)
        m.matvec(A, X, Y)
        m.matvec(B, X, Y2)
        assert bits_equal(Y.download(), Y2.download()), k
        for o in (A, B, X, Y, Y2):
            o.free()
    if dtype == np.float64:
        # mixed precision: FP32 iterations through the sliced FP32 values
        rp, ci, val = m.stencil3d(32, 32, 32, 27, 1e-2)
        n = len(rp) - 1
        b = oracle.random_vector(n, -1.0, 1.0, 3001)
        xs = []
        for sliced in (False, True):
            A = m.CRS(ctx, n, n, rp, ci, val)
            if sliced:
                A.set_plan(row_walk=2)
            X = m.Vector(ctx, np.zeros(n))
            B = m.Vector(ctx, b)
            r = m.mixed_cg(A, X, B, tol=1e-10, delta=1e-2)
            assert r.status == 0
            xs.append(X.download())
            for o in (A, X, B):
                o.free()
        assert bits_equal(xs[0], xs[1])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dot_nrm2_device_results(m, ctx, oracle, dtype):
    """mgx_dot_device / mgx_nrm2_device leave the result in device memory:
    the same bits mgx_dot / mgx_nrm2 return, and the oracle's."""
    say("dot/nrm2 device results", dtype.__name__)
    out = m.Vector(ctx, n=4)
    for k, n in enumerate((1, 777, 1_000_003)):
        x = oracle.random_vector(n, -1.0, 1.0, 300 + k, dtype=dtype)
        y = oracle.random_vector(n, -1.0, 1.0, 400 + k, dtype=dtype)
        X, Y = m.Vector(ctx, x, dtype=dtype), m.Vector(ctx, y, dtype=dtype)
        m.dot_device(X, Y, out, 1)
        m.nrm2_device(X, out, 3)
        got = out.download()
        assert got[1] == m.dot(X, Y) == float(oracle.dot(x, y))
        assert got[3] == m.nrm2(X) == float(oracle.nrm2(x))
        X.free()
        Y.free()
    out.free()


def test_caller_stream(m, oracle):
    """mgx_set_stream: a context queues its work on the caller's stream; an
    SpMV, a dot and a fused CG solve there give the bits they give on the
    library's own stream; NULL restores the library's stream."""
    import ctypes as C
    say("caller stream")
    hip = C.CDLL("libamdhip64.so")
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    st = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(st)) == 0
    c = m.Context(0)
    try:
        rp, ci, val = m.laplacian2d(60, 60)
        n = len(rp) - 1
        b = oracle.random_vector(n, -1.0, 1.0, 5)
        A, B = m.CRS(c, n, n, rp, ci, val), m.Vector(c, b)

        def run():
            Y, X = m.Vector(c, n=n), m.Vector(c, n=n)
            m.matvec(A, B, Y)
            d = m.dot(Y, B)
            r = m.cg(A, X, B, tol=1e-10, tol_is_relative=True, precond=m.PRECOND_JACOBI)
            return Y.download(), d, r.final_iter, X.download()

        own = c.stream
        base = run()
        m.set_stream(c, st.value)
        assert c.stream == st.value
        on = run()
        m.set_stream(c, None)
        assert c.stream == own
        assert np.array_equal(base[0], on[0]) and base[1] == on[1] and base[2] == on[2]
        assert np.array_equal(base[3].view(np.uint64), on[3].view(np.uint64))
        ref = oracle.solve("cg", rp, ci, val, np.zeros(n), b, precond=2,
                           tol=1e-10 * oracle.nrm2(b))
        assert np.array_equal(on[3].view(np.uint64), ref[1].view(np.uint64))
    finally:
        c.close()
        hip.hipStreamDestroy(st)
