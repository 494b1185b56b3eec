"""GPU parity of the MatrixMarket / COO ingest against the oracle's
restatement of COO::input_mm, COO::sort and CRS(COO&): structure and values
bit for bit, including the reference's own test matrices, unsorted and
duplicated entries, empty rows and the FP32 reader."""
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


def bits(a):
    a = np.asarray(a)
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def oracle_crs(oracle, rows, cols, r, c, v, sort):
    if sort:
        r, c, v = oracle.coo_sort(r, c, v, merge=sort == 2)
    return oracle.crs_from_coo(rows, cols, r, c, v)


def stable_merge(rows, r, c, v):
    """sort = 2 semantics: stable (row, col) order, the last duplicate in
    input order kept."""
    order = np.lexsort((np.arange(len(r)), c, r))
    r, c, v = r[order], c[order], v[order]
    last = np.ones(len(r), bool)
    if len(r) > 1:
        last[:-1] = (r[1:] != r[:-1]) | (c[1:] != c[:-1])
    return r[last], c[last], v[last]


@pytest.mark.parametrize("name", ["test.mtx", "test2.mtx"])
@pytest.mark.parametrize("sort", [0, 1, 2])
def test_read_mm_reference_matrices(m, ctx, oracle, name, sort):
    say("read_mm", name, sort)
    path = os.path.join(GOLDEN, name)
    rows, cols, r, c, v = oracle.input_mm(path)
    rp0, ci0, v0 = oracle_crs(oracle, rows, cols, r, c, v, sort)
    A = m.CRS.read_mm(ctx, path, sort=sort)
    assert (A.rows, A.cols, A.nnz) == (rows, cols, len(v0))
    rp, ci, val = A.download()
    assert np.array_equal(rp, rp0) and np.array_equal(ci, ci0)
    assert np.array_equal(bits(val), bits(v0))
    # float reader: the text rounded once to float
    _, _, _, _, vf = oracle.input_mm(path, is_float=True)
    Af = m.CRS.read_mm(ctx, path, dtype=np.float32, sort=sort)
    _, _, valf = Af.download()
    _, _, vf0 = oracle_crs(oracle, rows, cols, r, c, np.asarray(vf, np.float64), sort)
    assert np.array_equal(bits(valf), bits(vf0.astype(np.float32)))
    A.free()
    Af.free()


@pytest.mark.parametrize("n,nnz,seed", [(1, 1, 1), (37, 500, 2), (5000, 120000, 3),
                                        (300000, 2000000, 4)])
def test_from_coo_shuffled_sort(m, ctx, oracle, n, nnz, seed):
    """Distinct (row, col) in random order: sort = 1 gives the (row, col)
    order of COO::sort; sort = 2 changes nothing; every tile boundary of the
    device sort is crossed at the larger sizes."""
    say("from_coo shuffled", n, nnz, seed)
    rng = np.random.default_rng(seed)
    cols = max(n, 3)
    keys = np.unique(rng.integers(0, n * cols, nnz))
    rng.shuffle(keys)
    r = (keys // cols).astype(np.int32)
    c = (keys % cols).astype(np.int32)
    v = rng.uniform(-1, 1, len(keys))
    srt = np.argsort(r.astype(np.int64) * cols + c, kind="stable")
    rp0, ci0, v0 = oracle.crs_from_coo(n, cols, r[srt], c[srt], v[srt])
    if len(keys) <= 200000:
        ro, co, vo = oracle.coo_sort(r, c, v, merge=False)
        rpo, cio, vao = oracle.crs_from_coo(n, cols, ro, co, vo)
        assert np.array_equal(rpo, rp0) and np.array_equal(cio, ci0)
        assert np.array_equal(bits(vao), bits(v0))
    for sort in (1, 2):
        A = m.CRS.from_coo(ctx, n, cols, r, c, v, sort=sort)
        rp, ci, val = A.download()
        assert A.nnz == len(keys)
        assert np.array_equal(rp, rp0) and np.array_equal(ci, ci0), sort
        assert np.array_equal(bits(val), bits(v0)), sort
        A.free()


def test_from_coo_duplicates_merge(m, ctx, oracle):
    """The reference's duplicate known answer (matrix_common.cpp:185-209:
    a later (0,0) = 3 replaces the first) and random duplicates: each run of
    equal (row, col) keeps the value of the last one in input order."""
    say("duplicates")
    r = np.array([0, 0, 1, 1, 2, 2, 3, 0], np.int32)
    c = np.array([0, 1, 0, 2, 1, 3, 3, 0], np.int32)
    v = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 1.0, 3.0])
    ro, co, vo = oracle.coo_sort(r, c, v, merge=True)
    rp0, ci0, v0 = oracle.crs_from_coo(4, 4, ro, co, vo)
    A = m.CRS.from_coo(ctx, 4, 4, r, c, v, sort=2)
    rp, ci, val = A.download()
    assert A.nnz == 7 and np.array_equal(rp, rp0) and np.array_equal(ci, ci0)
    assert np.array_equal(val, v0) and val[0] == 3.0
    A.free()
    rng = np.random.default_rng(9)
    n, nnz = 2000, 400000
    r = rng.integers(0, n, nnz).astype(np.int32)
    c = rng.integers(0, 50, nnz).astype(np.int32)
    v = rng.standard_normal(nnz)
    rs, cs, vs = stable_merge(n, r, c, v)
    rp0, ci0, v0 = oracle.crs_from_coo(n, 50, rs, cs, vs)
    A = m.CRS.from_coo(ctx, n, 50, r, c, v, sort=2)
    rp, ci, val = A.download()
    assert A.nnz == len(vs) < nnz
    assert np.array_equal(rp, rp0) and np.array_equal(ci, ci0)
    assert np.array_equal(bits(val), bits(v0))
    # sort = 1 keeps every duplicate, in input order within each (row, col)
    A1 = m.CRS.from_coo(ctx, n, 50, r, c, v, sort=1)
    rp1, ci1, v1 = A1.download()
    srt = np.lexsort((np.arange(nnz), c, r))
    rp1r, ci1r, v1r = oracle.crs_from_coo(n, 50, r[srt], c[srt], v[srt])
    assert np.array_equal(rp1, rp1r) and np.array_equal(ci1, ci1r)
    assert np.array_equal(bits(v1), bits(v1r))
    A.free()
    A1.free()


def test_from_coo_empty_rows_and_errors(m, ctx, oracle):
    """Leading, interior and trailing empty rows (row_ptr of trailing rows
    reads back as 0, as CRS(COO&) leaves it); rows out of order without
    sort, indices outside the matrix and an empty COO."""
    say("empty rows / errors")
    r = np.array([1, 1, 3], np.int32)
    c = np.array([0, 2, 1], np.int32)
    v = np.array([1.0, 2.0, 3.0])
    rp0, ci0, v0 = oracle.crs_from_coo(6, 3, r, c, v)
    A = m.CRS.from_coo(ctx, 6, 3, r, c, v)
    rp, ci, val = A.download()
    assert rp.tolist() == rp0.tolist() == [0, 0, 2, 2, 3, 0, 0]
    x = m.Vector(ctx, np.array([1.0, 10.0, 100.0]))
    y = m.Vector(ctx, np.full(6, 7.0))
    m.matvec(A, x, y)
    assert y.download().tolist() == [0.0, 201.0, 0.0, 30.0, 0.0, 0.0]
    A.free()
    with pytest.raises(m.MgxError):
        m.CRS.from_coo(ctx, 6, 3, np.array([2, 1], np.int32), np.array([0, 0], np.int32),
                       np.ones(2), sort=0)
    with pytest.raises(m.MgxError):
        m.CRS.from_coo(ctx, 6, 3, np.array([0, 6], np.int32), np.array([0, 0], np.int32),
                       np.ones(2), sort=1)
    with pytest.raises(m.MgxError):
        m.CRS.from_coo(ctx, 6, 3, np.array([0], np.int32), np.array([-1], np.int32), np.ones(1))
    E = m.CRS.from_coo(ctx, 4, 4, np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0),
                       sort=2)
    rp, ci, val = E.download()
    assert E.nnz == 0 and rp.tolist() == [0] * 5
    E.free()


def test_read_mm_rejects(m, ctx, tmp_path):
    say("read_mm rejects")
    cases = {
        "sym.mtx": "%%MatrixMarket matrix coordinate real symmetric\n2 2 1\n1 1 1\n",
        "arr.mtx": "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",
        "short.mtx": "%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1\n2 2 2\n",
        "range.mtx": "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n",
        "text.mtx": "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 x\n",
    }
    for name, text in cases.items():
        p = tmp_path / name
        p.write_text(text)
        with pytest.raises(m.MgxError):
            m.CRS.read_mm(ctx, str(p))
    with pytest.raises(m.MgxError):
        m.CRS.read_mm(ctx, str(tmp_path / "missing.mtx"))


def test_read_mm_comments_signs_and_float(m, ctx, oracle, tmp_path):
    say("read_mm comments / float")
    p = tmp_path / "c.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n% c1\n%c2\n3 3 4\n"
                 "1 3 0.1\n2 1 -2.5e-3\n2 2 +7\n3 3 1e-320\n")
    rows, cols, r, c, v = oracle.input_mm(str(p))
    rp0, ci0, v0 = oracle.crs_from_coo(rows, cols, r, c, v)
    A = m.CRS.read_mm(ctx, str(p))
    rp, ci, val = A.download()
    assert np.array_equal(rp, rp0) and np.array_equal(ci, ci0)
    assert np.array_equal(bits(val), bits(v0))
    Af = m.CRS.read_mm(ctx, str(p), dtype=np.float32)
    _, _, vf = Af.download()
    assert vf.tolist()[:3] == [float(np.float32(0.1)), float(np.float32(-2.5e-3)), 7.0]
    A.free()
    Af.free()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_write_read_round_trip(m, ctx, oracle, tmp_path, dtype):
    """COO(A).output_mm then input_mm: max_digits10 scientific text brings
    back every bit; the bytes match a %.17e / %.9e formatting of the entries;
    the SpMV of the ingested matrix equals the original's bit for bit."""
    say("write/read round trip", dtype.__name__)
    rp, ci, val = m.random_lognormal(20000, 30000, 12.0, 1.0, 300, seed=21)
    val = val.astype(dtype)
    A = m.CRS(ctx, 20000, 30000, rp, ci, val)
    p = str(tmp_path / "rt.mtx")
    A.write_mm(p)
    with open(p) as f:
        head = [next(f) for _ in range(3)]
    fmt = "%.17e" if dtype == np.float64 else "%.9e"
    assert head[0] == "%%MatrixMarket matrix coordinate real general\n"
    assert head[1] == f"20000 30000 {len(val)}\n"
    assert np.all(np.diff(rp) >= 1)  # every row has an entry: the first line is row 1
    assert head[2] == f"1 {ci[0] + 1} {fmt % float(val[0])}\n"
    B = m.CRS.read_mm(ctx, p, dtype=dtype)
    rp2, ci2, v2 = B.download()
    assert np.array_equal(rp2, rp) and np.array_equal(ci2, ci)
    assert np.array_equal(bits(v2), bits(val))
    x = m.Vector(ctx, oracle.random_vector(30000, -1.0, 1.0, 5, dtype=dtype), dtype=dtype)
    y1 = m.Vector(ctx, n=20000, dtype=dtype)
    y2 = m.Vector(ctx, n=20000, dtype=dtype)
    m.matvec(A, x, y1)
    m.matvec(B, x, y2)
    assert np.array_equal(bits(y1.download()), bits(y2.download()))
    A.free()
    B.free()
