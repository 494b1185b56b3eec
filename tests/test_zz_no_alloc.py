"""The allocation-free dot (mgx_dot_device_noalloc) against the allocating
one (mgx_dot_device) and the oracle, through the C-ABI: the same bits on
two calls in a row, the second on a smaller input (the static scratch keeps
the first call's partials and counters), in FP64 and FP32; inputs beyond the
scratch's limit are refused before anything is launched."""
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


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_dot_equals_dot_device(m, oracle, dtype):
    c = m.Context(0)
    try:
        out = m.Vector(c, n=4)
        # first call large (many tiles, the general arrival tree), second
        # smaller (at most 256 tiles, the short arrival path), then a third
        # large one again after the small one
        for k, n in enumerate((3_000_017, 70_001, 1_048_576, 1)):
            say("noalloc dot", dtype.__name__, n)
            x = oracle.random_vector(n, -1.0, 1.0, 600 + k, dtype=dtype)
            y = oracle.random_vector(n, -1.0, 1.0, 700 + k, dtype=dtype)
            X, Y = m.Vector(c, x, dtype=dtype), m.Vector(c, y, dtype=dtype)
            m.dot_device(X, Y, out, 0)
            m.dot_device_noalloc(X, Y, out, 2)
            got = out.download()
            assert got[0] == got[2], (n, got[0], got[2])
            assert got[2] == float(oracle.dot(x, y))
            X.free()
            Y.free()
        out.free()
    finally:
        c.close()


def test_noalloc_dot_refuses_beyond_limit(m):
    c = m.Context(0)
    try:
        say("noalloc dot limit")
        n = (1 << 26) + 1
        X = m.Vector(c, n=n, dtype=np.float32)
        out = m.Vector(c, n=1)
        with pytest.raises(m.MgxError) as e:
            m.dot_device_noalloc(X, X, out, 0)
        assert e.value.args[0] == -101 or "limit" in str(e.value)
        X.free()
        out.free()
    finally:
        c.close()
