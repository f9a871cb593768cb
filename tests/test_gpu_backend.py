"""The DenseBackend entry points (crates/einfold-kernels/src/backend.rs:
124-143) through the Python callback form (einfold_amd.backend), against
the float64 NumPy formula of the reference's own backend test
(python/tests/test_dense_backend.py:21-49) at that test's tolerance
(rtol 1e-9, atol 1e-12)."""
import numpy as np
import pyarrow as pa
import pytest

from einfold_amd.backend import Mi355xBackend

pytestmark = pytest.mark.gpu


def numpy_backend(kind, arrays, shape, scale, ranges):
    arrays = [a.to_numpy() for a in arrays]
    if kind == "matmul":
        w, x, z, y = shape
        return (arrays[0].reshape(w, x, z) @ arrays[1].reshape(w, z, y)).reshape(-1)
    w, x, e, j, y = shape
    q, kt, v = arrays[0].reshape(w, x, e), arrays[1].reshape(w, e, j), arrays[2].reshape(w, j, y)
    code, c = scale
    s = q @ kt
    s = {0: s, 1: s * c, 2: s / c}[code]
    allowed = np.zeros((x, j), dtype=bool)
    for i, (lo, hi) in enumerate(ranges):
        allowed[i, lo:hi] = True
    s = np.where(allowed, s, -np.inf)
    with np.errstate(invalid="ignore"):
        p = np.exp(s - s.max(axis=2, keepdims=True))
        p /= p.sum(axis=2, keepdims=True)
    return (p @ v).reshape(-1)


@pytest.mark.parametrize("shape", [(1, 64, 64, 64), (3, 17, 33, 9), (2, 130, 257, 70), (1, 1, 5000, 3)])
def test_matmul(cuda, shape):
    rng = np.random.default_rng(sum(shape))
    w, x, z, y = shape
    arrays = [pa.array(rng.uniform(-1, 1, w * x * z)), pa.array(rng.uniform(-1, 1, w * z * y))]
    be = Mi355xBackend()
    got = be("matmul", arrays, shape).to_numpy()
    want = numpy_backend("matmul", arrays, shape, None, None)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12)
    assert be.calls == [("matmul", shape)]


# Scale::Over with c = 3 (no exact reciprocal): the device multiplies by the
# rounded 1/c, one more rounding per score than the division, inside the
# test's tolerance.
@pytest.mark.parametrize("code,c", [(0, 1.0), (1, 0.125), (2, 8.0), (2, 3.0), (1, 0.3)])
@pytest.mark.parametrize("shape,mask", [((2, 200, 64, 200, 64), "causal"), ((1, 100, 16, 150, 8), "window"),
                                        ((3, 70, 8, 90, 5), "none")])
def test_attention(cuda, shape, mask, code, c):
    rng = np.random.default_rng(shape[1] + code)
    w, x, e, j, y = shape
    arrays = [pa.array(rng.normal(size=w * x * e) * 0.3), pa.array(rng.normal(size=w * e * j) * 0.3),
              pa.array(rng.normal(size=w * j * y))]
    if mask == "causal":
        ranges = [(0, min(i + 1, j)) for i in range(x)]
    elif mask == "window":
        ranges = [(max(0, i - 20), min(i + 1, j)) for i in range(x)]
    else:
        ranges = [(0, j)] * x
    got = Mi355xBackend()("attention", arrays, shape, (code, c), ranges).to_numpy().reshape(w, x, y)
    want = numpy_backend("attention", arrays, shape, (code, c), ranges).reshape(w, x, y)
    rows = np.array([lo < hi for lo, hi in ranges])
    np.testing.assert_allclose(got[:, rows], want[:, rows], rtol=1e-9, atol=1e-12)


def test_may_run_declines_what_does_not_fit(cuda):
    be = Mi355xBackend()
    assert be.may_run("matmul", (1, 64, 64, 64))
    assert not be.may_run("matmul", (1, 1 << 30, 1 << 20, 1 << 20))
    assert be("matmul", [pa.array([0.0]), pa.array([0.0])], (1, 1 << 30, 1 << 20, 1 << 20)) is None
    assert not be.may_run("attention", (1 << 20, 1 << 20, 64, 1 << 20, 64))
