"""GPU: hip_dsyrk_gram_noalloc -- the allocation-free variant of
hip_dsyrk_gram whose k-split partials live in the library's fixed device
array -- against hip_dsyrk_gram through the C-ABI: bit for bit (same kernels,
same k-split) on two calls in a row, the second smaller, so it also reads a
scratch array that still holds the first call's partials."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gram(lib, fn, n, k, dA, dC, alpha=1.0, beta=0.0, mirror=1):
    return getattr(lib, fn)(n, k, alpha, dA.ptr, k, beta, dC.ptr, n, mirror)


@pytest.mark.parametrize("kind", ["random", "integer"])
def test_noalloc_equals_gram_on_two_calls(hip, kind):
    from ca3dmm_amd._lib import lib
    rnd = np.random.default_rng(77)
    for n, k in ((512, 300000), (200, 30001)):            # second call smaller
        print(f"[no_alloc] {kind} n={n} k={k}", flush=True)
        if kind == "integer":
            A = rnd.integers(-3, 4, (k, n)).astype(np.float64)
        else:
            A = rnd.standard_normal((k, n))
        dA = hip.DeviceArray.from_numpy(A)
        X = rnd.standard_normal((n, n))
        Cin = X + X.T            # symmetric: the mirrored lower triangle is beta-scaled from the upper one
        want = hip.DeviceArray.from_numpy(Cin)
        got = hip.DeviceArray.from_numpy(Cin)
        assert _gram(lib, "hip_dsyrk_gram", n, k, dA, want, 1.5, -0.25) == 0
        assert _gram(lib, "hip_dsyrk_gram_noalloc", n, k, dA, got, 1.5, -0.25) == 0
        w, g = want.numpy(), got.numpy()
        assert np.array_equal(g, w)
        if kind == "integer":
            assert np.array_equal(g, 1.5 * (A.T @ A) + (-0.25) * Cin)


def test_noalloc_refuses_beyond_its_scratch(hip):
    from ca3dmm_amd._lib import lib
    n, k = 4096, 256
    dA = hip.DeviceArray.zeros((k, n))
    dC = hip.DeviceArray.zeros((n, n))
    rc = _gram(lib, "hip_dsyrk_gram_noalloc", n, k, dA, dC)
    assert rc < 0
    assert b"partial" in lib.hip_proxy_last_error()
    # a shape inside the limit still works after the refusal
    n, k = 300, 5000
    A = np.random.default_rng(1).integers(-2, 3, (k, n)).astype(np.float64)
    dA = hip.DeviceArray.from_numpy(A)
    dC = hip.DeviceArray.zeros((n, n))
    assert _gram(lib, "hip_dsyrk_gram_noalloc", n, k, dA, dC) == 0
    assert np.array_equal(dC.numpy(), A.T @ A)
