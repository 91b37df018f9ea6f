"""Allocation-free dot (su2_tensor_dot_noalloc) against su2_tensor_dot (GPU).

Two calls in a row through the C ABI, the second on a smaller tensor, each
compared bit for bit with the allocating entry point (same kernels, same
summation order); also a fresh context whose first call is the no-alloc one.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pair(L, ctx, dims, seed):
    twoj = [1, 3, 5][:len(dims)]
    rng = np.random.default_rng(seed)
    x = L.wavefunction(ctx, twoj, dims, dims)
    y = L.wavefunction(ctx, twoj, dims, dims)
    x.upload(rng.standard_normal(x.size))
    y.upload(rng.standard_normal(y.size))
    return x, y


def test_dot_noalloc_matches_dot_twice():
    from su2dmrg_amd import lib as L
    import ctypes as C
    b = C.c_longlong()
    assert L.lib.su2_noalloc_scratch_bytes(C.byref(b)) == 0
    assert b.value == 8 * 513   # 512 partial sums and the result
    ctx = L.Context(0)
    # first call on a fresh context is the allocation-free one
    for dims, seed in (([900, 700, 300], 1), ([40, 17], 2)):
        _say(f"dot_noalloc dims={dims} seed={seed}")
        x, y = _pair(L, ctx, dims, seed)
        got = x.dot_noalloc(y)
        ref = x.dot(y)
        n = x.size
        assert got == ref, (got, ref, n)
        host = float(np.dot(x.download(), y.download()))
        assert abs(got - host) <= 1e-12 * max(1.0, abs(host)) * np.sqrt(n)
        x.close()
        y.close()
    ctx.close()


def test_dot_noalloc_empty_and_mismatch():
    from su2dmrg_amd import lib as L
    ctx = L.Context(0)
    x, y = _pair(L, ctx, [5, 3], 3)
    z, _ = _pair(L, ctx, [6, 3], 4)
    with pytest.raises(L.Su2Error):
        x.dot_noalloc(z)
    assert x.dot_noalloc(y) == x.dot(y)
    ctx.close()
