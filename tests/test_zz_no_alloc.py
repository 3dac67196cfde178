"""mcac_swept_sphere_batch_noalloc (static device staging, no allocation)
against mcac_swept_sphere_batch (allocating) through the C-ABI: bit-identical
results on two calls in a row, the second smaller, and the size limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    L = 2.4e-6
    p1 = rng.uniform(-0.2 * L, 1.2 * L, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r1 = rng.uniform(5e-9, 40e-9, n)
    r2 = rng.uniform(5e-9, 40e-9, n)
    dist = rng.uniform(1e-8, 1.5 * L, n)
    t = rng.uniform(0, 1, n)[:, None] * dist[:, None]
    p2 = p1 + t * d + rng.normal(scale=1, size=(n, 3)) * (r1 + r2)[:, None]
    p2[1::11] = p1[1::11] + d[1::11] * (r1 + r2)[1::11, None]  # touching
    p2[2::13] += L * rng.integers(-2, 3, (len(p2[2::13]), 3))  # other periodic images
    return p1, r1, p2, r2, d, dist, L


def _same(a, b):
    return bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def test_noalloc_matches_allocating_entry_point():
    import mcac_amd
    print("case: swept-sphere batch noalloc vs allocating, n=65536 then n=777", flush=True)
    for n, seed in ((65536, 1), (777, 2)):
        args = _inputs(n, seed)
        ref = mcac_amd.swept_sphere_batch(*args)
        got = mcac_amd.swept_sphere_batch(*args, noalloc=True)
        assert np.isfinite(ref).sum() > n // 10
        assert _same(got, ref), n


def test_noalloc_rejects_oversized_batch():
    import mcac_amd
    n = (1 << 16) + 1
    p1, r1, p2, r2, d, dist, L = _inputs(n, 3)
    out = np.full(n, -1.0)
    dp = lambda a: np.ascontiguousarray(a, np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rc = mcac_amd.load_gpu().mcac_swept_sphere_batch_noalloc(n, dp(p1), dp(r1), dp(p2), dp(r2), dp(d), dp(dist), L,
                                                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 0)
    assert rc == 33  # MCAC_E_CAPACITY
    assert np.all(out == -1.0)
