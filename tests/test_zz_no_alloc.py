"""xh_eval_bumps_noalloc (no device allocation: counts and bump parameters staged in the
library's static device array) against xh_eval_bumps (hipMalloc'd staging) through the C
ABI: bit-identical right-hand sides on two calls in a row, the second smaller, and sizes
beyond the staging array rejected before any launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _params(X, seed, nx, ny, nz, nv, max_bumps):
    count = np.zeros(nv, np.int32)
    P = np.zeros(nv * max_bumps * 5)
    X.check(X.host().xamg_bump_params(seed, nx, ny, nz, nv, 8, max_bumps, 2.0, 12.0,
                                      X._ptr(count), X._ptr(P)), "xamg_bump_params")
    return count, P


def _both(X, seed, nx, ny, nz, nv, max_bumps):
    count, P = _params(X, seed, nx, ny, nz, nv, max_bumps)
    n = nx * ny * nz * nv
    ref, out = X.DeviceArray(n), X.DeviceArray(n)
    X.check(X.hip().xh_eval_bumps(nx, ny, nz, nv, max_bumps, X._ptr(count), X._ptr(P), ref.ptr),
            "xh_eval_bumps")
    X.check(X.hip().xh_eval_bumps_noalloc(nx, ny, nz, nv, max_bumps, X._ptr(count), X._ptr(P),
                                          out.ptr), "xh_eval_bumps_noalloc")
    return ref.to_host(), out.to_host(), count


def test_noalloc_matches_allocating_entry_on_two_calls(gpu):
    X = gpu
    # first call: the largest staging the array holds (nv = 64, 64 bumps per RHS)
    ref, out, count = _both(X, 11, 24, 20, 16, 64, 64)
    assert count.max() > 32
    assert np.array_equal(ref, out)
    assert np.abs(out).max() > 0
    # second call, smaller: fewer RHS and bumps; whatever the first call left beyond the
    # second's extent must not be read
    ref2, out2, _ = _both(X, 12, 17, 13, 9, 4, 16)
    assert np.array_equal(ref2, out2)
    assert np.abs(out2).max() > 0


def test_noalloc_rejects_sizes_beyond_the_staging_array(gpu):
    X = gpu
    nv, maxb = 2, 65
    count = np.full(nv, 8, np.int32)
    P = np.zeros(nv * maxb * 5)
    out = X.DeviceArray(8 * 8 * 8 * 128)
    rc = X.hip().xh_eval_bumps_noalloc(8, 8, 8, nv, maxb, X._ptr(count), X._ptr(P), out.ptr)
    assert rc == -1
    assert b"at most 64" in X.hip().xh_last_error()
    rc = X.hip().xh_eval_bumps_noalloc(8, 8, 8, 128, 8, X._ptr(np.full(128, 8, np.int32)),
                                       X._ptr(np.zeros(128 * 8 * 5)), out.ptr)
    assert rc == -1
