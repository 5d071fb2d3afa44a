"""sc1d_march_noalloc, the 1D march that allocates nothing, against
sc1d_march (the handle's march) through the C ABI: two calls in a row, the
second on a smaller tube, each equal bit for bit; sizes beyond the scratch
maximum are rejected."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_N = 131071


def _lib():
    from solvcon_amd import _lib as L
    lib = L.lib()
    fn = lib.sc1d_march_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [L.P_dbl, L.P_dbl, L.P_dbl, L.P_dbl, L.P_dbl, L.c_i64,
                   L.c_dbl, L.c_int, L.c_i64, L.c_vp]
    return L, lib


def _dev(L, core, which):
    ptr = L.P_dbl()
    cnt = L.c_i64()
    L.check(L.lib().sc1d_device_ptr(core._h, which, ctypes.byref(ptr),
                                    ctypes.byref(cnt)))
    return ptr


def _tube(n):
    from solvcon_amd.gen import sod_tube
    dt = 0.4 / 600 * 2000 / (n - 1)
    return sod_tube(ncoord=n, time_increment=dt)


def _noalloc_march(L, lib, core, alpha, steps):
    # The tables of a handle serve as the caller's device memory; the
    # handle's own march is not used.  The set-up's uploads ran on the
    # handle's stream: wait for them before the default stream reads.
    L.check(lib.sc1d_sync(core._h))
    args = [_dev(L, core, w) for w in (0, 4, 2, 3, 1)]   # coord gamma so0 so1 cfl
    L.check(lib.sc1d_march_noalloc(*args, core.ncoord, core.time_increment,
                                   alpha, steps, None))
    L.check(lib.sc_device_sync())
    for m in core._mirrors.values():
        m.valid = False


@pytest.mark.parametrize('alpha', [1, 2])
def test_noalloc_equals_handle_march(alpha):
    L, lib = _lib()
    # First call: a large tube that takes several tiled launches (odd and
    # even counts); second: a smaller one, so the shared scratch still
    # holds the first call's values beyond the new tube's end.
    for n, steps in ((MAX_N, 130), (2001, 600), (9001, 65)):
        print('case: noalloc tube', n, 'points', steps, 'steps, alpha',
              alpha, flush=True)
        ref = _tube(n)
        getattr(ref, 'march_alpha%d' % alpha)(steps)
        got = _tube(n)
        _noalloc_march(L, lib, got, alpha, steps)
        assert np.isfinite(ref.so0).all()
        for name in ('so0', 'so1', 'cfl'):
            assert np.array_equal(getattr(got, name), getattr(ref, name)), \
                (n, name)


def test_noalloc_rejects_bad_sizes():
    L, lib = _lib()
    core = _tube(2001)
    L.check(lib.sc1d_sync(core._h))
    args = [_dev(L, core, w) for w in (0, 4, 2, 3, 1)]
    for n, alpha in ((MAX_N + 2, 2), (2000, 2), (3, 2), (2001, 3)):
        rc = lib.sc1d_march_noalloc(*args, n, 1e-3, alpha, 1, None)
        assert rc != 0, n
        with pytest.raises(ValueError):
            L.check(rc)
