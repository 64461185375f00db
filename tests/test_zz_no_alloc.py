"""The no-allocation entry point (slsim_shoot_device_noalloc), through the C
ABI: its results equal slsim_shoot_device's bit for bit on the same device
buffers, on two calls in a row (the second smaller, so that the fixed
scratch array holds the first call's contents), in both walk precisions, with
and without per-ray source redshifts; sizes beyond its maximum are refused."""
import ctypes

import numpy as np
import pytest

import slsimlib_amd as S
from slsimlib_amd import _abi
from slsimlib_amd import workloads as W

pytestmark = pytest.mark.gpu

P = ctypes.POINTER(ctypes.c_double)


def _dev(ctx, a):
    d = ctx.alloc(max(a.nbytes, 8))
    ctx.h2d(d, np.ascontiguousarray(a))
    return d


def _shoot(fn, lens, ctx, x, zs, prec):
    n = x.shape[0]
    xd = _dev(ctx, x)
    zd = _dev(ctx, zs) if zs is not None else None
    yd, Ad = ctx.alloc(16 * n), ctx.alloc(32 * n)
    st = _abi.slsim_shoot_stats()
    try:
        rc = fn(lens.h, ctypes.c_size_t(n), ctypes.cast(xd, P),
                ctypes.cast(zd, P) if zd is not None else None, ctypes.cast(yd, P),
                ctypes.cast(Ad, P), int(prec), ctypes.byref(st))
        if rc != 0:
            raise S.SlsimError(rc, S.lib.slsim_last_error().decode())
        y, A = np.zeros((n, 2)), np.zeros((n, 4))
        ctx.d2h(y, yd)
        ctx.d2h(A, Ad)
    finally:
        for d in (xd, zd, yd, Ad):
            if d is not None:
                ctx.free(d)
    return y, A, st.n_guarded


@pytest.fixture(scope="module")
def lens(ctx):
    lens, _, _ = W.build_field_lens(ctx, config=6, nplanes=5, n=1 << 15, mass_scale=1 / 32)
    return lens


@pytest.mark.parametrize("prec", [S.SLSIM_FP64, S.SLSIM_FP32C])
def test_noalloc_equals_allocating_shoot(ctx, lens, prec):
    rng = np.random.default_rng(31)
    for n, with_z in ((120000, False), (5000, True)):
        x = (rng.random((n, 2)) - 0.5) * W.grid_range(300.0)
        zs = rng.uniform(0.5, 3.5, n) if with_z else None
        y0, A0, g0 = _shoot(S.lib.slsim_shoot_device, lens, ctx, x, zs, prec)
        y1, A1, g1 = _shoot(S.lib.slsim_shoot_device_noalloc, lens, ctx, x, zs, prec)
        assert np.array_equal(y0, y1) and np.array_equal(A0, A1), (n, with_z)
        assert g0 == g1


def test_noalloc_refuses_sizes_beyond_its_maximum(ctx, lens):
    n = 131073
    x = np.zeros((n, 2))
    with pytest.raises(S.SlsimError):
        _shoot(S.lib.slsim_shoot_device_noalloc, lens, ctx, x, None, S.SLSIM_FP64)
