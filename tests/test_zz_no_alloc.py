"""hexo_price_noalloc through the C ABI: the same prices and standard errors as hexo_price, bit for bit
(the same kernels and reduction order, over a fixed device region instead of an allocated one), on
two calls in a row, the second with a smaller request; a request beyond the fixed region is refused
with HEXO_ERR_INVALID and a message naming the limit."""
import ctypes as C

import numpy as np
import pytest

import hestonexotics_amd as H
from hestonexotics_amd import workloads as W

pytestmark = pytest.mark.gpu


def noalloc(ctx, req):
    n = req.n_outputs
    p = np.full(n, np.nan)
    s = np.full(n, np.nan)
    cr = req.to_c()
    rc = ctx.lib.hexo_price_noalloc(ctx.h, C.byref(cr), p.ctypes.data_as(C.POINTER(C.c_double)),
                                    s.ctypes.data_as(C.POINTER(C.c_double)))
    return rc, p, s


def test_noalloc_equals_hexo_price_on_two_calls(ctx):
    first = W.c2(n_paths=300_000, precision=H.FP32)[3]  # 64 strikes, Asian + European, 52 steps
    second = W.c1(n_paths=20_000, precision=H.FP64, payoffs=H.ASIAN | H.EUROPEAN)  # smaller
    for r in (first, second):
        print("gpu noalloc:", r.n_paths, "paths,", r.n_outputs, "outputs, precision", r.precision, flush=True)
        rc, p, s = noalloc(ctx, r)
        assert rc == 0, ctx.lib.hexo_ctx_last_error(ctx.h)
        p_ref, s_ref = ctx.price(r)
        np.testing.assert_array_equal(p, p_ref)
        np.testing.assert_array_equal(s, s_ref)
    # the martingale-corrected kernel and the control variate (exact prices in the fixed region)
    r = W.c1(n_paths=50_000, precision=H.FP32, payoffs=H.ASIAN | H.ASIAN_GEO_CV | H.GEO_ASIAN)
    r.flags |= H.MARTINGALE | H.CV_REGRESSION
    rc, p, s = noalloc(ctx, r)
    assert rc == 0, ctx.lib.hexo_ctx_last_error(ctx.h)
    p_ref, s_ref = ctx.price(r)
    np.testing.assert_array_equal(p, p_ref)
    np.testing.assert_array_equal(s, s_ref)


def test_noalloc_refuses_beyond_its_region(ctx):
    # 2^30 paths in FP64: one tile sum per 64 paths, 2^24 x 16 bytes of partials > 64 MiB
    big = W.c1(n_paths=2 ** 30, precision=H.FP64)
    rc, p, _ = noalloc(ctx, big)
    assert rc == -1
    msg = ctx.lib.hexo_ctx_last_error(ctx.h).decode()
    assert "hexo_price_noalloc" in msg and str(64 << 20) in msg, msg
    # a later call still works
    r = W.c1(n_paths=10_000, precision=H.FP32)
    rc, p, s = noalloc(ctx, r)
    assert rc == 0
    np.testing.assert_array_equal(p, ctx.price(r)[0])
