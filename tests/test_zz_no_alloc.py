"""The allocation-free gradient entry point (lpk_lws_gradient_noalloc) against
lpk_lws_gradient, through the C ABI: the same kernels over a static device
workspace instead of library-allocated buffers, so the gradient and loss are
bit-identical. Two calls in a row, the second with fewer pairs (the workspace
keeps the first call's rows; the reduction reads only the rows of its own
call), and a batch beyond the workspace is refused before anything runs.
"""
import ctypes as C

import numpy as np
import pytest

from lipika_amd import _lib
from lipika_amd.pipeline import WarmStartNet

pytestmark = pytest.mark.gpu

MAX_PAIRS = 8192      # 128 workspace rows of 64 pairs


def _grad(fn, net, X, Y, n):
    g = np.full(_lib.LWS_PARAMS, np.nan, dtype=np.float32)
    loss = C.c_double(float("nan"))
    rc = fn(net.h, X.ptr, Y.ptr, n, g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss))
    return rc, g, loss.value


def test_noalloc_gradient_matches_allocating_one(backend):
    lib = backend.lib
    net = WarmStartNet(backend, init_seed=9)
    # one row more than the workspace holds, so that even the refused call's
    # pairs would be in bounds
    X, Y, _ = net.generate(MAX_PAIRS + 64, seed=5)
    try:
        for n in (MAX_PAIRS, 1000):
            print(f"gradient over {n} pairs", flush=True)
            rc, ga, la = _grad(lib.lpk_lws_gradient, net, X, Y, n)
            assert rc == 0
            rc, gb, lb = _grad(lib.lpk_lws_gradient_noalloc, net, X, Y, n)
            assert rc == 0, lib.lpk_last_error(backend.h)
            assert np.all(np.isfinite(gb)) and np.isfinite(lb)
            assert np.array_equal(ga.view(np.uint32), gb.view(np.uint32))
            assert la == lb
        rc, _, _ = _grad(lib.lpk_lws_gradient_noalloc, net, X, Y, MAX_PAIRS + 1)
        assert rc == -1
        assert b"static gradient workspace" in lib.lpk_last_error(backend.h)
    finally:
        X.free()
        Y.free()
        net.close()
