"""gx_mlp_gradient_noalloc (scratch in a fixed device array, no device
allocation in the call) against gx_mlp_gradient on the same network and
batch: identical gradient and loss bits, on two calls in a row, the second
with a smaller batch, and with the network's own stream and a caller's."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a):
    from genomix_amd import devmem
    return devmem.DeviceBytes(np.ascontiguousarray(a, np.float32).tobytes())


def _host_gradient(gx, m, x, t, n):
    g = np.zeros(n, np.float32)
    loss = ctypes.c_float()
    fp = ctypes.POINTER(ctypes.c_float)
    rc = gx.lib().gx_mlp_gradient(m, x.ctypes.data_as(fp), t.ctypes.data_as(fp), len(x), g.ctypes.data_as(fp),
                                  ctypes.byref(loss))
    assert rc == 0, gx.lib().gx_mlp_last_error(m)
    return g, loss.value


def _noalloc_gradient(gx, m, x, t, n, stream=None):
    from genomix_amd import devmem
    dx, dt = _dev(x), _dev(t)
    dg, dl = _dev(np.full(n, np.nan, np.float32)), _dev(np.full(1, np.nan, np.float32))
    rc = gx.lib().gx_mlp_gradient_noalloc(m, ctypes.c_void_p(dx.address), ctypes.c_void_p(dt.address), len(x),
                                          ctypes.c_void_p(dg.address), ctypes.c_void_p(dl.address), stream)
    assert rc == 0, gx.lib().gx_mlp_last_error(m)
    devmem.synchronize()
    g = np.frombuffer(devmem.to_host(dg.address, 4 * n), np.float32).copy()
    loss = float(np.frombuffer(devmem.to_host(dl.address, 4), np.float32)[0])
    return g, loss


@pytest.fixture
def net():
    from genomix_amd import gx
    L = gx.lib()
    m = ctypes.c_void_p()
    din, h, dout = 72, 64, 2
    assert L.gx_mlp_create(din, h, dout, 0, ctypes.byref(m)) == 0
    n = L.gx_mlp_num_params(m)
    rng = np.random.default_rng(11)
    p = (rng.standard_normal(n) * 0.1).astype(np.float32)
    assert L.gx_mlp_set_params(m, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n) == 0
    yield gx, m, din, dout, n
    L.gx_mlp_destroy(m)


def test_noalloc_gradient_matches_gradient(net):
    gx, m, din, dout, n = net
    rng = np.random.default_rng(12)
    for batch in (3000, 257):  # the second call smaller: the array keeps the first call's contents
        x = rng.standard_normal((batch, din)).astype(np.float32)
        t = rng.standard_normal((batch, dout)).astype(np.float32)
        g0, l0 = _host_gradient(gx, m, x, t, n)
        g1, l1 = _noalloc_gradient(gx, m, x, t, n)
        assert np.array_equal(g0.view(np.uint32), g1.view(np.uint32)), batch
        assert np.float32(l0).view(np.uint32) == np.float32(l1).view(np.uint32), batch


def test_noalloc_gradient_on_a_caller_stream(net):
    gx, m, din, dout, n = net
    from genomix_amd import devmem
    hip = devmem.hip()
    st = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(st)) == 0
    try:
        rng = np.random.default_rng(13)
        x = rng.standard_normal((1000, din)).astype(np.float32)
        t = rng.standard_normal((1000, dout)).astype(np.float32)
        g0, l0 = _host_gradient(gx, m, x, t, n)
        g1, l1 = _noalloc_gradient(gx, m, x, t, n, stream=st)
        assert np.array_equal(g0.view(np.uint32), g1.view(np.uint32))
        assert np.float32(l0) == np.float32(l1)
    finally:
        hip.hipStreamSynchronize(st)
        hip.hipStreamDestroy(st)


def test_noalloc_gradient_limits(net):
    gx, m, din, dout, n = net
    x = _dev(np.zeros((4097, din), np.float32))
    t = _dev(np.zeros((4097, dout), np.float32))
    L = gx.lib()
    assert L.gx_mlp_gradient_noalloc(m, ctypes.c_void_p(x.address), ctypes.c_void_p(t.address), 4097, None, None,
                                     None) == gx.GX_EINVAL
    assert b"4096" in L.gx_mlp_last_error(m)
    assert L.gx_mlp_gradient_noalloc(m, ctypes.c_void_p(x.address), ctypes.c_void_p(t.address), 0, None, None,
                                     None) == gx.GX_EINVAL
