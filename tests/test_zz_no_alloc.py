"""The allocation-free ragged-product entry point (dmx_ring_mul_batch_device_noalloc) through the C
ABI: its products equal those of the allocating entry point it mirrors (dmx_ring_mul_batch_device,
persistent mode) bit for bit, on two calls in a row, the second with a smaller batch (the shared
scratch keeps the first call's contents), and the oracle's on a sample; sizes past its limits
are refused before any launch."""
import ctypes

import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

dm = pytest.importorskip("damascus000_amd")


def _batch(rng, logs):
    a = [O.random_zq(rng, 1 << int(l)) for l in logs]
    b = [O.random_zq(rng, 1 << int(l)) for l in logs]
    return a, b


def _run(ctx, a, b, logs, noalloc):
    A = np.ascontiguousarray(np.concatenate(a), dtype=np.uint64)
    B = np.ascontiguousarray(np.concatenate(b), dtype=np.uint64)
    da, db = ctx.upload(A.tobytes()), ctx.upload(B.tobytes())
    do = dm.DeviceBuffer(ctx, A.nbytes)
    lg = np.ascontiguousarray(np.asarray(logs, dtype=np.uint32))
    if noalloc:
        rc = dm.lib().dmx_ring_mul_batch_device_noalloc(ctx.h, ctypes.c_void_p(da.ptr), ctypes.c_void_p(db.ptr),
                                                         ctypes.c_void_p(do.ptr), lg.ctypes.data_as(ctypes.c_void_p),
                                                         len(lg))
        dm._check(rc)
    else:
        dm.ring_mul_batch_device(da.ptr, db.ptr, do.ptr, lg, mode=0, ctx=ctx)
    return np.frombuffer(do.read(), dtype=np.uint64).reshape(-1, 2)


def test_noalloc_products_equal_the_allocating_entry_point():
    ctx = dm.Context.default()
    rng = np.random.default_rng(7007)
    first = [int(x) for x in rng.integers(0, 13, size=300)] + [12, 0]
    second = [int(x) for x in rng.integers(0, 10, size=40)]
    for logs in (first, second):
        print("noalloc batch", len(logs), "products, max degree", 1 << max(logs), flush=True)
        a, b = _batch(rng, logs)
        want = _run(ctx, a, b, logs, noalloc=False)
        got = _run(ctx, a, b, logs, noalloc=True)
        assert np.array_equal(got, want)
        pos = 0
        for x, y in zip(a, b):
            n = x.shape[0]
            if n <= 256:
                assert np.array_equal(got[pos: pos + n], O.negacyclic_multiply(x, y)), n
            pos += n


def test_noalloc_refuses_sizes_past_its_limits():
    ctx = dm.Context.default()
    buf = dm.DeviceBuffer(ctx, 1 << 20)
    fn = dm.lib().dmx_ring_mul_batch_device_noalloc
    many = np.zeros(65537, dtype=np.uint32)
    rc = fn(ctx.h, ctypes.c_void_p(buf.ptr), ctypes.c_void_p(buf.ptr), ctypes.c_void_p(buf.ptr),
            many.ctypes.data_as(ctypes.c_void_p), len(many))
    assert rc == dm.ERR_CODE["ERR_ARG"]
    deep = np.array([13], dtype=np.uint32)
    rc = fn(ctx.h, ctypes.c_void_p(buf.ptr), ctypes.c_void_p(buf.ptr), ctypes.c_void_p(buf.ptr),
            deep.ctypes.data_as(ctypes.c_void_p), 1)
    assert rc == dm.ERR_CODE["ERR_SHAPE"]
