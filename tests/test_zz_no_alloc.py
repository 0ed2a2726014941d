"""PSHipReduceGridNoAlloc (the reduction without allocation) on the GPU,
through the C ABI: its result equals PSHipReduceGrid's bit for bit, for every
operator and storage kind, on two calls in a row, the second on a smaller
grid (stale partials of the first call must not leak into the second)."""
import ctypes
import struct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OPS = ["PS_MAX", "PS_MIN", "PS_SUM", "PS_PROD"]


def bits(x):
    return struct.pack("<d", x)


@pytest.mark.parametrize("dtype,storage", [(np.float32, 0), (np.float64, 0), (np.float32, 1), (np.float32, 2)])
def test_no_alloc_reduce_equals_allocating_reduce(rt, dtype, storage):
    L = rt.lib()
    L.PSHipReduceGridNoAlloc.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int, rt._GP]
    L.PSHipReduceGridNoAlloc.restype = ctypes.c_int
    rng = np.random.default_rng(5)
    for shape in [(300, 211, 97), (37, 5, 3)]:
        print("shape", shape, "dtype", np.dtype(dtype).name, "storage", storage, flush=True)
        g = rt.Grid(shape, dtype=dtype, storage=storage)
        n = int(np.prod(shape))
        # PROD stays finite: values near 1.
        a = (1.0 + (rng.random(n) - 0.5) * 1e-6).astype(dtype)
        a[rng.integers(0, n, 7)] = rng.random(7) * 3.0
        g.copyin(a.reshape(tuple(reversed(shape))))
        for name in OPS:
            op = getattr(rt, name)
            want = ctypes.c_double()
            rt.check(L.PSHipReduceGrid(ctypes.byref(want), op, g.handle), "PSHipReduceGrid")
            got = ctypes.c_double()
            rt.check(L.PSHipReduceGridNoAlloc(ctypes.byref(got), op, g.handle), "PSHipReduceGridNoAlloc")
            assert bits(got.value) == bits(want.value), (name, got.value, want.value)
        g.free()


def test_no_alloc_reduce_rejects_bad_arguments(rt):
    L = rt.lib()
    L.PSHipReduceGridNoAlloc.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int, rt._GP]
    g = rt.Grid((8, 8, 8))
    out = ctypes.c_double()
    assert L.PSHipReduceGridNoAlloc(ctypes.byref(out), 9, g.handle) != 0
    assert L.PSHipReduceGridNoAlloc(None, rt.PS_SUM, g.handle) != 0
    g.free()
