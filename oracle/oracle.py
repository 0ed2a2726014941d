"""ctypes/numpy face of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline, never as a product path.  See oracle.h for what each function
restates from the reference and how it is pinned.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")

_lib = None


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.ps_ref_reduce_float_array.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_int, _f32p, ctypes.c_int64]
        L.ps_ref_reduce_double_array.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int, _f64p, ctypes.c_int64]
        L.ps_uniform_field_f32.argtypes = [_f32p, ctypes.c_int64, ctypes.c_uint64]
        L.ps_uniform_range_f32.argtypes = [_f32p, ctypes.c_int64, ctypes.c_uint64, ctypes.c_float, ctypes.c_float]
        L.ps_ref_diffusion7_f32.argtypes = [_f32p, _f32p, _i32p, _i32p, _f32p, ctypes.c_int, ctypes.c_int]
        L.ps_ref_diffusion7_f64.argtypes = [_f64p, _f64p, _i32p, _i32p, _f64p, ctypes.c_int, ctypes.c_int]
        L.ps_ref_diffusion7_f64_mt.argtypes = [_f64p, _f64p, _i32p, _i32p, _f64p, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int]
        L.ps_ref_himeno_f32.argtypes = [_f32p, _f32p, ctypes.POINTER(ctypes.c_void_p), _i32p, ctypes.c_float,
                                        ctypes.c_int, ctypes.c_void_p]
        L.ps_ref_himeno_f64.argtypes = [_f64p, _f64p, ctypes.POINTER(ctypes.c_void_p), _i32p, ctypes.c_double,
                                        ctypes.c_int, ctypes.c_void_p]
        _lib = L
    return _lib


PS_MAX, PS_MIN, PS_SUM, PS_PROD = 0, 1, 2, 3


def reduce(arr, op):
    """Sequential left fold from arr[0] in the array's own type (ref PSReduce)."""
    a = np.ascontiguousarray(arr).ravel()
    if a.dtype == np.float32:
        out = ctypes.c_float()
        rc = lib().ps_ref_reduce_float_array(ctypes.byref(out), op, a, a.size)
    else:
        a = a.astype(np.float64)
        out = ctypes.c_double()
        rc = lib().ps_ref_reduce_double_array(ctypes.byref(out), op, a, a.size)
    if rc != 0:
        raise ValueError("invalid reduce operator")
    return out.value


def uniform_field(shape, seed, lo=0.0, hi=1.0):
    """Seeded uniform field of float32 values, x fastest; shape = (nx, ny, nz).

    Returned array is indexed [z, y, x] (C order, x contiguous)."""
    n = int(np.prod(shape))
    out = np.empty(n, np.float32)
    if lo == 0.0 and hi == 1.0:
        lib().ps_uniform_field_f32(out, n, seed)
    else:
        lib().ps_uniform_range_f32(out, n, seed, lo, hi)
    return out.reshape(tuple(reversed(shape)))


def _dom6(dom, shape):
    if dom is None:
        dom = (0, shape[0], 0, shape[1], 0, shape[2])
    return np.asarray(dom, np.int32)


def diffusion7(field, coef, periodic, steps, dom=None, dtype=np.float32, threads=1):
    """Runs `steps` maps of 7-point diffusion; returns (g1, g2) after the run.

    field: array [z, y, x]; coef = (ce, cw, cn, cs, ct, cb, cc)."""
    nz, ny, nx = field.shape
    dim = np.asarray([nx, ny, nz], np.int32)
    d = _dom6(dom, (nx, ny, nz))
    g1 = np.ascontiguousarray(field, dtype=dtype).copy()
    g2 = np.zeros_like(g1)
    k = np.asarray(coef, dtype)
    if dtype == np.float32:
        lib().ps_ref_diffusion7_f32(g1.ravel(), g2.ravel(), dim, d, k, int(periodic), steps)
    elif threads > 1:
        lib().ps_ref_diffusion7_f64_mt(g1.ravel(), g2.ravel(), dim, d, k, int(periodic), steps, threads)
    else:
        lib().ps_ref_diffusion7_f64(g1.ravel(), g2.ravel(), dim, d, k, int(periodic), steps)
    return g1, g2


def himeno(p_init, coefs, omega, sweeps, dtype=np.float32, p1_init=None):
    """Runs Himeno sweeps; coefs = 12 arrays (a0,a1,a2,a3,b0,b1,b2,c0,c1,c2,bnd,wrk1).
    p0 starts as p_init, p1 as p1_init (default: p_init too).

    Returns (p0, p1, gosa[sweeps])."""
    nz, ny, nx = p_init.shape
    dim = np.asarray([nx, ny, nz], np.int32)
    p0 = np.ascontiguousarray(p_init, dtype=dtype).copy()
    p1 = p0.copy() if p1_init is None else np.ascontiguousarray(p1_init, dtype=dtype).copy()
    cs = [np.ascontiguousarray(c, dtype=dtype) for c in coefs]
    ptrs = (ctypes.c_void_p * 12)(*[c.ctypes.data for c in cs])
    gosa = np.zeros(max(sweeps, 1), np.float64)
    if dtype == np.float32:
        lib().ps_ref_himeno_f32(p0.ravel(), p1.ravel(), ptrs, dim, omega, sweeps, gosa.ctypes.data)
    else:
        lib().ps_ref_himeno_f64(p0.ravel(), p1.ravel(), ptrs, dim, omega, sweeps, gosa.ctypes.data)
    return p0, p1, gosa[:sweeps]
