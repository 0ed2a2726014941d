"""Python binding of libphysis_rt_hip's C ABI (ctypes, no torch).

Mirrors the reference runtime's user-visible operations for the single-grid
stencil path: PSInit, PSGrid3D{Float,Double}New, PSGridCopyin/Copyout,
PSGridFree, PSDomain3DNew, PSReduce (reference include/physis/physis_common.h:
76-99, include/physis/physis_cuda.h:131-145, runtime/libphysis_rt_cuda.cc:28-191).
The built-in stencil programs (diffusion, Himeno) are exposed as the run
functions physisc-hip emits for them.

The library is required: importing works without it, but every call raises
RuntimeError if libphysis_rt_hip.so is missing (there is no CPU fallback).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libphysis_rt_hip.so")

PS_MAX, PS_MIN, PS_SUM, PS_PROD = 0, 1, 2, 3
STORAGE_NATIVE, STORAGE_BF16, STORAGE_FP16 = 0, 1, 2
CLAMPED, PERIODIC = 0, 1
STATUS = {0: "ok", 1: "invalid argument", 2: "out of memory", 3: "device error", 4: "unsupported"}


class PhysisError(RuntimeError):
    def __init__(self, where, code):
        super().__init__("%s failed: %s (%d)" % (where, STATUS.get(code, "?"), code))
        self.code = code


class Domain(ctypes.Structure):
    """__PSDomain: half-open boxes [min, max) per axis."""
    _fields_ = [("min", ctypes.c_int32 * 3), ("max", ctypes.c_int32 * 3),
                ("local_min", ctypes.c_int32 * 3), ("local_max", ctypes.c_int32 * 3)]


class GridStruct(ctypes.Structure):
    """__PSGrid public prefix (include/physis/physis_hip.h)."""
    _fields_ = [("p0", ctypes.c_void_p), ("dim", ctypes.c_int * 3), ("elm_size", ctypes.c_int),
                ("num_dims", ctypes.c_int), ("num_elms", ctypes.c_int64), ("dev", ctypes.c_void_p)]


_GP = ctypes.POINTER(GridStruct)
_lib = None


def lib():
    """Loads libphysis_rt_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libphysis_rt_hip.so is not built (run __graft_entry__.build() "
                               "or make -C physis_amd/csrc); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        L.PSInit.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.POINTER(ctypes.c_char_p)),
                             ctypes.c_int]
        L.PSFinalize.argtypes = []
        L.PSHipGridNew.argtypes = [ctypes.POINTER(_GP), ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.PSHipGridFree.argtypes = [_GP]
        L.PSHipGridCopyin.argtypes = [_GP, ctypes.c_void_p]
        L.PSHipGridCopyout.argtypes = [_GP, ctypes.c_void_p]
        L.PSHipGridCopyDevice.argtypes = [_GP, _GP]
        L.PSHipGridStorage.argtypes = [_GP]
        L.PSHipReduceGrid.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int, _GP]
        L.__PSReduceGridFloat.argtypes = [ctypes.c_void_p, ctypes.c_int, _GP]
        L.__PSReduceGridDouble.argtypes = [ctypes.c_void_p, ctypes.c_int, _GP]
        L.PSHipSynchronize.argtypes = []
        L.PSHipGridBytesInUse.restype = ctypes.c_int64
        L.PSHipDeviceName.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.PSHipMemInfo.argtypes = [ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
        L.PSHipTimerElapsed.argtypes = [ctypes.POINTER(ctypes.c_float)]
        L.PSDomain3DNew.restype = Domain
        L.PSDomain3DNew.argtypes = [ctypes.c_int32] * 6
        L.PSHipDiffusion7Run.argtypes = [_GP, _GP, ctypes.POINTER(Domain), ctypes.POINTER(ctypes.c_float),
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.PSHipDiffusion7VariantName.restype = ctypes.c_char_p
        L.PSHipDiffusion7VariantName.argtypes = [ctypes.c_int]
        if hasattr(L, "PSHipHimenoRun"):
            L.PSHipHimenoRun.argtypes = [_GP, _GP, ctypes.POINTER(_GP), ctypes.c_float, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L.PSHipStreamCopy.argtypes = [_GP, _GP]
        L.PSHipGridFillPattern.argtypes = [_GP, ctypes.c_int, ctypes.c_float]
        if hasattr(L, "PSHipGridFillUniform"):
            L.PSHipGridFillUniform.argtypes = [_GP, ctypes.c_uint64, ctypes.c_float, ctypes.c_float]
        _lib = L
    return _lib


def check(code, where):
    if code != 0:
        raise PhysisError(where, code)


_inited = False


def init(argv=None):
    """PSInit(&argc, &argv, 3, ...); returns argv with runtime options removed."""
    global _inited
    args = [a.encode() for a in (argv or ["physis"])]
    argc = ctypes.c_int(len(args))
    arr = (ctypes.c_char_p * (len(args) + 1))(*args, None)
    parr = ctypes.cast(arr, ctypes.POINTER(ctypes.c_char_p))
    lib().PSInit(ctypes.byref(argc), ctypes.byref(parr), 3)
    _inited = True
    return [parr[i].decode() for i in range(argc.value)]


def finalize():
    global _inited
    lib().PSFinalize()
    _inited = False


def synchronize():
    check(lib().PSHipSynchronize(), "PSHipSynchronize")


def device_name():
    buf = ctypes.create_string_buffer(256)
    check(lib().PSHipDeviceName(buf, 256), "PSHipDeviceName")
    return buf.value.decode()


def mem_info():
    f, t = ctypes.c_int64(), ctypes.c_int64()
    check(lib().PSHipMemInfo(ctypes.byref(f), ctypes.byref(t)), "PSHipMemInfo")
    return f.value, t.value


def domain3d(minx, maxx, miny, maxy, minz, maxz):
    """PSDomain3DNew(minx, maxx, miny, maxy, minz, maxz)."""
    return lib().PSDomain3DNew(minx, maxx, miny, maxy, minz, maxz)


class Grid:
    """A device-resident grid (__PSGrid*).  shape is (nx, ny, nz), x fastest;
    host arrays exchanged with it are indexed [z, y, x]."""

    def __init__(self, shape, dtype=np.float32, storage=STORAGE_NATIVE):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.storage = storage
        dims = (ctypes.c_int * 3)(*(list(self.shape) + [1] * (3 - len(self.shape))))
        h = _GP()
        check(lib().PSHipGridNew(ctypes.byref(h), self.dtype.itemsize, len(self.shape), dims, storage),
              "PSHipGridNew")
        self.handle = h

    @property
    def host_shape(self):
        return tuple(reversed(self.shape))

    def copyin(self, arr):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        if a.size != int(np.prod(self.shape)):
            raise ValueError("array size does not match grid")
        check(lib().PSHipGridCopyin(self.handle, a.ctypes.data), "PSGridCopyin")

    def copyout(self):
        a = np.empty(self.host_shape, self.dtype)
        check(lib().PSHipGridCopyout(self.handle, a.ctypes.data), "PSGridCopyout")
        return a

    def fill_uniform(self, seed, lo=0.0, hi=1.0):
        """Device-side fill with the seeded uniform field (same values as the
        oracle's ps_uniform_field)."""
        check(lib().PSHipGridFillUniform(self.handle, seed, lo, hi), "PSHipGridFillUniform")

    def reduce(self, op):
        """PSReduce(&v, op, g): result in the grid's element type."""
        # getattr: a literal __name inside a class body would be name-mangled.
        if self.dtype == np.float32:
            v = ctypes.c_float()
            getattr(lib(), "__PSReduceGridFloat")(ctypes.byref(v), op, self.handle)
        else:
            v = ctypes.c_double()
            getattr(lib(), "__PSReduceGridDouble")(ctypes.byref(v), op, self.handle)
        return v.value

    def reduce64(self, op):
        """FP64-accumulated reduction result."""
        v = ctypes.c_double()
        check(lib().PSHipReduceGrid(ctypes.byref(v), op, self.handle), "PSHipReduceGrid")
        return v.value

    def free(self):
        if self.handle:
            check(lib().PSHipGridFree(self.handle), "PSGridFree")
            self.handle = None

    def __del__(self):
        try:
            if self.handle and _lib is not None:
                _lib.PSHipGridFree(self.handle)
        except Exception:
            pass


def stream_copy(dst, src):
    """One plain streaming copy kernel dst <- src (bandwidth yardstick)."""
    check(lib().PSHipStreamCopy(dst.handle, src.handle), "PSHipStreamCopy")


def diffusion7_run(g1, g2, coef, boundary, steps, dom=None, variant=0):
    """PSStencilRun(map(g1->g2), map(g2->g1), steps/2) of 7-point diffusion.
    coef = (ce, cw, cn, cs, ct, cb, cc).  After an even count the result is
    in g1."""
    if dom is None:
        nx, ny, nz = g1.shape
        dom = domain3d(0, nx, 0, ny, 0, nz)
    k = (ctypes.c_float * 7)(*coef)
    check(lib().PSHipDiffusion7Run(g1.handle, g2.handle, ctypes.byref(dom), k, boundary, steps, variant),
          "PSHipDiffusion7Run")


def diffusion7_variants():
    n = lib().PSHipDiffusion7Variants()
    return {i: lib().PSHipDiffusion7VariantName(i).decode() for i in range(1, n + 1)}


def himeno_run(p0, p1, coef_grids, omega, sweeps, with_gosa=True, variant=0):
    """Himeno sweeps p0->p1->p0...; coef_grids = 12 Grids (a0,a1,a2,a3,b0,b1,
    b2,c0,c1,c2,bnd,wrk1).  Returns the FP64 sum of ss^2 per sweep."""
    arr = (_GP * 12)(*[g.handle for g in coef_grids])
    gosa = (ctypes.c_double * max(sweeps, 1))()
    check(lib().PSHipHimenoRun(p0.handle, p1.handle, arr, omega, sweeps,
                               gosa if with_gosa else None, variant), "PSHipHimenoRun")
    return np.array(gosa[:sweeps], np.float64)


def himeno_run_tile(p0, p1, coef_grids, omega, sweeps, bx, by, zchunk, rows=1, with_gosa=True):
    """Himeno sweeps, one per launch, on the generic skeleton with bx x by
    threads of `rows` tile rows each (64 x 2/4/8, 128 x 2/4, 256 x 1/2 with
    one row; 64 x 4, 128 x 2, 256 x 1 with two) and zchunk planes per
    workgroup (0 = all).  Returns the FP64 sum of ss^2 per sweep."""
    L = lib()
    L.PSHipHimenoRunTile.argtypes = [_GP, _GP, ctypes.POINTER(_GP), ctypes.c_float, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_double)] + [ctypes.c_int] * 4
    arr = (_GP * 12)(*[g.handle for g in coef_grids])
    gosa = (ctypes.c_double * max(sweeps, 1))()
    check(L.PSHipHimenoRunTile(p0.handle, p1.handle, arr, omega, sweeps, gosa if with_gosa else None, bx, by,
                               rows, zchunk), "PSHipHimenoRunTile")
    return np.array(gosa[:sweeps], np.float64)


def himeno_run_tb(p0, p1, coef_grids, omega, sweeps, degree, by=16, rows_per_wave=1, zchunk=0, prefetch=0,
                  with_gosa=True):
    """Himeno sweeps with `degree` sweeps fused per HBM pass (tiles of 64 x by
    points, rows_per_wave rows per wave, zchunk output planes per workgroup,
    0 = all); same results and final grid contents as himeno_run.  Returns
    the FP64 sum of ss^2 per sweep."""
    L = lib()
    L.PSHipHimenoRunTB.argtypes = [_GP, _GP, ctypes.POINTER(_GP), ctypes.c_float, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_double)] + [ctypes.c_int] * 5
    arr = (_GP * 12)(*[g.handle for g in coef_grids])
    gosa = (ctypes.c_double * max(sweeps, 1))()
    check(L.PSHipHimenoRunTB(p0.handle, p1.handle, arr, omega, sweeps, gosa if with_gosa else None,
                             degree, by, rows_per_wave, zchunk, prefetch), "PSHipHimenoRunTB")
    return np.array(gosa[:sweeps], np.float64)


class Batch:
    """Many independent 7-point diffusion grid pairs advanced together: one
    launch per step (or per two fused steps) for all grids, captured into a
    hipGraph.  pairs: list of (g1, g2) Grids (nx <= 256); coefs: one 7-tuple
    (ce, cw, cn, cs, ct, cb, cc) per pair."""

    def __init__(self, pairs, coefs, boundary=CLAMPED):
        L = lib()
        for name in ("PSHipBatchCreate", "PSHipBatchRun", "PSHipBatchTune", "PSHipBatchConfigure",
                     "PSHipBatchConfigName", "PSHipBatchDestroy"):
            getattr(L, name).restype = ctypes.c_int
        L.PSHipBatchCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(_GP), ctypes.POINTER(_GP),
                                       ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.c_int]
        L.PSHipBatchRun.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.PSHipBatchTune.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
        L.PSHipBatchConfigure.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 4
        L.PSHipBatchConfigName.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
        L.PSHipBatchDestroy.argtypes = [ctypes.c_void_p]
        n = len(pairs)
        self.pairs = pairs
        a = (_GP * n)(*[p[0].handle for p in pairs])
        b = (_GP * n)(*[p[1].handle for p in pairs])
        k = (ctypes.c_float * (7 * n))(*[float(v) for c in coefs for v in c])
        h = ctypes.c_void_p()
        check(L.PSHipBatchCreate(ctypes.byref(h), a, b, n, k, boundary), "PSHipBatchCreate")
        self.handle = h

    def run(self, steps, graph=True):
        check(lib().PSHipBatchRun(self.handle, steps, 1 if graph else 0), "PSHipBatchRun")

    def configure(self, np_, by, zchunk, degree):
        check(lib().PSHipBatchConfigure(self.handle, np_, by, zchunk, degree), "PSHipBatchConfigure")

    def tune(self, steps):
        ms = ctypes.c_float()
        check(lib().PSHipBatchTune(self.handle, steps, ctypes.byref(ms)), "PSHipBatchTune")
        return ms.value

    def config_name(self):
        buf = ctypes.create_string_buffer(64)
        check(lib().PSHipBatchConfigName(self.handle, buf, 64), "PSHipBatchConfigName")
        return buf.value.decode()

    def destroy(self):
        if self.handle:
            check(lib().PSHipBatchDestroy(self.handle), "PSHipBatchDestroy")
            self.handle = None


class Request(ctypes.Structure):
    """PSHipRequest: kind 0 = 7-point diffusion, 1 = 19-point Himeno-form Jacobi."""
    _fields_ = [("kind", ctypes.c_int), ("n", ctypes.c_int), ("storage", ctypes.c_int), ("seed", ctypes.c_uint64),
                ("arrival", ctypes.c_double), ("max_steps", ctypes.c_int), ("tol", ctypes.c_double)]


class Result(ctypes.Structure):
    """PSHipResult."""
    _fields_ = [("status", ctypes.c_int), ("steps", ctypes.c_int), ("residual", ctypes.c_double),
                ("admitted", ctypes.c_double), ("finished", ctypes.c_double), ("checksum", ctypes.c_double)]


SERVE_BATCHED, SERVE_SEQUENTIAL = 0, 1


def _requests(requests):
    n = len(requests)
    return (Request * n)(*[Request(r["kind"], r["n"], r.get("storage", STORAGE_NATIVE), r.get("seed", 1),
                                   r.get("arrival", 0.0), r.get("max_steps", 2000), r.get("tol", 1e-5))
                           for r in requests])


def serve_tune(requests, budget_bytes, tick_steps=50):
    """Tunes the Jacobi launch shapes of the requests' size classes ahead of
    serving (PSHipServeTune); later traces reuse them."""
    L = lib()
    L.PSHipServeTune.argtypes = [ctypes.POINTER(Request), ctypes.c_int, ctypes.c_int64, ctypes.c_int]
    check(L.PSHipServeTune(_requests(requests), len(requests), budget_bytes, tick_steps), "PSHipServeTune")


def serve_trace(requests, budget_bytes, tick_steps=50, mode=SERVE_BATCHED):
    """Serves a list of request dicts (kind, n, storage, seed, arrival,
    max_steps, tol) in real time; returns one result dict per request."""
    L = lib()
    L.PSHipServeTrace.argtypes = [ctypes.POINTER(Request), ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(Result)]
    n = len(requests)
    reqs = (Request * n)(*[Request(r["kind"], r["n"], r.get("storage", STORAGE_NATIVE), r.get("seed", 1),
                                   r.get("arrival", 0.0), r.get("max_steps", 2000), r.get("tol", 1e-5))
                           for r in requests])
    res = (Result * n)()
    check(L.PSHipServeTrace(reqs, n, budget_bytes, tick_steps, mode, res), "PSHipServeTrace")
    return [{"status": r.status, "steps": r.steps, "residual": r.residual, "admitted": r.admitted,
             "finished": r.finished, "checksum": r.checksum} for r in res]


class Timer:
    """HIP events on the runtime stream."""

    def begin(self):
        check(lib().PSHipTimerBegin(), "PSHipTimerBegin")

    def end(self):
        check(lib().PSHipTimerEnd(), "PSHipTimerEnd")

    def elapsed_ms(self):
        ms = ctypes.c_float()
        check(lib().PSHipTimerElapsed(ctypes.byref(ms)), "PSHipTimerElapsed")
        return ms.value
