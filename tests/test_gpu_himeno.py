"""Himeno 19-point Jacobi on the GPU vs the oracle.

FP32 pressure fields are bitwise equal to the sequential ref restatement
(same expression order, no FMA contraction).  The FP64 residual sums differ
from the oracle's sequential sum only by summation order (relative 1e-12).
16-bit storage is compared with the FP64 oracle within a stated tolerance.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def himeno_inputs(shape, seed):
    """Inputs of the 257x129x129 Himeno workload: coefficients uniform in [0.9, 1.1] (a3 in [0.15,
    0.18]), bnd = 1, wrk1 = 0, p = (z/(nz-1))^2."""
    nx, ny, nz = shape
    coefs = []
    for i in range(12):
        if i == 3:
            c = O.uniform_field(shape, seed=seed + i, lo=0.15, hi=0.18)
        elif i == 10:
            c = np.ones(tuple(reversed(shape)), np.float32)
        elif i == 11:
            c = np.zeros(tuple(reversed(shape)), np.float32)
        else:
            c = O.uniform_field(shape, seed=seed + i, lo=0.9, hi=1.1)
        coefs.append(c)
    z = np.arange(nz, dtype=np.float32)
    p = np.broadcast_to((z * z / np.float32((nz - 1) * (nz - 1)))[:, None, None], (nz, ny, nx)).astype(np.float32)
    return p, coefs


def run_hip(rt, p, coefs, omega, sweeps, storage=0, variant=0):
    nz, ny, nx = p.shape
    shape = (nx, ny, nz)
    p0, p1 = rt.Grid(shape, storage=storage), rt.Grid(shape, storage=storage)
    p0.copyin(p)
    p1.copyin(p)
    grids = []
    for c in coefs:
        g = rt.Grid(shape, storage=storage)
        g.copyin(c)
        grids.append(g)
    gosa = rt.himeno_run(p0, p1, grids, omega, sweeps, variant=variant)
    out0, out1 = p0.copyout(), p1.copyout()
    for g in grids + [p0, p1]:
        g.free()
    return out0, out1, gosa


# Default tile widths: 64 (65, 19, 3 wide), 128 (129 wide: 127 interior
# points) and 256 (258 wide: 256 interior points).
@pytest.mark.parametrize("shape", [(65, 33, 33), (19, 7, 5), (3, 3, 3), (129, 17, 9), (258, 6, 10)])
def test_himeno_bitwise_vs_oracle(rt, shape):
    p, coefs = himeno_inputs(shape, seed=100)
    want0, want1, wgosa = O.himeno(p, coefs, 0.8, 6)
    for variant in range(8):
        print("shape", shape, "variant", variant, flush=True)
        got0, got1, gosa = run_hip(rt, p, coefs, 0.8, 6, variant=variant)
        assert np.array_equal(got0, want0) and np.array_equal(got1, want1)
        np.testing.assert_allclose(gosa, wgosa, rtol=1e-12, atol=1e-300)


def test_himeno_config3_shape_400_sweeps_vs_fp64(rt):
    shape = (257, 129, 129)
    p, coefs = himeno_inputs(shape, seed=200)
    got0, _, gosa = run_hip(rt, p, coefs, 0.8, 40)
    want0, _, wgosa = O.himeno(p.astype(np.float64), [c.astype(np.float64) for c in coefs], 0.8, 40,
                               dtype=np.float64)
    rel = np.max(np.abs(got0 - want0)) / np.max(np.abs(want0))
    print("FP32 storage: max rel error vs FP64 after 40 sweeps", rel, flush=True)
    assert rel < 1e-5
    np.testing.assert_allclose(gosa, wgosa, rtol=1e-4)


# Unit roundoff of 16-bit storage (BF16: 8 significant bits, FP16: 11).  The
# workload's iteration diverges (|p| ~ 3e3 after 40 sweeps), and a rounding
# error grows with the field, so the relative error accumulates about one
# rounding per sweep: bound 2 * sweeps * u for the field (max error / peak)
# and twice that for the residual sums of squares.  Measured at 40 sweeps:
# BF16 7.1e-2 (bound 0.156).
UNIT_ROUNDOFF16 = {1: 2.0 ** -9, 2: 2.0 ** -12}


@pytest.mark.parametrize("storage", [1, 2])
def test_himeno_config3_shape_16bit_storage_vs_fp64(rt, storage):
    """The 257x129x129 workload with 16-bit grid storage (FP32 arithmetic):
    error of the one-sweep-per-launch path against the FP64 oracle within
    2 * sweeps * u after 40 sweeps (400 sweeps diverge in FP64 too:
    |p| ~ 1e63), and the fused-sweep path (two and three sweeps per pass)
    equal to it bitwise."""
    shape = (257, 129, 129)
    p, coefs = himeno_inputs(shape, seed=200)
    sweeps = 40
    got0, got1, gosa = run_hip(rt, p, coefs, 0.8, sweeps, storage=storage)
    want0, _, wgosa = O.himeno(p.astype(np.float64), [c.astype(np.float64) for c in coefs], 0.8, sweeps,
                               dtype=np.float64)
    rel = np.max(np.abs(got0 - want0)) / np.max(np.abs(want0))
    print("storage", storage, "257x129x129, %d sweeps: max error / peak vs FP64" % sweeps, rel, flush=True)
    assert rel < 2 * sweeps * UNIT_ROUNDOFF16[storage]
    np.testing.assert_allclose(gosa, wgosa, rtol=4 * sweeps * UNIT_ROUNDOFF16[storage])
    for degree, by in ((2, 8), (3, 16)):
        print("fused degree", degree, flush=True)
        g0, g1 = rt.Grid(shape, storage=storage), rt.Grid(shape, storage=storage)
        g0.copyin(p)
        g1.copyin(p)
        grids = []
        for c in coefs:
            g = rt.Grid(shape, storage=storage)
            g.copyin(c)
            grids.append(g)
        fgosa = rt.himeno_run_tb(g0, g1, grids, 0.8, sweeps, degree, by, 1, 32)
        assert np.array_equal(g0.copyout(), got0) and np.array_equal(g1.copyout(), got1)
        np.testing.assert_allclose(fgosa, gosa, rtol=1e-12)
        for g in grids + [g0, g1]:
            g.free()


def test_himeno_convergent_form_400_sweeps_vs_fp64(rt):
    """With the workload's coefficient ranges (b-terms ~1) the relaxation
    diverges (|p| ~ 3e3 after 40 sweeps, ~1e63 after 400 in FP64), so 400
    sweeps are checked on the original Himeno coefficients (a=c=1, b=0,
    a3=1/6), which converge."""
    shape = (65, 33, 33)
    nx, ny, nz = shape
    vals = [1.0, 1.0, 1.0, 1.0 / 6.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0, 0.0]
    coefs = [np.full((nz, ny, nx), np.float32(v), np.float32) for v in vals]
    z = np.arange(nz, dtype=np.float32)
    p = np.broadcast_to((z * z / np.float32((nz - 1) * (nz - 1)))[:, None, None], (nz, ny, nx)).astype(np.float32)
    got0, _, gosa = run_hip(rt, p, coefs, 0.8, 400)
    want0, _, wgosa = O.himeno(p.astype(np.float64), [c.astype(np.float64) for c in coefs], 0.8, 400,
                               dtype=np.float64)
    rel = np.max(np.abs(got0 - want0)) / np.max(np.abs(want0))
    print("FP32 storage, 400 sweeps: max rel error vs FP64", rel, flush=True)
    assert rel < 1e-5
    assert gosa[-1] < gosa[0]
    np.testing.assert_allclose(gosa, wgosa, rtol=1e-3)


@pytest.mark.parametrize("storage,tol", [(1, 3e-2), (2, 5e-3)])
def test_himeno_16bit_storage_vs_fp64(rt, storage, tol):
    shape = (65, 33, 33)
    p, coefs = himeno_inputs(shape, seed=300)
    got0, _, gosa = run_hip(rt, p, coefs, 0.8, 10, storage=storage)
    want0, _, _ = O.himeno(p.astype(np.float64), [c.astype(np.float64) for c in coefs], 0.8, 10,
                           dtype=np.float64)
    err = np.max(np.abs(got0 - want0))
    print("storage", storage, "max abs error vs FP64", err, flush=True)
    assert err < tol


# Fused-sweep (temporal blocking) configurations: (degree, tile rows, rows
# per wave, z chunk).
TB_CONFIGS = [(1, 8, 1, 0), (1, 16, 2, 5), (2, 8, 1, 0), (2, 8, 2, 7), (2, 16, 2, 3), (2, 16, 1, 0), (3, 8, 1, 4),
              (3, 8, 2, 0), (3, 16, 1, 0), (3, 16, 2, 6), (4, 16, 1, 0), (4, 16, 2, 2)]


@pytest.mark.parametrize("shape", [(65, 33, 33), (19, 7, 5), (3, 3, 3), (129, 17, 9), (70, 40, 23), (4, 30, 12)])
def test_himeno_fused_sweeps_bitwise_vs_oracle(rt, shape):
    """T sweeps per pass leave p0, p1 and every sweep's residual exactly as
    T single sweeps do.  p1 starts different from p0 (its boundary shell
    too), so a pass that took the shell of the wrong iterate would show."""
    p, coefs = himeno_inputs(shape, seed=400)
    nz, ny, nx = p.shape
    p1_init = O.uniform_field(shape, seed=77, lo=-1.0, hi=1.0)
    for sweeps in (1, 2, 5, 6, 11):
        want0, want1, wgosa = O.himeno(p, coefs, 0.8, sweeps, p1_init=p1_init)
        for degree, by, rpw, zchunk in TB_CONFIGS:
            print("shape", shape, "sweeps", sweeps, "config", (degree, by, rpw, zchunk), flush=True)
            g0, g1 = rt.Grid(shape), rt.Grid(shape)
            g0.copyin(p)
            g1.copyin(p1_init)
            grids = []
            for c in coefs:
                g = rt.Grid(shape)
                g.copyin(c)
                grids.append(g)
            gosa = rt.himeno_run_tb(g0, g1, grids, 0.8, sweeps, degree, by, rpw, zchunk)
            got0, got1 = g0.copyout(), g1.copyout()
            for g in grids + [g0, g1]:
                g.free()
            assert np.array_equal(got0, want0), "p0 differs"
            assert np.array_equal(got1, want1), "p1 differs"
            np.testing.assert_allclose(gosa, wgosa, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("storage", [1, 2])
def test_himeno_fused_sweeps_16bit_equal_single_sweeps(rt, storage):
    """With 16-bit storage every level is rounded to storage, so fused and
    single sweeps agree bitwise too."""
    shape = (97, 45, 21)
    p, coefs = himeno_inputs(shape, seed=500)
    results = []
    for degree, by, rpw, zchunk in [(1, 16, 1, 0), (2, 16, 2, 0), (3, 8, 1, 5), (4, 16, 1, 0), (4, 16, 2, 9)]:
        print("storage", storage, "config", (degree, by, rpw, zchunk), flush=True)
        g0, g1 = rt.Grid(shape, storage=storage), rt.Grid(shape, storage=storage)
        g0.copyin(p)
        g1.copyin(p)
        grids = []
        for c in coefs:
            g = rt.Grid(shape, storage=storage)
            g.copyin(c)
            grids.append(g)
        gosa = rt.himeno_run_tb(g0, g1, grids, 0.8, 14, degree, by, rpw, zchunk)
        results.append((g0.copyout(), g1.copyout(), gosa))
        for g in grids + [g0, g1]:
            g.free()
    ref = run_hip(rt, p, coefs, 0.8, 14, storage=storage)
    for got in results:
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
        np.testing.assert_allclose(got[2], ref[2], rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("cfg", [(1, 8, 1, 16, 2), (1, 8, 1, 0, 3), (1, 8, 1, 5, 4), (1, 4, 1, 16, 2),
                                 (2, 8, 1, 32, 2), (2, 8, 1, 3, 3), (3, 16, 1, 0, 2)])
def test_himeno_prefetch_distances_bitwise_vs_oracle(rt, cfg):
    """Deeper coefficient prefetch only changes when loads are issued."""
    shape = (70, 40, 23)
    p, coefs = himeno_inputs(shape, seed=600)
    want0, want1, wgosa = O.himeno(p, coefs, 0.8, 9)
    g0, g1 = rt.Grid(shape), rt.Grid(shape)
    g0.copyin(p)
    g1.copyin(p)
    grids = []
    for c in coefs:
        g = rt.Grid(shape)
        g.copyin(c)
        grids.append(g)
    print("config", cfg, flush=True)
    gosa = rt.himeno_run_tb(g0, g1, grids, 0.8, 9, *cfg)
    got0, got1 = g0.copyout(), g1.copyout()
    for g in grids + [g0, g1]:
        g.free()
    assert np.array_equal(got0, want0) and np.array_equal(got1, want1)
    np.testing.assert_allclose(gosa, wgosa, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("tile", [(64, 2, 1), (64, 2, 8), (64, 4, 3), (64, 4, 8), (64, 4, 0), (64, 8, 2), (64, 8, 16),
                                  (128, 2, 4), (128, 4, 5), (256, 1, 8), (256, 2, 0),
                                  (64, 4, 5, 2), (128, 2, 3, 2), (256, 1, 0, 2)])
@pytest.mark.parametrize("storage", [0, 1])
def test_himeno_tile_shapes_vs_oracle(rt, tile, storage):
    """Every tile height and z-chunk of the generic skeleton gives the same
    fields: bitwise for FP32, equal to variant 1 for BF16 storage."""
    shape = (70, 21, 37)
    bx, by, zchunk = tile[:3]
    rows = tile[3] if len(tile) > 3 else 1
    p, coefs = himeno_inputs(shape, seed=700)
    g0, g1 = rt.Grid(shape, storage=storage), rt.Grid(shape, storage=storage)
    grids = []
    for c in coefs:
        g = rt.Grid(shape, storage=storage)
        g.copyin(c)
        grids.append(g)
    outs = []
    for mode in ("tile", "variant1"):
        g0.copyin(p)
        g1.copyin(p)
        print("tile", tile, "storage", storage, mode, flush=True)
        if mode == "tile":
            gosa = rt.himeno_run_tile(g0, g1, grids, 0.8, 7, bx, by, zchunk, rows=rows)
        else:
            gosa = rt.himeno_run(g0, g1, grids, 0.8, 7, variant=1)
        outs.append((g0.copyout(), g1.copyout(), gosa))
    for g in grids + [g0, g1]:
        g.free()
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    np.testing.assert_allclose(outs[0][2], outs[1][2], rtol=1e-12, atol=1e-300)
    if storage == 0:
        want0, want1, wgosa = O.himeno(p, coefs, 0.8, 7)
        assert np.array_equal(outs[0][0], want0) and np.array_equal(outs[0][1], want1)
        np.testing.assert_allclose(outs[0][2], wgosa, rtol=1e-12, atol=1e-300)
