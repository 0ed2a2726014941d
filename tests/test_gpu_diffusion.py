"""7-point diffusion on the GPU through the C ABI, checked against the oracle.

FP32 results must be bitwise equal to the sequential ref restatement (both
evaluate the DSL expression left to right with no FMA contraction).  At the
full 512^3 size the checks are a bitwise comparison over 2 steps and mass
conservation over many steps.  16-bit storage is checked against the FP64
oracle within a stated tolerance.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

K_BENCH = [0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.4]          # ce, cw, cn, cs, ct, cb, cc
K_ASYM = [0.11, 0.07, 0.05, 0.13, 0.02, 0.17, 0.45]


def run_hip(rt, field, coef, boundary, steps, variant=0, dom=None, storage=0):
    nz, ny, nx = field.shape
    g1 = rt.Grid((nx, ny, nz), storage=storage)
    g2 = rt.Grid((nx, ny, nz), storage=storage)
    g1.copyin(field)
    g2.copyin(np.zeros_like(field))
    d = rt.domain3d(*dom) if dom is not None else None
    rt.diffusion7_run(g1, g2, coef, boundary, steps, dom=d, variant=variant)
    a, b = g1.copyout(), g2.copyout()
    g1.free()
    g2.free()
    return a, b


def variant_ids(rt):
    """0 (the default: fused steps with a shape chosen from the grid size)
    and every fixed kernel variant."""
    return [0] + list(rt.diffusion7_variants().keys())


@pytest.mark.parametrize("boundary", [0, 1])
def test_64cubed_100_steps_bitwise_all_variants(rt, boundary):
    field = O.uniform_field((64, 64, 64), seed=1)
    want1, want2 = O.diffusion7(field, K_BENCH, periodic=bool(boundary), steps=100)
    for v in variant_ids(rt):
        print("variant", v, rt.diffusion7_variants().get(v, "default"), flush=True)
        got1, got2 = run_hip(rt, field, K_BENCH, boundary, 100, variant=v)
        assert np.array_equal(got1, want1), "variant %d" % v
        assert np.array_equal(got2, want2), "variant %d" % v


# (516, 5, 9) and (1024, 3, 4): whole 16-byte rows wider than 512 points,
# which the default variant fuses with four waves per row.
@pytest.mark.parametrize("shape", [(37, 23, 19), (65, 5, 3), (1, 1, 1), (2, 3, 2), (130, 9, 70),
                                   (36, 23, 19), (4, 5, 3), (260, 7, 9), (512, 3, 5), (8, 1, 1),
                                   (516, 5, 9), (1024, 3, 4)])
@pytest.mark.parametrize("boundary", [0, 1])
def test_ragged_shapes_bitwise(rt, shape, boundary):
    field = O.uniform_field(shape, seed=11)
    want1, want2 = O.diffusion7(field, K_ASYM, periodic=bool(boundary), steps=7)
    for v in variant_ids(rt):
        got1, got2 = run_hip(rt, field, K_ASYM, boundary, 7, variant=v)
        assert np.array_equal(got1, want1) and np.array_equal(got2, want2), "variant %d" % v


def test_subdomain_bitwise(rt):
    shape = (40, 33, 21)
    dom = (3, 37, 1, 30, 2, 20)
    field = O.uniform_field(shape, seed=4)
    want1, want2 = O.diffusion7(field, K_ASYM, periodic=False, steps=4, dom=dom)
    for v in variant_ids(rt):
        got1, got2 = run_hip(rt, field, K_ASYM, 0, 4, variant=v, dom=dom)
        assert np.array_equal(got1, want1) and np.array_equal(got2, want2), "variant %d" % v


def test_empty_domain_is_a_no_op(rt):
    field = O.uniform_field((8, 8, 8), seed=4)
    got1, got2 = run_hip(rt, field, K_BENCH, 0, 3, dom=(4, 4, 0, 8, 0, 8))
    assert np.array_equal(got1, field) and np.all(got2 == 0)


def test_neumann_cond_fixture_on_gpu(rt):
    """The reference test test_7-pt-neumann-cond (N=32, 10 steps) on the GPU."""
    from tests.test_oracle import GOLDEN, neumann_coefficients, sha
    n = 32
    field = np.arange(n ** 3, dtype=np.float32).reshape(n, n, n)
    got1, _ = run_hip(rt, field, neumann_coefficients(n), 0, 10)
    assert sha(got1) == GOLDEN["test_7-pt-neumann-cond"]["sha256"]


def fused_variants(rt):
    return [v for v, name in rt.diffusion7_variants().items() if name.startswith("tb2")]


@pytest.mark.parametrize("boundary", [0, 1])
@pytest.mark.parametrize("shape", [(64, 48, 40), (37, 29, 23), (130, 17, 19), (383, 6, 9), (512, 9, 7)])
def test_every_step_count_leaves_both_grids_as_the_reference(rt, shape, boundary):
    """The default run fuses every pair of steps, the last pass also storing
    its intermediate step, along a path of three buffers that depends on the
    parity of the pair count; odd counts run one step first.  For each count
    from 0 to 9 and 100/101, on rows of 64, 37, 130 and 383 points
    (packed-row tiles: fused pairs and single steps) and of 512 points
    (vectorised tiles), both grids equal the
    oracle bitwise: g1 and g2 hold the last two time levels exactly as the
    reference's ping-pong leaves them."""
    field = O.uniform_field(shape, seed=11)
    for steps in list(range(10)) + [100, 101]:
        print("shape", shape, "boundary", boundary, "steps", steps, flush=True)
        want1, want2 = O.diffusion7(field, K_ASYM, periodic=bool(boundary), steps=steps)
        got1, got2 = run_hip(rt, field, K_ASYM, boundary, steps)
        if steps == 0:
            want2 = np.zeros_like(field)
        assert np.array_equal(got1, want1), "g1, %d steps" % steps
        assert np.array_equal(got2, want2), "g2, %d steps" % steps


@pytest.mark.parametrize("boundary", [0, 1])
@pytest.mark.parametrize("shape,steps", [((257, 193, 97), 12), ((383, 64, 130), 13), ((300, 120, 70), 10),
                                         ((510, 40, 33), 9)])
def test_large_packed_grids_bitwise(rt, shape, steps, boundary):
    """The packed-row tiles at sizes with many row tiles and z chunks per
    launch, both workgroup widths (8 waves up to 256 points, 16 above),
    ragged and aligned rows, even and odd step counts: both grids equal the
    oracle bitwise."""
    field = O.uniform_field(shape, seed=5 + shape[0])
    print("shape", shape, "steps", steps, "boundary", boundary, flush=True)
    want1, want2 = O.diffusion7(field, K_ASYM, periodic=bool(boundary), steps=steps)
    got1, got2 = run_hip(rt, field, K_ASYM, boundary, steps)
    assert np.array_equal(got1, want1) and np.array_equal(got2, want2)


def test_512cubed_six_steps_bitwise(rt):
    """Full size: the default kernel and the fused two-step kernels, 6 steps
    (three fused pairs, the last also storing its intermediate step)."""
    n = 512
    g1 = rt.Grid((n, n, n))
    g2 = rt.Grid((n, n, n))
    g1.fill_uniform(7)
    field = O.uniform_field((n, n, n), seed=7)
    assert np.array_equal(g1.copyout(), field)
    variants = [0] + fused_variants(rt)[:2]
    for boundary in (0, 1):
        want1, want2 = O.diffusion7(field, K_BENCH, periodic=bool(boundary), steps=6)
        for v in variants:
            print("512^3 bitwise, 6 steps, boundary", boundary, "variant", v, flush=True)
            g1.copyin(field)
            rt.diffusion7_run(g1, g2, K_BENCH, boundary, 6, variant=v)
            assert np.array_equal(g1.copyout(), want1), "boundary %d variant %d" % (boundary, v)
            assert np.array_equal(g2.copyout(), want2), "boundary %d variant %d" % (boundary, v)
    g1.free()
    g2.free()


@pytest.mark.parametrize("boundary", [0, 1])
def test_512cubed_mass_conservation(rt, boundary):
    n = 512
    g1 = rt.Grid((n, n, n))
    g2 = rt.Grid((n, n, n))
    g1.fill_uniform(3)
    before = g1.reduce64(rt.PS_SUM)
    rt.diffusion7_run(g1, g2, K_BENCH, boundary, 200)
    after = g1.reduce64(rt.PS_SUM)
    # FP32 rounding per update; the operator itself conserves mass exactly.
    assert abs(after - before) / before < 1e-5
    assert abs(before / n ** 3 - 0.5) < 1e-3
    g1.free()
    g2.free()


@pytest.mark.parametrize("storage,tol", [(1, 2e-2), (2, 2e-3)])
def test_16bit_storage_vs_fp64(rt, storage, tol):
    shape = (48, 40, 32)
    field = O.uniform_field(shape, seed=9)
    want, _ = O.diffusion7(field.astype(np.float64), K_BENCH, periodic=False, steps=20, dtype=np.float64)
    got, _ = run_hip(rt, field, K_BENCH, 0, 20, storage=storage)
    err = np.max(np.abs(got - want))
    print("storage", storage, "max abs error vs FP64", err, flush=True)
    assert err < tol


@pytest.mark.parametrize("storage", [1, 2])
@pytest.mark.parametrize("boundary", [0, 1])
def test_16bit_fused_equals_unfused(rt, storage, boundary):
    """With 16-bit storage the fused kernel rounds its on-chip intermediate
    exactly as a store/load would, so it matches separate steps bitwise."""
    field = O.uniform_field((64, 48, 40), seed=21)
    names = rt.diffusion7_variants()
    one = [v for v, n in names.items() if n == "star4_by8_z128"][0]
    fused = [v for v, n in names.items() if n.startswith("tb2_w1")][0]
    a1, a2 = run_hip(rt, field, K_ASYM, boundary, 10, variant=one, storage=storage)
    b1, b2 = run_hip(rt, field, K_ASYM, boundary, 10, variant=fused, storage=storage)
    assert np.array_equal(a1, b1) and np.array_equal(a2, b2)


@pytest.mark.parametrize("storage", [1, 2])
@pytest.mark.parametrize("boundary", [0, 1])
@pytest.mark.parametrize("shape", [(37, 29, 23), (130, 17, 19), (383, 6, 9)])
def test_16bit_packed_rows_fused_equal_unfused(rt, storage, boundary, shape):
    """16-bit grids whose rows do not fill 256-point waves: the default run
    (fused pairs on the packed-row tile, the last pass storing its
    intermediate step) and one step per launch (the packed one-step tile)
    leave both grids bitwise equal, for 10 and 11 steps."""
    field = O.uniform_field(shape, seed=31)
    names = rt.diffusion7_variants()
    one = [v for v, n in names.items() if n == "star4_by8_z128"][0]
    for steps in (10, 11):
        a1, a2 = run_hip(rt, field, K_ASYM, boundary, steps, variant=one, storage=storage)
        b1, b2 = run_hip(rt, field, K_ASYM, boundary, steps, variant=0, storage=storage)
        assert np.array_equal(a1, b1) and np.array_equal(a2, b2), "%d steps" % steps


def test_invalid_arguments_are_rejected(rt):
    g1 = rt.Grid((8, 8, 8))
    g2 = rt.Grid((8, 8, 4))
    with pytest.raises(rt.PhysisError):
        rt.diffusion7_run(g1, g2, K_BENCH, 0, 1)
    g3 = rt.Grid((8, 8, 8))
    with pytest.raises(rt.PhysisError):
        rt.diffusion7_run(g1, g3, K_BENCH, 5, 1)
    with pytest.raises(rt.PhysisError):
        rt.diffusion7_run(g1, g3, K_BENCH, 0, 1, dom=rt.domain3d(0, 9, 0, 8, 0, 8))
    for g in (g1, g2, g3):
        g.free()


@pytest.mark.parametrize("boundary", [0, 1])
def test_512cubed_1000_steps_vs_fp64_closed_form(rt, boundary):
    """The full benchmark run (512^3, 1000 steps) from a separable cosine
    mode, whose exact FP64 solution is the mode times lambda^1000: clamped,
    cos(pi k (j + 1/2) / n) with lambda = cc + 2 c_nb sum_d cos(pi k / n);
    periodic, cos(2 pi k j / n) with lambda = cc + 2 c_nb sum_d cos(2 pi k / n).
    The operator is a stochastic matrix (max-norm 1), so rounding the initial
    mode to FP32 adds at most half an ulp.  The error printed is the FP32
    target's own (the kernel equals the ref restatement bitwise); 64^3 and
    128^3 runs of the oracle give 1.1e-5 to 1.2e-5 of the peak value."""
    n, steps = 512, 1000
    j = np.arange(n, dtype=np.float64)
    if boundary:
        k = 1
        c = np.cos(2 * np.pi * k * j / n)
        lam = 0.4 + 0.6 * np.cos(2 * np.pi * k / n)
    else:
        k = 2
        c = np.cos(np.pi * k * (j + 0.5) / n)
        lam = 0.4 + 0.6 * np.cos(np.pi * k / n)
    cyx = (c[:, None] * c[None, :])
    field = (c[:, None, None] * cyx[None]).astype(np.float32)
    g1 = rt.Grid((n, n, n))
    g2 = rt.Grid((n, n, n))
    g1.copyin(field)
    rt.diffusion7_run(g1, g2, K_BENCH, boundary, steps)
    got = g1.copyout()
    g1.free()
    g2.free()
    decay = lam ** steps
    err = 0.0
    for z in range(0, n, 64):
        want = c[z:z + 64, None, None] * cyx[None] * decay
        err = max(err, float(np.max(np.abs(got[z:z + 64].astype(np.float64) - want))))
    rel = err / decay
    print("boundary", boundary, "lambda^1000", decay, "max error / peak vs FP64 closed form", rel, flush=True)
    assert rel < 5e-5
