"""Batched multi-grid diffusion on the GPU: every grid of the batch must end
bitwise equal to the oracle run on that grid alone, for every launch
configuration (one and two steps per launch), with and without hipGraph
capture, and across repeated graph replays with eager work in between."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def make_batch(rt, shapes, seed, boundary):
    rng = np.random.default_rng(seed)
    pairs, fields, coefs = [], [], []
    for i, shp in enumerate(shapes):
        f = O.uniform_field(shp, seed=seed * 1000 + i)
        g1, g2 = rt.Grid(shp), rt.Grid(shp)
        g1.copyin(f)
        g2.copyin(np.zeros_like(f))
        nb = rng.uniform(0.02, 0.15, 6).astype(np.float32)
        c = [float(v) for v in nb] + [float(np.float32(1.0) - np.float32(nb.sum()))]
        pairs.append((g1, g2))
        fields.append(f)
        coefs.append(c)
    return pairs, fields, coefs, rt.Batch(pairs, coefs, boundary)


# Rows of up to 128 points (every configuration below covers them), then
# whole 16-byte rows longer than 128 points (132, 256, 300, 200), which the
# mixed mode runs in its star classes (one or two waves per row).
NARROW = [(5, 7, 3), (70, 9, 11), (33, 40, 6), (64, 64, 64), (1, 1, 1), (17, 3, 29), (128, 5, 4), (2, 31, 8),
          (101, 12, 9), (8, 8, 8), (66, 2, 2), (127, 6, 5)]
SHAPES = NARROW + [(132, 9, 7), (256, 6, 5), (300, 4, 6), (200, 17, 33)]


def check_against_oracle(pairs, fields, coefs, boundary, steps):
    for (g1, g2), f, c in zip(pairs, fields, coefs):
        w1, w2 = O.diffusion7(f, c, periodic=bool(boundary), steps=steps)
        assert np.array_equal(g1.copyout(), w1), "grid %s" % (g1.shape,)
        assert np.array_equal(g2.copyout(), w2), "grid %s" % (g2.shape,)


@pytest.mark.parametrize("boundary", [0, 1])
# np = 0: the mixed mode (each class with its own kernel and stream); -1: the
# same with 16-row tiles for the 1- and 2-point classes.
@pytest.mark.parametrize("cfg", [(2, 4, 16, 1), (2, 8, 64, 1), (4, 8, 32, 1), (2, 8, 16, 2), (2, 16, 32, 2),
                                 (4, 8, 8, 2), (0, 8, 16, 1), (0, 16, 32, 2), (0, 8, 8, 2), (-1, 8, 32, 2),
                                 (-1, 8, 16, 1)])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_bitwise_vs_oracle(rt, boundary, cfg, graph):
    steps = 9
    shapes = SHAPES if cfg[0] <= 0 else NARROW
    pairs, fields, coefs, batch = make_batch(rt, shapes, seed=5, boundary=boundary)
    batch.configure(*cfg)
    print("config", batch.config_name(), "graph", graph, flush=True)
    batch.run(steps, graph=graph)
    check_against_oracle(pairs, fields, coefs, boundary, steps)
    batch.destroy()
    for g1, g2 in pairs:
        g1.free()
        g2.free()


@pytest.mark.parametrize("cfg", [(2, 8, 16, 2), (0, 8, 16, 2), (-2, 8, 12, 2)])
def test_graph_replays_with_eager_work_between(rt, cfg):
    steps = 6
    shapes = (PACKED[:6] if cfg[0] == -2 else NARROW[:6] + (SHAPES[-2:] if cfg[0] == 0 else []))
    pairs, fields, coefs, batch = make_batch(rt, shapes, seed=9, boundary=0)
    batch.configure(*cfg)
    for rep in range(1, 4):
        batch.run(steps, graph=True)
        # Eager work on the same stream between replays.
        junk = rt.Grid((16, 16, 16))
        junk.fill_uniform(rep)
        junk.reduce64(rt.PS_SUM)
        junk.free()
        rt.synchronize()
        check_against_oracle(pairs, fields, coefs, 0, steps * rep)
    batch.destroy()


def test_tune_restores_grids_and_picks_a_config(rt):
    pairs, fields, coefs, batch = make_batch(rt, SHAPES, seed=3, boundary=1)
    ms = batch.tune(4)
    print("tuned:", batch.config_name(), "%.3f ms" % ms, flush=True)
    assert ms > 0
    for (g1, g2), f in zip(pairs, fields):
        assert np.array_equal(g1.copyout(), f)
    batch.run(8)
    check_against_oracle(pairs, fields, coefs, 1, 8)
    batch.destroy()


def test_batch_rejects_wide_or_mismatched_grids(rt):
    a, b = rt.Grid((600, 4, 4)), rt.Grid((600, 4, 4))
    with pytest.raises(rt.PhysisError):
        rt.Batch([(a, b)], [[0.1] * 6 + [0.4]])
    c, d = rt.Grid((8, 4, 4)), rt.Grid((8, 4, 5))
    with pytest.raises(rt.PhysisError):
        rt.Batch([(c, d)], [[0.1] * 6 + [0.4]])


@pytest.mark.parametrize("cfg", [(0, 8, 16, 2), (0, 8, 64, 1)])
@pytest.mark.parametrize("boundary", [0, 1])
def test_bf16_mixed_batch_equals_single_grid_runs(rt, cfg, boundary):
    # BF16 storage, every class of the mixed mode (the star classes too):
    # each grid must end bitwise equal to the same grid advanced alone by
    # the plain single-grid run (both round to BF16 at every step).
    steps = 9
    rng = np.random.default_rng(11)
    batch_pairs, solo_pairs, coefs = [], [], []
    for i, shp in enumerate(SHAPES):
        f = O.uniform_field(shp, seed=700 + i)
        nb = rng.uniform(0.02, 0.15, 6).astype(np.float32)
        coefs.append([float(v) for v in nb] + [float(np.float32(1.0) - np.float32(nb.sum()))])
        for pairs in (batch_pairs, solo_pairs):
            g1, g2 = rt.Grid(shp, storage=rt.STORAGE_BF16), rt.Grid(shp, storage=rt.STORAGE_BF16)
            g1.copyin(f)
            g2.copyin(np.zeros_like(f))
            pairs.append((g1, g2))
    batch = rt.Batch(batch_pairs, coefs, boundary)
    batch.configure(*cfg)
    print("config", batch.config_name(), flush=True)
    batch.run(steps, graph=True)
    for (s1, s2), c in zip(solo_pairs, coefs):
        rt.diffusion7_run(s1, s2, c, boundary, steps)
    for (b1, b2), (s1, s2) in zip(batch_pairs, solo_pairs):
        assert np.array_equal(b1.copyout(), s1.copyout()), "grid %s" % (b1.shape,)
        assert np.array_equal(b2.copyout(), s2.copyout()), "grid %s" % (b2.shape,)
    batch.destroy()
    for g1, g2 in batch_pairs + solo_pairs:
        g1.free()
        g2.free()


@pytest.mark.parametrize("steps", [2, 4, 5, 8])
@pytest.mark.parametrize("cfg", [(2, 8, 16, 2), (0, 8, 16, 2), (0, 8, 64, 2)])
def test_batch_step_counts_bitwise_vs_oracle(rt, cfg, steps):
    """Fused runs of every pair-count parity: the passes follow a three-buffer
    path back to A, the last storing its intermediate step into B (even
    counts) or followed by one step A -> B (odd counts)."""
    shapes = SHAPES if cfg[0] <= 0 else NARROW
    pairs, fields, coefs, batch = make_batch(rt, shapes, seed=13, boundary=0)
    batch.configure(*cfg)
    print("config", batch.config_name(), "steps", steps, flush=True)
    batch.run(steps, graph=True)
    check_against_oracle(pairs, fields, coefs, 0, steps)
    batch.destroy()
    for g1, g2 in pairs:
        g1.free()
        g2.free()


# Rows of 4 to 512 points of every residue mod 4 (the packed tile aligns a
# row's last vector to its end), one to many rows per workgroup, thin and
# deep z.
PACKED = [(4, 9, 5), (5, 3, 7), (7, 17, 3), (13, 2, 9), (32, 5, 7), (33, 40, 6), (37, 1, 9), (63, 2, 3),
          (64, 64, 33), (65, 9, 1), (127, 6, 5), (139, 17, 12), (200, 3, 70), (255, 11, 4), (256, 6, 5)]
PACKED_WIDE = PACKED + [(257, 5, 6), (300, 4, 6), (511, 9, 3), (512, 3, 4)]


@pytest.mark.parametrize("boundary", [0, 1])
@pytest.mark.parametrize("cfg,shapes", [((-2, 8, 16, 2), PACKED), ((-2, 16, 64, 2), PACKED),
                                        ((-2, 16, 8, 2), PACKED_WIDE)])
@pytest.mark.parametrize("steps", [4, 9])
@pytest.mark.parametrize("graph", [False, True])
def test_packed_batch_bitwise_vs_oracle(rt, boundary, cfg, shapes, steps, graph):
    """The packed mode (rows of any length packed across lanes, four points
    per lane, two fused steps per launch over all grids): every grid equals
    the oracle bitwise, for even counts (the last pass stores the
    intermediate step) and odd ones (a single step follows)."""
    pairs, fields, coefs, batch = make_batch(rt, shapes, seed=21, boundary=boundary)
    batch.configure(*cfg)
    print("config", batch.config_name(), "steps", steps, "graph", graph, flush=True)
    batch.run(steps, graph=graph)
    check_against_oracle(pairs, fields, coefs, boundary, steps)
    batch.destroy()
    for g1, g2 in pairs:
        g1.free()
        g2.free()


@pytest.mark.parametrize("cfg,shapes", [((-2, 8, 16, 2), PACKED), ((-2, 16, 12, 2), PACKED_WIDE)])
@pytest.mark.parametrize("boundary", [0, 1])
@pytest.mark.parametrize("steps", [6, 9])
def test_bf16_packed_batch_equals_single_grid_runs(rt, cfg, shapes, boundary, steps):
    """BF16 storage on the packed mode: each grid ends bitwise equal to the
    same grid advanced alone by the plain single-grid run (both round to
    BF16 at every step and evaluate the body with the same fused
    multiply-adds), for rows of 4 to 512 points."""
    rng = np.random.default_rng(17)
    batch_pairs, solo_pairs, coefs = [], [], []
    for i, shp in enumerate(shapes):
        f = O.uniform_field(shp, seed=900 + i)
        nb = rng.uniform(0.02, 0.15, 6).astype(np.float32)
        coefs.append([float(v) for v in nb] + [float(np.float32(1.0) - np.float32(nb.sum()))])
        for pairs in (batch_pairs, solo_pairs):
            g1, g2 = rt.Grid(shp, storage=rt.STORAGE_BF16), rt.Grid(shp, storage=rt.STORAGE_BF16)
            g1.copyin(f)
            g2.copyin(np.zeros_like(f))
            pairs.append((g1, g2))
    batch = rt.Batch(batch_pairs, coefs, boundary)
    batch.configure(*cfg)
    print("config", batch.config_name(), "steps", steps, flush=True)
    batch.run(steps, graph=True)
    for (s1, s2), c in zip(solo_pairs, coefs):
        rt.diffusion7_run(s1, s2, c, boundary, steps)
    for (b1, b2), (s1, s2) in zip(batch_pairs, solo_pairs):
        assert np.array_equal(b1.copyout(), s1.copyout()), "grid %s" % (b1.shape,)
        assert np.array_equal(b2.copyout(), s2.copyout()), "grid %s" % (b2.shape,)
    batch.destroy()
    for g1, g2 in batch_pairs + solo_pairs:
        g1.free()
        g2.free()


def test_packed_mode_rejects_rows_it_does_not_serve(rt):
    # Rows shorter than 4 points, rows over 256 points for 8 waves per
    # workgroup, and one step per launch are not served by the packed mode.
    for shapes, cfg in ((NARROW, (-2, 8, 16, 2)), (PACKED_WIDE, (-2, 8, 16, 2)), (PACKED, (-2, 16, 16, 1))):
        pairs, _, _, batch = make_batch(rt, shapes, seed=1, boundary=0)
        with pytest.raises(rt.PhysisError):
            batch.configure(*cfg)
        batch.destroy()
        for g1, g2 in pairs:
            g1.free()
            g2.free()


@pytest.mark.parametrize("mode", ["tuned", "packed"])
def test_batched_workload_256_grids_200_steps_equal_single_grid_runs(rt, mode):
    """The batched workload at its size: 256 FP32 grids with edges drawn
    log-uniform in [32, 256] per axis (seed 2024), fields uniform [0, 1),
    clamped, 200 steps with the autotuned launch shape (or the packed mode
    with 8 waves and 64-plane chunks) and one hipGraph.  Every grid (both
    buffers) ends bitwise equal to the same grid advanced alone by the plain
    single-grid run, and the three smallest grids equal the oracle."""
    coef = [0.1] * 6 + [0.4]
    steps = 200
    rng = np.random.default_rng(2024)
    shapes = [tuple(int(round(v)) for v in np.exp(rng.uniform(np.log(32), np.log(256), 3))) for _ in range(256)]
    pairs = [(rt.Grid(s), rt.Grid(s)) for s in shapes]
    for i, (a, b) in enumerate(pairs):
        a.fill_uniform(1000 + i)
        b.fill_uniform(5000 + i)
    batch = rt.Batch(pairs, [coef] * len(pairs), rt.CLAMPED)
    if mode == "tuned":
        batch.tune(8)
    else:
        batch.configure(-2, 8, 64, 2)
    print(mode, batch.config_name(), flush=True)
    for i, (a, b) in enumerate(pairs):
        a.fill_uniform(1000 + i)
        b.fill_uniform(5000 + i)
    batch.run(steps)
    rt.synchronize()
    print("batched run done", flush=True)
    smallest = sorted(range(len(shapes)), key=lambda i: np.prod(shapes[i]))[:3]
    for i, (a, b) in enumerate(pairs):
        s1, s2 = rt.Grid(shapes[i]), rt.Grid(shapes[i])
        s1.fill_uniform(1000 + i)
        s2.fill_uniform(5000 + i)
        rt.diffusion7_run(s1, s2, coef, rt.CLAMPED, steps)
        ga, gb, w1, w2 = a.copyout(), b.copyout(), s1.copyout(), s2.copyout()
        assert np.array_equal(ga, w1), "grid %d %s" % (i, shapes[i])
        assert np.array_equal(gb, w2), "grid %d %s" % (i, shapes[i])
        if i in smallest:
            o1, o2 = O.diffusion7(O.uniform_field(shapes[i], seed=1000 + i), coef, periodic=False, steps=steps)
            assert np.array_equal(ga, o1) and np.array_equal(gb, o2), "grid %d vs oracle" % i
        s1.free()
        s2.free()
    batch.destroy()
    for a, b in pairs:
        a.free()
        b.free()
