"""The reference's DSL system tests, translated by physisc-hip, run on the GPU.

The executables are built from the reference checkout by
tests/golden/build_corpus.py into tests/_corpus_bin/.  As in the reference's
run_system_tests.sh (run_system_tests.sh.cmake:808-856), a program passes if
its stdout equals the output of the reference's hand-written ref oracle
(compared through the SHA-256 in tests/golden/manual_ref.json); the
self-checking programs (test_01, 02, 03, 08-14, multi-kernels and the
user-defined point type checks) must exit 0.
"""
import hashlib
import json
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "_corpus_bin")
GOLDEN = json.load(open(os.path.join(HERE, "golden", "manual_ref.json")))

# Programs whose output is compared with a manual.ref oracle.
COMPARED = ["test_7-pt", "test_7-pt-periodic", "test_7-pt-neumann-cond", "test_7-pt-multi-iterations",
            "test_7-pt-double-type", "test_7-pt-int-type", "test_7-pt-type-mix", "test_15", "test_16",
            "test_27-pt", "test_27-pt-periodic", "test_3-pt-periodic", "test_3-pt-1d", "test_9-pt-2d",
            "test_asymmetric", "test_asymmetric-periodic", "test_stencil-hole", "test_mixed-dim",
            "test_mixed-dim2", "test_mixed-dim3", "test_redblack", "test_redblack-periodic",
            "test_redblack-separated", "test_user-defined-type1", "test_user-defined-type3",
            "test_user-defined-type5", "test_user-defined-type-multi-dim-member",
            "test_user-defined-type-multi-members", "test_7-pt.module"]
# Programs without a manual.ref of their own, compared with the oracle of an
# equivalent program: MapRed followed by MapBlack is one MapRedBlack.
# The module test (a DSL unit driven from plain C) runs test_7-pt's program.
GOLDEN_OF = {"test_redblack-separated": "test_redblack", "test_7-pt.module": "test_7-pt"}
SELF_CHECKING = ["test_01", "test_02", "test_03", "test_08", "test_09", "test_10", "test_11", "test_12",
                 "test_13", "test_14", "test_multi-kernels", "test_user-defined-type2", "test_user-defined-type4",
                 "test_user-defined-type-copyin-copyout", "test_user-defined-type-kernel-copy",
                 "test_user-defined-type-transpose"]


def exe(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.skip("%s not built (tests/golden/build_corpus.py needs the reference checkout)" % name)
    return path


@pytest.mark.parametrize("name", COMPARED)
def test_output_matches_reference_oracle(name):
    print("running", name, flush=True)
    out = subprocess.run([exe(name)], capture_output=True, timeout=120)
    assert out.returncode == 0, out.stderr.decode()[-2000:]
    text = out.stdout
    g = GOLDEN[GOLDEN_OF.get(name, name)]
    if hashlib.sha256(text).hexdigest() != g["sha256"]:
        lines = text.decode().splitlines()
        assert len(lines) == g["lines"], "line count %d != %d" % (len(lines), g["lines"])
        for i, want in enumerate(g["samples"]):
            got = lines[i * g["sample_every"]]
            assert got == want, "line %d: %s != %s" % (i * g["sample_every"], got, want)
        pytest.fail("output differs from the reference oracle outside the sampled lines")


@pytest.mark.parametrize("name", SELF_CHECKING)
def test_self_checking_program_passes(name):
    print("running", name, flush=True)
    out = subprocess.run([exe(name)], capture_output=True, timeout=120)
    assert out.returncode == 0, (out.stdout.decode()[-1000:], out.stderr.decode()[-2000:])


def run_benchmark(path, size, count, cwd):
    out = subprocess.run([path, "--size", str(size), "--count", str(count), "--dump"], capture_output=True,
                         timeout=120, cwd=cwd)
    assert out.returncode == 0, out.stderr.decode()[-2000:]
    acc = [l for l in out.stdout.decode().splitlines() if l.startswith("Accuracy")]
    return acc


@pytest.mark.parametrize("size,count", [(64, 100), (48, 20), (100, 8)])
def test_reference_diffusion_benchmark_matches_its_cpu_baseline(tmp_path, size, count):
    """The reference's diffusion benchmark (examples/diffusion-benchmark): its
    Physis program translated by physisc-hip and linked with the benchmark's
    own harness (main.cc built for the Physis variant) dumps the same field,
    to the last printed digit of every element, as the same harness built
    for its sequential CPU baseline (diffusion3d.cc), and reports the same
    accuracy against the analytic solution."""
    bench = exe("diffusion_benchmark_physis")
    base = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "diffusion_baseline.exe")
    if not os.path.exists(base):
        pytest.skip("the reference's CPU baseline is not built (oracle/Makefile ref)")
    print("benchmark", size, count, flush=True)
    d_hip, d_cpu = tmp_path / "hip", tmp_path / "cpu"
    d_hip.mkdir()
    d_cpu.mkdir()
    acc_hip = run_benchmark(bench, size, count, d_hip)
    acc_cpu = run_benchmark(base, size, count, d_cpu)
    assert acc_hip and acc_hip == acc_cpu
    got = (d_hip / "diffusion3d_result.physis.out").read_bytes()
    want = (d_cpu / "diffusion3d_result.baseline.out").read_bytes()
    assert len(got) == len(want) and got == want


RAGGED = [(37, 29, 23, 5), (130, 17, 19, 4), (383, 6, 9, 3), (510, 7, 5, 3), (19, 3, 2, 6), (64, 48, 40, 4),
          (256, 9, 6, 3)]


@pytest.mark.parametrize("periodic", [0, 1])
@pytest.mark.parametrize("shape", RAGGED)
def test_star_program_ragged_rows_bitwise_vs_oracle(tmp_path, shape, periodic):
    """A star-form DSL program (tests/golden/programs/star_ragged.c)
    translated by physisc-hip, on rows of any length (the fused pairs of
    FP32 grids run on the packed-row tile unless the rows fill whole
    256-point waves of the vectorised tile): both grids equal the oracle
    bitwise after 2*ITER maps."""
    import numpy as np
    from oracle import oracle as O
    nx, ny, nz, it = shape
    field = O.uniform_field((nx, ny, nz), seed=40 + nx)
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    field.astype(np.float32).tofile(src)
    out = subprocess.run([exe("star_ragged"), str(nx), str(ny), str(nz), str(it), str(periodic), str(src), str(dst),
                          "--physis-trace"], capture_output=True, timeout=120)
    assert out.returncode == 0, out.stderr.decode()[-2000:]
    got = np.fromfile(dst, np.float32).reshape(2, nz, ny, nx)
    want1, want2 = O.diffusion7(field, [0.11, 0.07, 0.05, 0.13, 0.02, 0.17, 0.45], periodic=bool(periodic),
                                steps=2 * it)
    assert np.array_equal(got[0], want1) and np.array_equal(got[1], want2)
    # Every pair ran fused, on the tile the row width calls for.
    tile = "vectorised" if nx % 256 == 0 else "packed"
    assert ("Physis: fused pairs: %d passes on the %s tile" % (it, tile)) in out.stderr.decode()


def wave_oracle(u, v, k, iters):
    """The leapfrog program of tests/golden/programs/wave_pingpong.c evaluated
    on the CPU in FP32, operation for operation as its kernel body is written
    (w + e + n + s + d + t - 6c, then 2c - b + k*lap), with separate steps:
    map(u -> v) then map(v -> u), iters times."""
    import numpy as np
    k = np.float32(k)

    def step(a, b):
        c = a
        lap = (np.roll(a, 1, 2) + np.roll(a, -1, 2) + np.roll(a, 1, 1) + np.roll(a, -1, 1) + np.roll(a, 1, 0)
               + np.roll(a, -1, 0) - np.float32(6.0) * c)
        return np.float32(2.0) * c - b + k * lap

    for _ in range(iters):
        v = step(u, v)
        u = step(v, u)
    return u, v


@pytest.mark.parametrize("shape", [(64, 24, 20, 5), (37, 11, 9, 4)])
def test_kernel_reading_its_output_grid_runs_unfused_and_matches(tmp_path, shape):
    """A ping-pong pair whose kernel reads the grid it emits (leapfrog wave,
    tests/golden/programs/wave_pingpong.c) is not fused: both grids equal a
    step-by-step FP32 evaluation of the program bitwise, and no fused pass
    is traced."""
    import numpy as np
    from oracle import oracle as O
    nx, ny, nz, it = shape
    u = O.uniform_field((nx, ny, nz), seed=7)
    v = (O.uniform_field((nx, ny, nz), seed=8) * np.float32(0.5)).astype(np.float32)
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    np.concatenate([u.ravel(), v.ravel()]).astype(np.float32).tofile(src)
    print("wave", shape, flush=True)
    out = subprocess.run([exe("wave_pingpong"), str(nx), str(ny), str(nz), str(it), str(src), str(dst),
                          "--physis-trace"], capture_output=True, timeout=120)
    assert out.returncode == 0, out.stderr.decode()[-2000:]
    assert "fused pairs" not in out.stderr.decode()
    got = np.fromfile(dst, np.float32).reshape(2, nz, ny, nx)
    want_u, want_v = wave_oracle(u, v, 0.1, it)
    assert np.array_equal(got[0], want_u) and np.array_equal(got[1], want_v)


@pytest.mark.parametrize("size,dims", [("XS", (64, 32, 32)), ("S", (128, 64, 64)), ("M", (256, 128, 128))])
def test_reference_himeno_program_translated(tmp_path, size, dims):
    """The reference's Himeno program (examples/himeno/himenobmtxpa_physis.c,
    the 19-point Jacobi of the Himeno workload), translated unchanged by
    physisc-hip (tests/golden/build_corpus.py): its grids, initialisation
    (p = (k/(nz-1))^2, a = c = bnd = 1, a3 = 1/6, b = wrk1 = 0,
    himenobmtxpa_physis.c:120-151, 297-329), its ping-pong PSStencilRun of
    jacobi_kernel over the interior domain (:331-393) and its own dump of p
    after the four rehearsal sweeps (ENABLE_DUMP, :177-194, "%f" per point).  Every dumped
    line equals the oracle's FP32 value printed the same way, and the
    measured loop (18 sweeps under the stand-in clock) completes."""
    import numpy as np
    from oracle import oracle as O
    nx, ny, nz = dims
    print("himeno", size, flush=True)
    out = subprocess.run([exe("himeno_physis"), size], capture_output=True, timeout=120, cwd=tmp_path)
    assert out.returncode == 0, out.stderr.decode()[-2000:]
    text = out.stdout.decode()
    assert "mimax = %d mjmax = %d mkmax = %d" % dims in text
    assert "Loop executed for 18 times" in text
    got = (tmp_path / "himeno.physis.dat").read_text().split("\n")[:-1]
    assert len(got) == nx * ny * nz
    k = np.arange(nz, dtype=np.int64)
    col = (k * k).astype(np.float32) / np.float32((nz - 1) * (nz - 1))
    p = np.ascontiguousarray(np.broadcast_to(col[:, None, None], (nz, ny, nx)), dtype=np.float32)
    vals = [1.0, 1.0, 1.0, 1.0 / 6.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0, 0.0]
    coefs = [np.full((nz, ny, nx), np.float32(v), np.float32) for v in vals]
    want, _, _ = O.himeno(p, coefs, np.float32(0.8), 4)
    want_lines = ["%f" % v for v in want.ravel().tolist()]
    bad = [i for i, (g, w) in enumerate(zip(got, want_lines)) if g != w]
    assert not bad, "%d points differ, first at %d: %s vs %s" % (len(bad), bad[0], got[bad[0]], want_lines[bad[0]])


def bf16_round(x):
    """FP32 values rounded to the nearest BF16 (ties to even), kept as FP32."""
    import numpy as np
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("dims,sweeps,storage", [((65, 33, 33), 8, "fp32"), ((66, 34, 17), 6, "fp32"),
                                                 ((65, 33, 33), 8, "bf16")])
def test_reference_himeno_jacobi_timing_harness(dims, sweeps, storage):
    """The benchmark's drop-in Himeno leg: the reference's Himeno program
    translated with TRACE_KERNEL and driven through its own mat_set,
    mat_set_init and jacobi (tests/golden/himeno_timing.cc, 4 then `sweeps`
    sweeps).  FP32 takes the fused point-form run; BF16 storage runs one map
    per launch.  The FP64 sum of p in memory order equals, exactly, that of
    the oracle's FP32 field after the same sweeps (for BF16: the oracle's
    sweep on the BF16 values, its result rounded to BF16 after each sweep)."""
    import numpy as np
    from oracle import oracle as O
    nx, ny, nz = dims
    print("himeno timing harness", dims, sweeps, storage, flush=True)
    args = [exe("himeno_dropin_timing"), str(nx), str(ny), str(nz), str(sweeps)]
    out = subprocess.run(args + ([storage] if storage == "bf16" else []) + ["--physis-trace"],
                         capture_output=True, timeout=120)
    assert out.returncode == 0, out.stderr.decode()[-2000:]
    err = out.stderr.decode()
    assert len(re.findall(r"Stencil finished \(time: ([0-9.]+)\)", err)) == 2
    assert ("fused sweeps" in err) == (storage == "fp32")
    got = float(re.search(r"p0 sum: (\S+)", err).group(1))
    k = np.arange(nz, dtype=np.int64)
    col = (k * k).astype(np.float32) / np.float32((nz - 1) * (nz - 1))
    p = np.ascontiguousarray(np.broadcast_to(col[:, None, None], (nz, ny, nx)), dtype=np.float32)
    vals = [1.0, 1.0, 1.0, 1.0 / 6.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0, 0.0]
    coefs = [np.full((nz, ny, nx), np.float32(v), np.float32) for v in vals]
    if storage == "fp32":
        want, _, _ = O.himeno(p, coefs, np.float32(0.8), 4 + sweeps)
    else:
        coefs = [bf16_round(c) for c in coefs]
        src, dst = bf16_round(p), bf16_round(p)
        for _ in range(4 + sweeps):
            _, nxt, _ = O.himeno(src, coefs, np.float32(0.8), 1, p1_init=dst)
            src, dst = bf16_round(nxt), src
        want = src
    total = 0.0
    for v in want.ravel().tolist():
        total += v
    assert got == total


def coef_extents_oracle(p, a, b, c, w, iters):
    """FP32 evaluation of tests/golden/programs/coef_extents.c in its own
    order: ((a*p(x+1) + b*p(y-1)) + c(x, y, z+1)*p) + w*a(x-1) over the
    interior, maps p -> q then q -> p (arrays indexed [z, y, x]; c may have
    more rows than p)."""
    import numpy as np
    w = np.float32(w)
    nz, ny, nx = p.shape
    I = (slice(1, nz - 1), slice(1, ny - 1), slice(1, nx - 1))
    c1 = c[2:nz, 1:ny - 1, 1:nx - 1]  # c(x, y, z + 1) on the interior

    def relax(src, dst):
        out = dst.copy()
        v = a[I] * src[1:-1, 1:-1, 2:] + b[I] * src[1:-1, :-2, 1:-1]
        v = v + c1 * src[I]
        v = v + w * a[1:-1, 1:-1, :-2]
        out[I] = v
        return out

    q = p.copy()
    for _ in range(iters):
        q = relax(p, q)
        p = relax(q, p)
    return p, q


@pytest.mark.parametrize("mode", [0, 1])
def test_translated_plan_with_direct_grids_both_extents_paths(tmp_path, mode):
    """A map reading three grids besides the staged one (one of them at
    z + 1 and one through the periodic form), run as a ping-pong over the
    interior: with every grid of one shape the launch takes the uniform-
    extents plan, with grid c one row taller (MODE 1) the per-grid one; both
    equal an FP32 evaluation in the program's order bitwise."""
    import numpy as np
    from oracle import oracle as O
    nx, ny, nz, it = 70, 23, 19, 3
    p = O.uniform_field((nx, ny, nz), seed=11)
    a = O.uniform_field((nx, ny, nz), seed=12)
    b = O.uniform_field((nx, ny, nz), seed=13)
    c = O.uniform_field((nx, ny + mode, nz), seed=14)
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    np.concatenate([p.ravel(), a.ravel(), b.ravel(), c.ravel()]).astype(np.float32).tofile(src)
    print("coef_extents mode", mode, flush=True)
    out = subprocess.run([exe("coef_extents"), str(nx), str(ny), str(nz), str(mode), str(it), str(src), str(dst)],
                         capture_output=True, timeout=120)
    assert out.returncode == 0, out.stderr.decode()[-2000:]
    got = np.fromfile(dst, np.float32).reshape(2, nz, ny, nx)
    want_p, want_q = coef_extents_oracle(p, a, b, c, 0.25, it)
    assert np.array_equal(got[0], want_p) and np.array_equal(got[1], want_q)


def point_relax_oracle(p, q, a, b, c, w, iters):
    """FP32 evaluation of tests/golden/programs/point_relax.c in its own
    order over the interior: s = (a*(p(x+1) + p(x-1)) + b*(p(y+1,z-1) -
    p(y-1,z+1))) + c*p(x+1,y+1,z+1); v = p + w*(s - p); maps p -> q then
    q -> p (arrays indexed [z, y, x])."""
    import numpy as np
    w = np.float32(w)
    nz, ny, nx = p.shape
    I = (slice(1, nz - 1), slice(1, ny - 1), slice(1, nx - 1))

    def at(g, dx, dy, dz):
        return g[1 + dz:nz - 1 + dz, 1 + dy:ny - 1 + dy, 1 + dx:nx - 1 + dx]

    def relax(src, dst):
        out = dst.copy()
        s = a[I] * (at(src, 1, 0, 0) + at(src, -1, 0, 0))
        s = s + b[I] * (at(src, 0, 1, -1) - at(src, 0, -1, 1))
        s = s + c[I] * at(src, 1, 1, 1)
        out[I] = src[I] + w * (s - src[I])
        return out

    for _ in range(iters):
        q = relax(p, q)
        p = relax(q, p)
    return p, q


@pytest.mark.parametrize("shape,iters", [((70, 40, 23), 1), ((70, 40, 23), 2), ((70, 40, 23), 3),
                                         ((129, 17, 9), 5), ((19, 7, 5), 4), ((3, 3, 3), 2), ((200, 61, 45), 6)])
def test_point_form_map_fuses_two_sweeps_bitwise(tmp_path, shape, iters):
    """A ping-pong of a point-form kernel (staged grid within radius 1 with
    diagonals, three coefficient grids at the point, the interior domain)
    runs two sweeps per pass from two iterations on (the trace says so),
    one sweep per launch at one; both grids equal an FP32 evaluation in
    the program's order bitwise (q starts different from p, shell
    included, so a pass that took the wrong iterate's shell would show)."""
    import numpy as np
    from oracle import oracle as O
    nx, ny, nz = shape
    p = O.uniform_field(shape, seed=21)
    q = O.uniform_field(shape, seed=22, lo=-1.0, hi=1.0)
    a, b, c = (O.uniform_field(shape, seed=s, lo=0.1, hi=0.4) for s in (23, 24, 25))
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    np.concatenate([x.ravel() for x in (p, q, a, b, c)]).astype(np.float32).tofile(src)
    print("point_relax", shape, "iter", iters, flush=True)
    out = subprocess.run([exe("point_relax"), str(nx), str(ny), str(nz), str(iters), str(src), str(dst),
                          "--physis-trace"], capture_output=True, timeout=120)
    assert out.returncode == 0, out.stderr.decode()[-2000:]
    assert ("fused sweeps: %d" % (2 * iters) in out.stderr.decode()) == (iters > 1)
    got = np.fromfile(dst, np.float32).reshape(2, nz, ny, nx)
    want_p, want_q = point_relax_oracle(p, q, a, b, c, 0.7, iters)
    assert np.array_equal(got[0], want_p) and np.array_equal(got[1], want_q)


def reduce_kernel_oracle(f, bf16):
    """Each reduction of tests/golden/programs/reduce_kernel.c restated with
    numpy: the kernel's value at every point of its domain, evaluated in the
    kernel's own order (FP32 for the float kernels, FP64 for the double
    ones; the float grid rounded to BF16 first when it is stored so), then
    folded in FP64 and converted to the kernel's type."""
    import numpy as np
    g = bf16_round(f) if bf16 else f.astype(np.float32)
    h = 0.5 * f.astype(np.float64) + 0.25
    d = g[:, :, 1:] - g[:, :, :-1]
    sq = (np.float32(0.75) * d) * d
    cv = g - np.float32(0.5) * (np.roll(g, 1, axis=1) + np.roll(g, -1, axis=1))
    cv = np.abs(cv[1:-1])
    out = {"sqdiff_sum": np.float32(np.sum(sq, dtype=np.float64)),
           "curv_max": np.float32(cv.max()), "curv_min": np.float32(cv.min()),
           "dsq_sum": float(np.sum(h * h + g.astype(np.float64))),
           "dprod_prod": float(np.prod(1.0 + 1e-3 * h[0:2, 0:2, :])),
           "row2_sum": np.float32(np.sum(np.float32(0.25) * f[0].astype(np.float32), dtype=np.float64)),
           "line1_min": np.float32((f[0, 0] - np.roll(f[0, 0], -1)).min()),
           "empty_sum": np.float32(0.0), "empty_max": np.float32(-np.inf)}
    return out


@pytest.mark.parametrize("shape,bf16", [((67, 33, 21), 0), ((300, 17, 9), 1), ((1030, 5, 4), 0),
                                        ((16, 300, 200), 0), ((5, 4, 3), 0)])
def test_kernel_reductions_match_oracle(tmp_path, shape, bf16):
    """PSReduce(&v, op, kernel, dom, args...) lowered by physisc-hip (the
    interface the reference documents but does not implement): float SUM
    (an x difference, squared and weighted), MAX and MIN (a periodic y
    curvature over a z-interior), double SUM and PROD (a double grid mixed
    with the float one), a 2-D kernel, a 1-D one with a periodic read, and
    empty domains (the operator's identity).  MAX and MIN equal the oracle exactly; the FP64-folded sums
    and products agree to FP64 rounding (the float results to one FP32
    ulp).  Rows longer than one workgroup item (1030 points) and more items
    than workgroups (60000 rows) are covered."""
    import numpy as np
    from oracle import oracle as O
    nx, ny, nz = shape
    f = O.uniform_field(shape, seed=60 + nx)
    src = tmp_path / "in.bin"
    f.astype(np.float32).tofile(src)
    print("reduce_kernel", shape, "bf16", bf16, flush=True)
    out = subprocess.run([exe("reduce_kernel"), str(nx), str(ny), str(nz), str(src), str(bf16)],
                         capture_output=True, timeout=120)
    assert out.returncode == 0, out.stderr.decode()[-2000:]
    got = dict(l.split() for l in out.stdout.decode().splitlines())
    want = reduce_kernel_oracle(f, bf16)
    assert set(got) == set(want)
    for name in ("curv_max", "curv_min", "line1_min", "empty_sum", "empty_max"):
        assert np.float32(float(got[name])) == want[name], (name, got[name], want[name])
    for name in ("sqdiff_sum", "row2_sum"):
        g32, w32 = np.float32(float(got[name])), want[name]
        assert abs(float(g32) - float(w32)) <= 2.0 ** -23 * abs(float(w32)), (name, got[name], w32)
    for name in ("dsq_sum", "dprod_prod"):
        assert abs(float(got[name]) - want[name]) <= 1e-13 * abs(want[name]), (name, got[name], want[name])


@pytest.mark.parametrize("n", [9, 70])
def test_kernel_reduction_over_a_point_type_grid(n):
    """A reduction kernel reading members of a user-defined point type grid
    (tests/golden/programs/reduce_point_type.c): the SUM of x(p)^2 +
    y(p + e_x) * z(p) agrees with the FP32 values folded in FP64 to one FP32
    ulp, the MAX of x - y exactly."""
    import numpy as np
    print("reduce_point_type", n, flush=True)
    out = subprocess.run([exe("reduce_point_type"), str(n)], capture_output=True, timeout=120)
    assert out.returncode == 0, out.stderr.decode()[-2000:]
    got = dict(l.split() for l in out.stdout.decode().splitlines())
    i = np.arange(n ** 3, dtype=np.float32).reshape(n, n, n)
    px = np.float32(0.01) * i
    py = np.float32(1.0) - np.float32(0.001) * i
    pz = np.full_like(px, np.float32(0.5))
    vals = px[:, :, :-1] * px[:, :, :-1] + py[:, :, 1:] * pz[:, :, :-1]
    want_s = np.float32(np.sum(vals, dtype=np.float64))
    want_m = np.float32((px - py).max())
    assert abs(float(got["norm2_sum"]) - float(want_s)) <= 2.0 ** -23 * abs(float(want_s))
    assert np.float32(float(got["spread_max"])) == want_m
