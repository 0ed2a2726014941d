"""The continuous-batching solve server on the GPU.

Each request's result (step count and the FP64 checksum of its final field)
must be bitwise the same served alone, in the mix, under a tight HBM budget,
and in the one-at-a-time sequential mode; FP32 requests also match the CPU
oracle run for the same number of steps."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

MIX = [
    dict(kind=0, n=24, storage=0, seed=11, arrival=0.000, max_steps=120, tol=1e-5),
    dict(kind=1, n=19, storage=0, seed=12, arrival=0.002, max_steps=120, tol=1e-5),
    dict(kind=0, n=37, storage=1, seed=13, arrival=0.004, max_steps=80, tol=1e-5),
    dict(kind=0, n=50, storage=0, seed=14, arrival=0.006, max_steps=120, tol=1e-5),
    dict(kind=1, n=33, storage=1, seed=15, arrival=0.008, max_steps=60, tol=1e-5),
    dict(kind=0, n=8, storage=0, seed=16, arrival=0.010, max_steps=2000, tol=1e-3),
    dict(kind=0, n=61, storage=0, seed=17, arrival=0.012, max_steps=40, tol=1e-5),
    dict(kind=1, n=26, storage=0, seed=18, arrival=0.014, max_steps=100, tol=1e-5),
]
TICK = 20
BIG_BUDGET = 8 << 30


def key(r):
    return (r["status"], r["steps"], r["checksum"])


@pytest.fixture(scope="module")
def mixed(rt):
    return rt.serve_trace(MIX, BIG_BUDGET, TICK, rt.SERVE_BATCHED)


def test_all_requests_complete(rt, mixed):
    for r, q in zip(mixed, MIX):
        assert r["status"] == 0
        assert r["steps"] % TICK == 0 and r["steps"] >= TICK
        assert r["steps"] <= q["max_steps"] + TICK - 1
        assert r["finished"] >= r["admitted"] >= 0.0
        if r["steps"] < q["max_steps"]:
            assert r["residual"] <= q["tol"]
    # The small diffusion request converges to its loose tolerance early.
    assert mixed[5]["steps"] < 2000


def test_same_alone_and_in_the_mix(rt, mixed):
    for q, r in zip(MIX, mixed):
        alone = rt.serve_trace([dict(q, arrival=0.0)], BIG_BUDGET, TICK, rt.SERVE_BATCHED)[0]
        assert key(alone) == key(r), q


def test_sequential_mode_gives_identical_results(rt, mixed):
    seq = rt.serve_trace(MIX, BIG_BUDGET, TICK, rt.SERVE_SEQUENTIAL)
    for q, a, b in zip(MIX, mixed, seq):
        assert key(a) == key(b), q


def test_tight_budget_admits_one_at_a_time_with_same_results(rt, mixed):
    biggest = max(q["n"] ** 3 * (4 if q["storage"] == 0 else 2) * (3 if q["kind"] == 0 else 15) for q in MIX)
    res = rt.serve_trace(MIX, biggest, TICK, rt.SERVE_BATCHED)
    for q, a, b in zip(MIX, mixed, res):
        assert key(a) == key(b), q


def test_request_larger_than_budget_is_rejected(rt):
    res = rt.serve_trace([dict(kind=0, n=64, storage=0, seed=1, arrival=0.0, max_steps=20, tol=0.0)],
                         64 ** 3 * 4 * 3 - 1, TICK, rt.SERVE_BATCHED)
    assert res[0]["status"] == 2


def test_fp32_results_match_oracle(rt, mixed):
    for q, r in zip(MIX, mixed):
        if q["storage"] != 0:
            continue
        n, steps = q["n"], r["steps"]
        if q["kind"] == 0:
            field = O.uniform_field((n, n, n), seed=q["seed"])
            g1, _ = O.diffusion7(field, [0.1] * 6 + [0.4], periodic=False, steps=steps)
            want = float(np.sum(g1.astype(np.float64)))
        else:
            want = jacobi_checksum(q["seed"], n, steps)
        assert abs(r["checksum"] - want) <= 1e-12 * max(1.0, abs(want)), (q, r["checksum"], want)


def jacobi_checksum(seed, n, steps):
    """FP64 sum of the oracle's p0 after `steps` sweeps of a served Jacobi
    request: a0..a2 and c0..c2 uniform in [0.9, 1.0], a3 = 1/6, b = 0,
    bnd = 1, wrk1 = 0, p = (z/(n-1))^2 on an n^3 grid."""
    coefs = []
    for i in range(12):
        if i == 3:
            coefs.append(np.full((n, n, n), np.float32(1.0) / np.float32(6.0), np.float32))
        elif 4 <= i <= 6 or i == 11:
            coefs.append(np.zeros((n, n, n), np.float32))
        elif i == 10:
            coefs.append(np.ones((n, n, n), np.float32))
        else:
            coefs.append(O.uniform_field((n, n, n), seed=seed + i, lo=0.9, hi=1.0))
    z = np.arange(n, dtype=np.float32)
    p = np.broadcast_to((z * z / np.float32((n - 1) * (n - 1)))[:, None, None], (n, n, n)).astype(np.float32)
    p0, _, _ = O.himeno(p, coefs, 0.8, steps)
    return float(np.sum(p0.astype(np.float64)))


def test_large_jacobi_requests_run_in_arrival_order_and_small_work_overlaps(rt):
    """Large Jacobi requests (grids of 256 MiB or more: 15 FP32 grids of
    176^3) sweep one at a time in arrival order, so the earlier one finishes
    first; a small Jacobi request and a diffusion request that arrive with
    them are served alongside and finish before either.  Results are the
    same as served alone."""
    reqs = [
        dict(kind=1, n=176, storage=0, seed=31, arrival=0.000, max_steps=400, tol=0.0),
        dict(kind=1, n=176, storage=0, seed=32, arrival=0.001, max_steps=400, tol=0.0),
        dict(kind=1, n=40, storage=0, seed=33, arrival=0.002, max_steps=100, tol=0.0),
        dict(kind=0, n=48, storage=0, seed=34, arrival=0.002, max_steps=100, tol=0.0),
    ]
    res = rt.serve_trace(reqs, BIG_BUDGET, TICK, rt.SERVE_BATCHED)
    assert all(r["status"] == 0 for r in res)
    assert [r["steps"] for r in res] == [400, 400, 100, 100]
    print("finished", [round(r["finished"], 4) for r in res], flush=True)
    assert res[0]["finished"] < res[1]["finished"]
    assert max(res[2]["finished"], res[3]["finished"]) < res[0]["finished"]
    for q, r in zip(reqs, res):
        alone = rt.serve_trace([dict(q, arrival=0.0)], BIG_BUDGET, TICK, rt.SERVE_BATCHED)[0]
        assert key(alone) == key(r), q


def workload_trace():
    """The served workload: 24 requests with Poisson arrivals at 40/s (seed
    7), 70% 7-point diffusion and 30% 19-point Jacobi, cubic edges uniform in
    [32, 384], FP32 or BF16 storage per request, each to residual 1e-5 or
    2000 steps; plus two requests of 350^3 or more so that both kinds come in
    both storage kinds at that size (the seeded trace has a 377^3 FP32
    diffusion and a 383^3 BF16 Jacobi request)."""
    rng = np.random.default_rng(7)
    t, reqs = 0.0, []
    for i in range(24):
        t += rng.exponential(1.0 / 40.0)
        reqs.append(dict(kind=0 if rng.random() < 0.7 else 1, n=int(rng.integers(32, 385)),
                         storage=int(rng.random() < 0.5), seed=1000 + i, arrival=t, max_steps=2000, tol=1e-5))
    reqs.append(dict(kind=0, n=360, storage=1, seed=2000, arrival=0.30, max_steps=2000, tol=1e-5))
    reqs.append(dict(kind=1, n=352, storage=0, seed=2001, arrival=0.31, max_steps=2000, tol=1e-5))
    reqs.sort(key=lambda r: r["arrival"])
    return reqs


WORKLOAD_BUDGET = 48 << 30
WORKLOAD_TICK = 50


@pytest.fixture(scope="module")
def workload(rt):
    reqs = workload_trace()
    big = [(q["kind"], q["storage"]) for q in reqs if q["n"] >= 350]
    assert set(big) >= {(0, 0), (0, 1), (1, 0), (1, 1)}
    print("serving the workload trace", flush=True)
    return reqs, rt.serve_trace(reqs, WORKLOAD_BUDGET, WORKLOAD_TICK, rt.SERVE_BATCHED)


def test_workload_mix_same_alone_in_the_mix_and_sequential(rt, workload):
    """Every request of the served workload completes, and its result (status,
    step count, FP64 checksum of the final field) is bitwise the same in the
    mix, served alone, and in the sequential plain-run loop."""
    reqs, mixed = workload
    for q, r in zip(reqs, mixed):
        assert r["status"] == 0, q
        assert r["steps"] <= q["max_steps"] + WORKLOAD_TICK - 1
    seq = rt.serve_trace(reqs, WORKLOAD_BUDGET, WORKLOAD_TICK, rt.SERVE_SEQUENTIAL)
    for q, a, b in zip(reqs, mixed, seq):
        assert key(a) == key(b), q
    for q, r in zip(reqs, mixed):
        print("alone", q["kind"], q["n"], q["storage"], flush=True)
        alone = rt.serve_trace([dict(q, arrival=0.0)], WORKLOAD_BUDGET, WORKLOAD_TICK, rt.SERVE_BATCHED)[0]
        assert key(alone) == key(r), q


def test_workload_fp32_checksums_match_oracle(rt, workload):
    """FP32 requests of the workload whose CPU oracle run is short (up to
    3e9 point updates for diffusion, 2e8 for Jacobi) have the oracle's FP64
    checksum of the final field, to FP64 summation order."""
    reqs, mixed = workload
    checked = 0
    for q, r in zip(reqs, mixed):
        n, steps = q["n"], r["steps"]
        if q["storage"] != 0 or n ** 3 * steps > (3e9 if q["kind"] == 0 else 2e8):
            continue
        print("oracle", q["kind"], n, steps, flush=True)
        if q["kind"] == 0:
            field = O.uniform_field((n, n, n), seed=q["seed"])
            g1, _ = O.diffusion7(field, [0.1] * 6 + [0.4], periodic=False, steps=steps)
            want = float(np.sum(g1.astype(np.float64)))
        else:
            want = jacobi_checksum(q["seed"], n, steps)
        assert abs(r["checksum"] - want) <= 1e-12 * max(1.0, abs(want)), (q, r["checksum"], want)
        checked += 1
    assert checked >= 2
