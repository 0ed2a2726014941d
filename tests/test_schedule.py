"""The pass schedule of fused Himeno runs (host code, no GPU).

PSHipHimenoRunTB fuses T sweeps per HBM pass through a third buffer; the
schedule it follows (PSHipHimenoTBSchedule, the function the runtime itself
calls) must leave p0 and p1 exactly as single sweeps leave them: after n
sweeps of the reference's ping-pong (himenobmtxpa_physis.c:372-392,
p0 -> p1 -> p0 ...), iterate n sits in p0 for even n and in p1 for odd n, and
iterate n-1 in the other grid.  The checks below simulate the buffers."""
import ctypes
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "physis_amd", "lib", "libphysis_rt_hip.so")


@pytest.fixture(scope="module")
def schedule():
    if not os.path.exists(LIB):
        pytest.skip("libphysis_rt_hip.so not built")
    L = ctypes.CDLL(LIB)
    f = L.PSHipHimenoTBSchedule
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    f.restype = ctypes.c_int

    def run(sweeps, degree, max_passes=None):
        cap = sweeps + 1 if max_passes is None else max_passes
        buf = (ctypes.c_int * (4 * max(cap, 1)))()
        n = f(sweeps, degree, buf, cap)
        return None if n < 0 else [tuple(buf[4 * i:4 * i + 4]) for i in range(n)]

    return run


@pytest.mark.parametrize("degree", [1, 2, 3, 4])
def test_schedule_leaves_grids_as_single_sweeps(schedule, degree):
    for n in range(0, 90):
        passes = schedule(n, degree)
        held = {0: 0, 1: None, 2: None}  # buffer -> iterate it holds
        k = 0
        for first, t, src, dst in passes:
            assert first == k, (n, degree, passes)
            assert held[src] == k, "pass reads a buffer that does not hold iterate %d" % k
            assert dst != src
            assert t in (1, degree)
            if degree == 1:
                assert 2 not in (src, dst), "one-sweep runs need no third buffer"
            held[dst] = k + t
            k += t
        assert k == n
        if n >= 1:
            assert held[n & 1] == n and held[(n - 1) & 1] == n - 1, (n, degree, passes)
        else:
            assert held[0] == 0


@pytest.mark.parametrize("degree", [2, 3, 4])
def test_schedule_fuses_all_but_the_tail(schedule, degree):
    # Fused passes run while at least degree + 2 sweeps remain, so at most
    # degree + 1 single sweeps finish the run and at least two do (n >= 2).
    for n in range(0, 90):
        passes = schedule(n, degree)
        singles = [p for p in passes if p[1] == 1]
        fused = [p for p in passes if p[1] == degree]
        assert len(fused) == max(0, (n - 2) // degree)
        assert len(singles) == n - degree * len(fused)
        assert all(a[0] < b[0] for a, b in zip(fused, singles[:1]))  # fused first
        if n >= 2:
            assert 2 <= len(singles) <= degree + 1


def test_schedule_reports_a_short_buffer(schedule):
    assert schedule(10, 2, max_passes=2) is None
    assert schedule(0, 3, max_passes=0) == []
