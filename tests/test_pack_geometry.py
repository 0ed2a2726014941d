"""The packed-row tiles' indexing, checked on the CPU for every row length.

The kernels (include/physis/hip/skeleton_pack.h) place lane L of a workgroup
at row L // V, vector L % V (V = ceil(nx/4) vectors per row, the last one
aligned to the row's end), publish a plane slot-major to LDS (point j of
every vector in slot array j at the vector's position; u0 rows one slot
below the two halo rows) and read each point's y neighbours V positions away
and its x neighbours across the vector's ends at per-lane (slot, offset)
words.  PSHipPackGeometry returns exactly the geometry the kernels compute
(the same host/device functions); this test simulates the LDS plane with
those places and asserts, for every nx in [4, 512], both boundary forms and
both workgroup sizes, that every lane reads the neighbour the stencil means
(at a clamped row end the point itself, at a periodic one the row's other
end), that no LDS word receives two different points, and that the lanes
cover every point of their rows."""
import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "physis_amd", "lib", "libphysis_rt_hip.so")


@pytest.fixture(scope="module")
def geometry():
    if not os.path.exists(LIB):
        pytest.skip("libphysis_rt_hip.so not built")
    L = ctypes.CDLL(LIB)
    f = L.PSHipPackGeometry
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                  ctypes.POINTER(ctypes.c_int)]
    f.restype = ctypes.c_int

    def get(nx, nw, periodic):
        lanes = np.zeros(8 * 64 * nw, dtype=np.int32)
        halo = np.zeros(4 * nx, dtype=np.int32)
        rows = f(nx, nw, periodic, lanes.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                 halo.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return rows, lanes.reshape(-1, 8), halo.reshape(-1, 4)

    return get


def check(nx, nw, periodic, rows, lanes, halo):
    V = (nx + 3) // 4
    R = rows
    assert R == 64 * nw // V and R * V <= 64 * nw
    npos = (R + 2) * V
    lds = np.full((4, npos), -1, dtype=np.int64)

    def tag(r, x):  # row r in [-1, R], point x
        return (np.asarray(r) + 1) * 4096 + np.asarray(x)

    def put(slot, pos, t):
        cur = lds[slot, pos]
        clash = (cur != -1) & (cur != t)
        assert not np.any(clash), "LDS word written with two points (nx %d)" % nx
        lds[slot, pos] = t

    x = np.arange(nx)
    for row, base in ((-1, 0), (R, (R + 1) * V)):
        put(halo[:, 0], halo[:, 1] + base, tag(row, x))
        put(halo[:, 2], halo[:, 3] + base, tag(row, x))
    r, vi, xs, shadow, wj, wd, ej, ed = (lanes[:, k] for k in range(8))
    assert np.all((xs >= 0) & (xs + 3 < nx)), "vector outside its row"
    pos = (r + 1) * V + vi
    for j in range(4):
        put(np.full_like(pos, j), pos, tag(r, xs + j))
    live = shadow == 0
    # Coverage: every point of the R rows belongs to a live lane.
    cov = np.zeros((R, nx), dtype=bool)
    for j in range(4):
        cov[r[live], xs[live] + j] = True
    assert cov.all(), "points not covered (nx %d)" % nx
    r, vi, xs, wj, wd, ej, ed, pos = (a[live] for a in (r, vi, xs, wj, wd, ej, ed, pos))
    for j in range(4):
        px = xs + j
        assert np.all(lds[j, pos - V] == tag(r - 1, px)), "north neighbour (nx %d)" % nx
        assert np.all(lds[j, pos + V] == tag(r + 1, px)), "south neighbour (nx %d)" % nx
    # x neighbours across the vector's ends.
    w_want = np.where(xs > 0, xs - 1, nx - 1 if periodic else 0)
    assert np.all(lds[wj, pos + wd] == tag(r, w_want)), "west neighbour (nx %d)" % nx
    e_x = xs + 3
    e_want = np.where(e_x < nx - 1, e_x + 1, 0 if periodic else nx - 1)
    assert np.all(lds[ej, pos + ed] == tag(r, e_want)), "east neighbour (nx %d)" % nx
    # The LDS arrays the kernels declare hold every position used.
    nxmax = 256 if nx <= 256 else 512
    ks0 = (64 * nw + 2 * (nxmax // 4) + 63) // 64 * 64
    assert npos <= ks0


@pytest.mark.parametrize("periodic", [0, 1])
@pytest.mark.parametrize("nw,nx_max", [(8, 256), (16, 512)])
def test_every_row_length_reads_the_right_neighbours(geometry, periodic, nw, nx_max):
    for nx in range(4, nx_max + 1):
        rows, lanes, halo = geometry(nx, nw, periodic)
        if 64 * nw // ((nx + 3) // 4) < 3:
            continue
        check(nx, nw, periodic, rows, lanes, halo)


def test_planted_index_mistake_is_caught(geometry):
    """The check itself: an east-neighbour slot off by one on one lane (the
    kind of mistake the overlap of a row's last two vectors invites) fails."""
    nx = 139
    rows, lanes, halo = geometry(nx, 8, 0)
    V = (nx + 3) // 4
    bad = lanes.copy()
    bad[V - 2, 6] = (bad[V - 2, 6] + 1) % 4  # second-to-last vector's east slot
    with pytest.raises(AssertionError):
        check(nx, 8, 0, rows, bad, halo)
