"""Pins the CPU oracle (test infrastructure) against the reference's own
fixtures and against closed forms.  CPU only.

Each check is chosen so that a plausible restatement mistake fails it:
  - bit-exact equality with the reference's hand-written ref oracles
    (tests/system_tests/test_cases/*.manual.ref.c, via tests/golden);
  - exact mass conservation (telescoping sums) in FP64;
  - cosine / Fourier eigenmode decay with per-axis coefficients, so a
    dropped, swapped or mis-signed neighbour changes the decay rate;
  - an independent dense-matrix formulation of the clamped operator on a
    tiny grid with asymmetric coefficients (catches w/e, n/s, b/t swaps);
  - reduction closed forms (reference test_11..test_14);
  - Himeno: sine-mode decay of its 7-point Jacobi special case and
    polynomial fields that isolate each b-term and each a/c neighbour.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "manual_ref.json")))


def text_of(values):
    return "".join("%f\n" % v for v in np.asarray(values).ravel())


def sha(values):
    return hashlib.sha256(text_of(values).encode()).hexdigest()


def neumann_coefficients(n):
    """Coefficients exactly as test_7-pt-neumann-cond.c:58-71 computes them
    (float variables, double literals promote)."""
    f32 = np.float32
    l, kappa = f32(1.0), f32(0.1)
    dx = f32(l / f32(n))
    dt = f32(0.1 * np.float64(dx) * np.float64(dx) / np.float64(kappa))
    ce = f32(kappa * dt / (dx * dx))
    cn = ce
    ct = ce
    s = f32(f32(f32(f32(f32(ce + ce) + cn) + cn) + ct) + ct)
    cc = f32(1.0 - np.float64(s))
    return [ce, ce, cn, cn, ct, ct, cc]


def test_golden_fixture_set_is_complete():
    for name in ["test_7-pt", "test_7-pt-periodic", "test_7-pt-neumann-cond",
                 "test_7-pt-multi-iterations", "test_7-pt-double-type"]:
        assert name in GOLDEN and GOLDEN[name]["lines"] == 32 ** 3


def test_diffusion_clamped_matches_reference_manual_oracle():
    n = 32
    field = np.arange(n ** 3, dtype=np.float32).reshape(n, n, n)
    g1, _ = O.diffusion7(field, neumann_coefficients(n), periodic=False, steps=10)
    assert sha(g1) == GOLDEN["test_7-pt-neumann-cond"]["sha256"]


def test_diffusion_periodic_matches_reference_manual_oracle():
    n = 32
    field = np.arange(n ** 3, dtype=np.float32).reshape(n, n, n)
    _, g2 = O.diffusion7(field, [1.0] * 7, periodic=True, steps=1)
    assert sha(g2) == GOLDEN["test_7-pt-periodic"]["sha256"]


def test_sha_detects_a_planted_mistake():
    # A mistake planted in the oracle's inputs (one neighbour coefficient
    # dropped) must change the fixture hash.
    n = 32
    field = np.arange(n ** 3, dtype=np.float32).reshape(n, n, n)
    k = neumann_coefficients(n)
    k[4] = np.float32(0.0)
    g1, _ = O.diffusion7(field, k, periodic=False, steps=10)
    assert sha(g1) != GOLDEN["test_7-pt-neumann-cond"]["sha256"]


@pytest.mark.parametrize("periodic", [False, True])
def test_mass_conservation_fp64(periodic):
    field = O.uniform_field((13, 11, 9), seed=7).astype(np.float64)
    k = [0.1, 0.1, 0.05, 0.05, 0.12, 0.12, 0.46]
    g1, _ = O.diffusion7(field, k, periodic=periodic, steps=20, dtype=np.float64)
    assert abs(g1.sum() - field.sum()) < 1e-10 * field.sum()


def test_clamped_cosine_mode_decay_fp64():
    nx, ny, nz = 12, 10, 8
    kx, ky, kz = 2, 1, 3
    cw = ce = 0.1
    cn = cs = 0.05
    cb = ct = 0.12
    cc = 1 - 2 * (0.1 + 0.05 + 0.12)
    x = np.cos(np.pi * kx * (np.arange(nx) + 0.5) / nx)
    y = np.cos(np.pi * ky * (np.arange(ny) + 0.5) / ny)
    z = np.cos(np.pi * kz * (np.arange(nz) + 0.5) / nz)
    field = z[:, None, None] * y[None, :, None] * x[None, None, :]
    lam = cc + 2 * ce * np.cos(np.pi * kx / nx) + 2 * cn * np.cos(np.pi * ky / ny) + 2 * ct * np.cos(np.pi * kz / nz)
    steps = 12
    g1, _ = O.diffusion7(field, [ce, cw, cn, cs, ct, cb, cc], periodic=False, steps=steps, dtype=np.float64)
    np.testing.assert_allclose(g1, lam ** steps * field, atol=1e-12)


def test_periodic_fourier_mode_fp64():
    nx, ny, nz = 9, 8, 6
    kx, ky, kz = 2, 3, 1
    k = [0.11, 0.11, 0.07, 0.07, 0.13, 0.13, 1 - 2 * (0.11 + 0.07 + 0.13)]
    tx, ty, tz = 2 * np.pi * kx / nx, 2 * np.pi * ky / ny, 2 * np.pi * kz / nz
    X, Y, Z = np.arange(nx), np.arange(ny), np.arange(nz)
    field = np.cos(tz * Z[:, None, None] + ty * Y[None, :, None] + tx * X[None, None, :] + 0.3)
    lam = k[6] + 2 * k[0] * np.cos(tx) + 2 * k[2] * np.cos(ty) + 2 * k[4] * np.cos(tz)
    g1, _ = O.diffusion7(field, k, periodic=True, steps=10, dtype=np.float64)
    np.testing.assert_allclose(g1, lam ** 10 * field, atol=1e-12)


def clamped_matrix(n, cminus, cplus):
    """1-D operator of one axis: neighbour j-1 weighted cminus, j+1 cplus,
    reflecting (the out-of-grid neighbour is the point itself)."""
    A = np.zeros((n, n))
    for j in range(n):
        A[j, max(j - 1, 0)] += cminus
        A[j, min(j + 1, n - 1)] += cplus
    return A


def test_clamped_asymmetric_vs_dense_operator():
    nx, ny, nz = 4, 3, 5
    ce, cw, cn, cs, ct, cb, cc = 0.11, 0.07, 0.05, 0.13, 0.02, 0.17, 0.45
    Ax = clamped_matrix(nx, cw, ce)
    Ay = clamped_matrix(ny, cn, cs)
    Az = clamped_matrix(nz, cb, ct)
    Ix, Iy, Iz = np.eye(nx), np.eye(ny), np.eye(nz)
    # Flattened index = x + nx*(y + ny*z): kron order z, y, x.
    A = (cc * np.kron(Iz, np.kron(Iy, Ix)) + np.kron(Iz, np.kron(Iy, Ax)) + np.kron(Iz, np.kron(Ay, Ix))
         + np.kron(Az, np.kron(Iy, Ix)))
    field = O.uniform_field((nx, ny, nz), seed=3).astype(np.float64)
    steps = 5
    want = np.linalg.matrix_power(A, steps) @ field.ravel()
    g1, g2 = O.diffusion7(field, [ce, cw, cn, cs, ct, cb, cc], periodic=False, steps=steps, dtype=np.float64)
    got = g2 if steps % 2 else g1
    np.testing.assert_allclose(got.ravel(), want, atol=1e-13)


def test_subdomain_leaves_outside_untouched():
    field = O.uniform_field((10, 9, 8), seed=5)
    dom = (2, 7, 1, 8, 3, 6)
    g1, g2 = O.diffusion7(field, [0.1] * 6 + [0.4], periodic=False, steps=1, dom=dom)
    mask = np.zeros(field.shape, bool)
    mask[3:6, 1:8, 2:7] = True
    assert np.all(g2[~mask] == 0.0)
    assert np.all(g1 == field)


def test_reductions_closed_forms():
    # test_11: PS_SUM of 0..511 (N=8) is exact in float.
    a = np.arange(8 ** 3, dtype=np.float32)
    assert O.reduce(a, O.PS_SUM) == 130816.0
    # test_12: PS_MAX of 0..63.
    assert O.reduce(np.arange(64, dtype=np.float32), O.PS_MAX) == 63.0
    # test_14: PS_MIN of i+10.
    assert O.reduce(np.arange(512, dtype=np.float32) + 10, O.PS_MIN) == 10.0
    # test_13: PS_PROD of 64 x 1.1 folded in float.
    v = np.float32(1.1)
    acc = v
    for _ in range(63):
        acc = np.float32(acc * v)
    assert O.reduce(np.full(64, 1.1, np.float32), O.PS_PROD) == acc
    assert abs(acc - 1.1 ** 64) / 1.1 ** 64 < 1e-5
    with pytest.raises(ValueError):
        O.reduce(a, 7)


def test_uniform_field_is_seeded_and_in_range():
    a = O.uniform_field((16, 8, 4), seed=1)
    b = O.uniform_field((16, 8, 4), seed=1)
    c = O.uniform_field((16, 8, 4), seed=2)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= 0.0 and a.max() < 1.0
    assert abs(a.mean() - 0.5) < 0.05
    # multiples of 2^-24
    assert np.all((a * 2 ** 24) == np.round(a * 2 ** 24))


def himeno_coefs(shape, vals):
    return [np.full(tuple(reversed(shape)), v, np.float64) for v in vals]


def test_himeno_jacobi_special_case_sine_decay():
    nx, ny, nz = 17, 13, 11
    kx, ky, kz = 2, 1, 3
    X = np.sin(np.pi * kx * np.arange(nx) / (nx - 1))
    Y = np.sin(np.pi * ky * np.arange(ny) / (ny - 1))
    Z = np.sin(np.pi * kz * np.arange(nz) / (nz - 1))
    p = Z[:, None, None] * Y[None, :, None] * X[None, None, :]
    coefs = himeno_coefs((nx, ny, nz), [1, 1, 1, 1 / 6, 0, 0, 0, 1, 1, 1, 1, 0])
    omega = 0.8
    mu = 1 - omega + omega / 3 * (np.cos(np.pi * kx / (nx - 1)) + np.cos(np.pi * ky / (ny - 1))
                                 + np.cos(np.pi * kz / (nz - 1)))
    sweeps = 6
    p0, p1, gosa = O.himeno(p, coefs, omega, sweeps, dtype=np.float64)
    np.testing.assert_allclose(p0, mu ** sweeps * p, atol=1e-12)
    # ss = (mu - 1)/omega * p at sweep s (before the update): closed-form gosa.
    inner = p[1:-1, 1:-1, 1:-1]
    for s in range(sweeps):
        want = ((mu - 1) / omega) ** 2 * (mu ** s) ** 2 * float((inner ** 2).sum())
        assert abs(gosa[s] - want) <= 1e-10 * max(want, 1e-30)


@pytest.mark.parametrize("field,term", [("xy", 5), ("yz", 4), ("xz", 6)])
def test_himeno_polynomial_isolates_each_term(field, term):
    nx, ny, nz = 6, 5, 4
    X, Y, Z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="xy")
    X = np.transpose(X, (2, 0, 1)).astype(np.float64)  # [z, y, x]
    Y = np.transpose(Y, (2, 0, 1)).astype(np.float64)
    Z = np.transpose(Z, (2, 0, 1)).astype(np.float64)
    p = {"xy": X * Y, "yz": Y * Z, "xz": X * Z}[field]
    vals = [2.0, 3.0, 5.0, 0.25, 7.0, 11.0, 13.0, 17.0, 19.0, 23.0, 1.0, 0.5]
    a0, a1, a2, a3, b0, b1, b2, c0, c1, c2, bnd, wrk1 = vals
    coefs = himeno_coefs((nx, ny, nz), vals)
    omega = 0.8
    p0, p1, _ = O.himeno(p, coefs, omega, 1, dtype=np.float64)

    def P(dx, dy, dz):
        x, y, z = X + dx, Y + dy, Z + dz
        return {"xy": x * y, "yz": y * z, "xz": x * z}[field]

    cross = {"xy": b1 * 4.0, "yz": b0 * 4.0, "xz": b2 * 4.0}[field]
    s0 = (a0 * P(0, 0, 1) + a1 * P(0, 1, 0) + a2 * P(1, 0, 0) + cross
          + c0 * P(0, 0, -1) + c1 * P(0, -1, 0) + c2 * P(-1, 0, 0) + wrk1)
    ss = (s0 * a3 - p) * bnd
    want = p + omega * ss
    np.testing.assert_allclose(p1[1:-1, 1:-1, 1:-1], want[1:-1, 1:-1, 1:-1], atol=1e-12)
    # Boundary cells are outside the interior domain and keep their values.
    assert np.array_equal(p1[0], p[0]) and np.array_equal(p1[:, :, -1], p[:, :, -1])


def periodic_matrix(n, cminus, cplus):
    """1-D periodic operator of one axis: neighbour j-1 weighted cminus, j+1
    cplus, indices wrapped (i + n) % n."""
    A = np.zeros((n, n))
    for j in range(n):
        A[j, (j - 1) % n] += cminus
        A[j, (j + 1) % n] += cplus
    return A


def test_periodic_asymmetric_vs_dense_operator():
    """Periodic branch with six different neighbour coefficients: a w/e, n/s
    or b/t swap, or a wrong wrap, changes the result (the Fourier-mode test
    uses symmetric coefficients and would not see a swap).  The FP32 term
    order is the clamped branch's (one shared expression), pinned bit for bit
    by the reference's Neumann-condition fixture."""
    nx, ny, nz = 5, 4, 3
    ce, cw, cn, cs, ct, cb, cc = 0.11, 0.07, 0.05, 0.13, 0.02, 0.17, 0.45
    Ax = periodic_matrix(nx, cw, ce)
    Ay = periodic_matrix(ny, cn, cs)
    Az = periodic_matrix(nz, cb, ct)
    Ix, Iy, Iz = np.eye(nx), np.eye(ny), np.eye(nz)
    A = (cc * np.kron(Iz, np.kron(Iy, Ix)) + np.kron(Iz, np.kron(Iy, Ax)) + np.kron(Iz, np.kron(Ay, Ix))
         + np.kron(Az, np.kron(Iy, Ix)))
    field = O.uniform_field((nx, ny, nz), seed=4).astype(np.float64)
    for steps in (1, 4):
        want = np.linalg.matrix_power(A, steps) @ field.ravel()
        g1, g2 = O.diffusion7(field, [ce, cw, cn, cs, ct, cb, cc], periodic=True, steps=steps, dtype=np.float64)
        got = g2 if steps % 2 else g1
        np.testing.assert_allclose(got.ravel(), want, atol=1e-13)


HIMENO_ORIGINAL = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "himeno_original_jacobi.exe")


@pytest.mark.parametrize("shape,sweeps", [((17, 9, 7), 1), ((17, 9, 7), 2), ((33, 12, 10), 5), ((5, 4, 3), 3)])
def test_himeno_oracle_equals_reference_original_jacobi(tmp_path, shape, sweeps):
    """The Himeno restatement (of himenobmtxpa_physis.c:331-361) against the
    benchmark's original C solver, jacobi() of himenobmtxpa_original.c:299-346,
    built from the reference sources (oracle/Makefile ref) and run on the
    same random inputs: the pressure field is bitwise equal after every
    sweep count.  The original's indices (i, j, k), k fastest, are z, y, x
    here, and its terms come in the port's order under that permutation, so
    the FP32 roundings agree one for one; its residual is a float sum in
    loop order, the oracle's an FP64 sum."""
    import subprocess
    if not os.path.exists(HIMENO_ORIGINAL):
        pytest.skip("the reference's Himeno original is not built (make -C oracle ref)")
    nx, ny, nz = shape
    p = O.uniform_field(shape, seed=90, lo=-1.0, hi=1.0)
    coefs = []
    for i in range(12):
        if i == 3:
            c = O.uniform_field(shape, seed=91 + i, lo=0.15, hi=0.18)
        elif i == 10:
            c = (O.uniform_field(shape, seed=91 + i) < 0.8).astype(np.float32)  # bnd: 0 or 1
        elif i == 11:
            c = O.uniform_field(shape, seed=91 + i, lo=-0.1, hi=0.1)
        else:
            c = O.uniform_field(shape, seed=91 + i, lo=0.9, hi=1.1)
        coefs.append(c)
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    np.concatenate([a.ravel() for a in [p] + coefs]).astype(np.float32).tofile(src)
    subprocess.run([HIMENO_ORIGINAL, str(nx), str(ny), str(nz), str(sweeps), str(src), str(dst)], check=True,
                   timeout=60)
    out = np.fromfile(dst, np.float32)
    got_p, got_gosa = out[:-1].reshape(nz, ny, nx), float(out[-1])
    p0, p1, gosa = O.himeno(p, coefs, 0.8, sweeps)
    want = p1 if sweeps % 2 else p0
    assert np.array_equal(got_p, want)
    assert abs(got_gosa - gosa[-1]) <= 1e-4 * abs(gosa[-1])
