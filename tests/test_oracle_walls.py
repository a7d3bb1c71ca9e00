"""Pin the wall-bounded oracle (CPU only).

The first five tests restate the reference's own known-answer tests
(``tests/test_banded_solver.py:68-197``) with the same seeds, sizes and
tolerances, applied to the oracle's functions.
"""

import math

import numpy as np
import pytest
from numpy.testing import assert_allclose
from scipy.linalg.lapack import get_lapack_funcs

from oracle import walls as W
from dnsjax_amd.fields import wall_field, wall_quadrature_weights


def _make_random_banded(Ny, kl, ku, seed):
    """``tests/test_banded_solver.py:55-65``."""
    rng = np.random.default_rng(seed)
    A = np.zeros((Ny, Ny))
    for i in range(Ny):
        j_lo, j_hi = max(0, i - kl), min(Ny, i + ku + 1)
        A[i, j_lo:j_hi] = rng.standard_normal(j_hi - j_lo)
    return A + 10.0 * np.eye(Ny)


def test_banded_solve_random_real():
    """``tests/test_banded_solver.py:68-82``."""
    Ny, kl, ku = 16, 4, 4
    A = _make_random_banded(Ny, kl, ku, seed=0)
    b = np.random.default_rng(10).standard_normal(Ny)
    ab, ipiv = W.banded_lu(A, kl, ku)
    assert_allclose(W.banded_solve(ab, ipiv, b, kl, ku), np.linalg.solve(A, b), atol=1e-10, rtol=1e-10)


def test_banded_operator_complex_rhs():
    """``tests/test_banded_solver.py:85-100``."""
    Ny, kl, ku = 16, 4, 4
    A = _make_random_banded(Ny, kl, ku, seed=1)
    rng = np.random.default_rng(11)
    b = rng.standard_normal(Ny) + 1j * rng.standard_normal(Ny)
    ab, ipiv = W.banded_lu(A, kl, ku)
    x = W.banded_solve_modes(ab[None, None], ipiv[None, None], b[None, None])[0, 0]
    assert_allclose(x, np.linalg.solve(A, b), atol=1e-10, rtol=1e-10)


MODES = [(0.0, 0.0), (0.0, 1.7), (2.0, 3.0)]


def test_banded_vs_dense_on_imm_operators():
    """``tests/test_banded_solver.py:103-145``."""
    Ny, p = 17, 4
    y = W.cgl_grid(Ny)
    D1, D2 = W.build_diff_matrices(y, p)
    dt, c, nu = 0.01, 0.5, 1.0 / 1000.0
    rng = np.random.default_rng(20)
    for kz, kx in MODES:
        k2 = kx**2 + kz**2
        Lk = W.build_Lk_neumann(k2, D1, D2)
        Hk, _ = W.build_Hk_dirichlet(k2, D2, dt, c, nu)
        b = rng.standard_normal(Ny) + 1j * rng.standard_normal(Ny)
        for M in (Lk, Hk):
            ab, piv = W.banded_lu(M, p, p)
            xb = W.banded_solve_modes(ab[None, None], piv[None, None], b[None, None])[0, 0]
            assert_allclose(xb, np.linalg.solve(M, b), atol=1e-9, rtol=1e-9)


def test_lk_matvec_matches_dense():
    """``tests/test_banded_solver.py:148-173``."""
    Ny, p = 17, 4
    y = W.cgl_grid(Ny)
    D1, D2 = W.build_diff_matrices(y, p)
    rng = np.random.default_rng(30)
    for kz, kx in MODES:
        k2v = kx**2 + kz**2
        u = rng.standard_normal(Ny) + 1j * rng.standard_normal(Ny)
        k2 = np.array([[[k2v]]])
        got = W.lk_matvec(u[None, None], D2, D1[[0, -1]], k2, k2 == 0.0)[0, 0]
        assert_allclose(got, W.build_Lk_neumann(k2v, D1, D2) @ u, atol=1e-10, rtol=1e-10)


def test_hk_minus_matvec_matches_dense():
    """``tests/test_banded_solver.py:176-197``."""
    Ny, p = 17, 4
    dt, c, nu = 0.01, 0.5, 1.0 / 1000.0
    y = W.cgl_grid(Ny)
    _, D2 = W.build_diff_matrices(y, p)
    rng = np.random.default_rng(40)
    for kz, kx in MODES:
        k2v = kx**2 + kz**2
        u = rng.standard_normal(Ny) + 1j * rng.standard_normal(Ny)
        got = W.hk_minus_matvec(u[None, None], D2, np.array([[[k2v]]]), dt, c, nu)[0, 0]
        assert_allclose(got, W.build_Hk_dirichlet(k2v, D2, dt, c, nu)[1] @ u, atol=1e-10, rtol=1e-10)


# ------------------------------------------------------------ beyond the reference
def test_banded_solve_matches_scipy_dgbtrs():
    Ny, kl, ku = 33, 6, 6
    A = _make_random_banded(Ny, kl, ku, seed=7)
    b = np.random.default_rng(8).standard_normal(Ny)
    ab, ipiv = W.banded_lu(A, kl, ku)
    (gbtrs,) = get_lapack_funcs(("gbtrs",), (ab,))
    ref, info = gbtrs(ab, kl, ku, b, ipiv)
    assert info == 0
    assert_allclose(W.banded_solve(ab, ipiv, b, kl, ku), ref, rtol=1e-13, atol=1e-13)


def test_fornberg_textbook_uniform():
    w = W.fornberg_weights(0.0, np.array([-1.0, 0.0, 1.0]), 2)
    assert_allclose(w[:, 1], [-0.5, 0.0, 0.5], atol=1e-15)
    assert_allclose(w[:, 2], [1.0, -2.0, 1.0], atol=1e-15)
    w = W.fornberg_weights(0.0, np.array([-2.0, -1.0, 0.0, 1.0, 2.0]), 1)
    assert_allclose(w[:, 1], [1 / 12, -2 / 3, 0, 2 / 3, -1 / 12], atol=1e-15)


@pytest.mark.parametrize("p", [2, 4, 6])
def test_fd_polynomial_exactness_on_cgl(p):
    y = W.cgl_grid(33)
    D1, D2 = W.build_diff_matrices(y, p)
    for d in range(p + 1):
        f = y**d
        df = d * y ** max(d - 1, 0) if d else 0 * y
        d2f = d * (d - 1) * y ** max(d - 2, 0) if d > 1 else 0 * y
        assert np.max(np.abs(D1 @ f - df)) < 1e-9
        assert np.max(np.abs(D2 @ f - d2f)) < 1e-7


@pytest.mark.parametrize("ny", [17, 33, 16, 32, 129, 128])
def test_quadrature_exact_for_quadratics(ny):
    """Non-uniform Simpson panels integrate quadratics exactly (odd cubics too when Ny is odd)."""
    y = W.cgl_grid(ny)
    for d in range(4 if ny % 2 else 3):
        exact = (1 - (-1) ** (d + 1)) / (d + 1)
        assert abs(W.integrate_scalar_in_y(y**d, y) - exact) < 1e-13
    w = wall_quadrature_weights(y)
    f = np.cos(3 * y) + y**2
    assert abs(np.dot(w, f) - W.integrate_scalar_in_y(f, y)) < 1e-13


def _small_flow(ny=17, nx=8, nz=8, backend="banded", **kw):
    p = W.WallParams(nx=nx, ny=ny, nz=nz, **kw)
    return W.WallFlow(p, backend=backend)


def test_imm_invariants_no_slip_and_wall_divergence():
    flow = _small_flow()
    p = flow.p
    rng = np.random.default_rng(0)
    shp = (3,) + p.spec_shape
    un = wall_field(p.nx, p.ny, p.nz, p.lx, p.lz, 0.05, seed=1)
    Nn = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    Nj = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    uj = un + 0.01 * (rng.standard_normal(shp) + 1j * rng.standard_normal(shp))
    vel, corr = W.imm_iteration(un, uj, Nn, Nj, flow)
    assert np.max(np.abs(vel[..., 0])) < 1e-14 and np.max(np.abs(vel[..., -1])) < 1e-14
    div = 1j * flow.kx * vel[0] + W.dy(flow.D1, vel[1]) + 1j * flow.kz * vel[2]
    scale = np.max(np.abs(vel))
    top = div[..., -1]
    bot = np.where(flow.k2_is_zero[..., 0], 0.0, div[..., 0])
    assert np.max(np.abs(top)) < 1e-10 * scale * 100
    assert np.max(np.abs(bot)) < 1e-10 * scale * 100
    assert np.max(np.abs(corr - (vel - uj))) == 0


def test_imm_banded_equals_dense():
    fb = _small_flow(backend="banded")
    fd = _small_flow(backend="dense")
    p = fb.p
    rng = np.random.default_rng(5)
    shp = (3,) + p.spec_shape
    un = wall_field(p.nx, p.ny, p.nz, p.lx, p.lz, 0.05, seed=2)
    Nn = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    vb, _ = W.imm_iteration(un, un, Nn, Nn, fb)
    vd, _ = W.imm_iteration(un, un, Nn, Nn, fd)
    assert np.max(np.abs(vb - vd)) < 1e-11 * np.max(np.abs(vd))


def test_laminar_couette_is_fixed_point():
    flow = _small_flow()
    zero = np.zeros((3,) + flow.p.spec_shape, complex)
    rhs = W.get_rhs(zero, flow)
    # only U x curl U = (0, y, 0) in the mean mode survives
    assert np.max(np.abs(rhs[1, 0, 0] - flow.ys)) < 1e-14
    rhs2 = rhs.copy()
    rhs2[1, 0, 0] = 0
    assert np.max(np.abs(rhs2)) < 1e-15
    pred = W.predict(zero, rhs, flow)
    assert np.max(np.abs(pred)) < 1e-11


def test_wall_generator_no_slip_and_rms():
    nx, ny, nz = 8, 17, 8
    u = wall_field(nx, ny, nz, 2 * math.pi, math.pi, 0.05, seed=2)
    assert u.shape == (3, nz - 1, nx // 2, ny)
    assert np.all(u[..., 0] == 0) and np.all(u[..., -1] == 0)
    flow = _small_flow(ny=ny, nx=nx, nz=nz)
    ep = W.get_norm2(u, flow) / 2
    assert abs(math.sqrt(2 * ep / 3) - 0.05) < 1e-14
    # Hermitian in (kz) at kx = 0: the 2D round trip is the identity
    phys = W.spec_to_phys_2d(u, flow.p)
    assert np.max(np.abs(W.phys_to_spec_2d(phys, flow.p) - u)) < 1e-15


def test_batched_banded_solve_is_bitwise_the_scalar_solve():
    """The oracle's batched solve does the scalar dgbtrs-order solve's
    operations per right-hand side: identical bits."""
    rng = np.random.default_rng(3)
    for Ny, p in ((17, 4), (33, 6), (20, 2)):
        y = W.cgl_grid(Ny)
        D1, D2 = W.build_diff_matrices(y, p)
        # Neumann L_k (its wall rows pivot) and random diagonally dominant bands
        mats = [W.build_Lk_neumann(k2, D1, D2) for k2 in (0.0, 1.0, 7.5)]
        mats += [_make_random_banded(Ny, p, p, seed=s) for s in range(2)]
        facs = [W.banded_lu(A, p, p) for A in mats]
        ab = np.stack([f[0] for f in facs])
        piv = np.stack([f[1] for f in facs])
        assert any(np.any(pv != np.arange(Ny)) for pv in piv)  # some rows swap
        b = rng.standard_normal((5, Ny))
        got = W.banded_solve_batch(ab, piv, b, p, p)
        for i in range(5):
            assert np.array_equal(got[i], W.banded_solve(ab[i], piv[i], b[i], p, p))


def test_setup_tables_shared_per_k2_equal_the_per_mode_loop():
    """precompute_imm computes each table once per distinct k^2: at a box
    where +-kz share k^2, the copies equal a fresh per-mode computation."""
    flow = _small_flow(ny=17, nx=8, nz=8, lx=3.0, lz=1.7)
    imm = flow.imm
    for iz, ix in ((1, 2), (6, 2), (3, 1)):
        k2 = float(flow.k2[iz, ix, 0])
        Lk = W.build_Lk_neumann(k2, flow.D1, flow.D2)
        e1 = np.zeros(17)
        e1[0] = 1.0
        assert np.array_equal(imm["p1"][iz, ix], np.linalg.solve(Lk, e1))
        ab, piv = W.banded_lu(Lk, 4, 4)
        assert np.array_equal(imm["Lk_ab"][iz, ix], ab) and np.array_equal(imm["Lk_piv"][iz, ix], piv)
