"""Pin the MLP oracle: analytic gradients against central finite differences."""

import numpy as np
import pytest

from oracle import mlp as M


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    d_in, h, d_out, B = 16, 8, 6, 5
    params = M.init(d_in, h, d_out, seed=1)
    params[1] = rng.standard_normal(h) * 0.1
    params[3] = rng.standard_normal(h) * 0.1
    params[5] = rng.standard_normal(d_out) * 0.1
    x = rng.standard_normal((B, d_in))
    t = rng.standard_normal((B, d_out))
    _, grads = M.loss_and_grads(params, x, t)
    eps = 1e-6
    for pi, (p, g) in enumerate(zip(params, grads)):
        flat = p.reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            lp, _ = M.loss_and_grads(params, x, t)
            flat[idx] = old - eps
            lm, _ = M.loss_and_grads(params, x, t)
            flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g.reshape(-1)[idx]) <= 1e-7 + 1e-6 * abs(fd), (pi, idx, fd, g.reshape(-1)[idx])


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.0 + 2**-8 + 2**-9, -3.14159, 1e-20])
    r = M.to_bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0078125
    assert abs(r[3] + 3.140625) < 1e-12
    assert abs(r[4] - 1e-20) / 1e-20 < 2**-8


def test_adam_decreases_loss():
    rng = np.random.default_rng(2)
    params = M.init(4, 8, 2, seed=3)
    x = rng.standard_normal((64, 4))
    t = np.stack([np.sin(x[:, 0]), x[:, 1] * x[:, 2]], 1)
    m = [np.zeros_like(p) for p in params]
    v = [np.zeros_like(p) for p in params]
    l0, _ = M.loss_and_grads(params, x, t)
    for step in range(1, 301):
        _, g = M.loss_and_grads(params, x, t)
        M.adam(params, g, m, v, step, 1e-2)
    l1, _ = M.loss_and_grads(params, x, t)
    assert l1 < 0.3 * l0


def test_bf16_gradients_track_full_precision():
    """Straight-through bf16 gradients stay within bf16 precision of the exact ones."""
    rng = np.random.default_rng(4)
    params = M.init(16, 32, 6, seed=5)
    x = rng.standard_normal((40, 16))
    t = rng.standard_normal((40, 6))
    _, g = M.loss_and_grads(params, x, t)
    _, gb = M.loss_and_grads(params, x, t, bf16=True)
    for a, b in zip(g, gb):
        assert np.linalg.norm(a - b) <= 3e-2 * np.linalg.norm(a)


# ---------------------------------------------------------------- corrector guess
import math  # noqa: E402

from oracle import periodic as P  # noqa: E402
from dnsjax_amd.fields import periodic_field  # noqa: E402


def _flow(n=8, re=40.0):
    L = 2 * math.pi
    return P.Flow(P.PeriodicParams(nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, re=re, dt=0.01, system="kolmogorov"))


def _f32(a):
    return a.astype(np.complex64).astype(np.complex128)


def test_guess_mode0_is_the_plain_predictor():
    flow = _flow()
    u = periodic_field(8, 8, 8, 2 * math.pi, 2 * math.pi, 2 * math.pi, 0.1, 3)
    z = np.zeros_like(u)
    pred, rhs, hist = M.guess(u, z, z, flow, 0)
    n0 = P.get_rhs(u, flow)
    assert np.array_equal(rhs, n0) and np.array_equal(hist, n0)
    assert np.array_equal(pred, P.predict(u, n0, flow))


@pytest.mark.parametrize("mode", [1, 2, 3, 4])
def test_guess_keeps_the_corrector_map(mode):
    """One corrector pass from the guess equals the plain iteration map
    u1 + c ildt_2 (N(pred) - N^n) at every mode, whatever dN is."""
    flow = _flow()
    L = 2 * math.pi
    u = periodic_field(8, 8, 8, L, L, L, 0.1, 3)
    h1 = _f32(P.get_rhs(periodic_field(8, 8, 8, L, L, L, 0.1, 4), flow))
    h2 = _f32(P.get_rhs(periodic_field(8, 8, 8, L, L, L, 0.1, 5), flow))
    h3 = _f32(P.get_rhs(periodic_field(8, 8, 8, L, L, L, 0.1, 6), flow)) if mode == 4 else None
    params = M.init(16, 32, 6, seed=1)
    pred, rhs_prev, _ = M.guess(u, h1, h2, flow, mode, params, (1.0, 2.0, 0.5), h3=h3)
    n0 = P.get_rhs(u, flow)
    u1 = P.predict(u, n0, flow)
    assert np.linalg.norm(pred - u1) > 1e-3 * np.linalg.norm(u1 - u)  # the guess moved
    nxt, _ = P.correct(u, pred, rhs_prev, P.get_rhs(pred, flow), flow)
    plain = u1 + flow.p.implicitness * (P.get_rhs(pred, flow) - n0) * flow.ildt_2
    assert np.max(np.abs(nxt - plain)) <= 1e-13 * np.max(np.abs(plain))
    # ... which is the full form ildt_2 (ldt_1 u^n + c N(pred) + (1 - c) N^n)
    c = flow.p.implicitness
    full = flow.ildt_2 * (flow.ldt_1 * u + c * P.get_rhs(pred, flow) + (1 - c) * n0)
    full[:, 0, 0, 0] = nxt[:, 0, 0, 0]  # the mean mode is held by correct_velocity
    assert np.max(np.abs(nxt - full)) <= 1e-13 * np.max(np.abs(full))


def test_extrapolation_is_exact_for_polynomial_histories():
    """A nonlinear term moving quadratically in time is extrapolated exactly
    by mode 3 (target zero), linearly by mode 1 only when linear."""
    flow = _flow()
    L = 2 * math.pi
    a = _f32(P.get_rhs(periodic_field(8, 8, 8, L, L, L, 0.1, 3), flow))
    b = _f32(P.get_rhs(periodic_field(8, 8, 8, L, L, L, 0.1, 9), flow)) * 0.25
    e = _f32(P.get_rhs(periodic_field(8, 8, 8, L, L, L, 0.1, 11), flow)) * 0.125
    seq = [a + k * b + k * k * e for k in range(4)]  # N^{n-2}, N^{n-1}, N^n, N^{n+1}
    modes = np.arange(a[0].size)
    X, T = M.features(seq[0], seq[1], seq[2], seq[3], modes, flow, 1.0, 1.0, 1.0)
    scale = np.max(np.abs(b))
    assert np.max(np.abs(T)) <= 1e-12 * scale
    # D1 = N^n - N^{n-1} = b + 3 e ; D2 = 2 e
    assert np.allclose(X[:, 0], (b[0] + 3 * e[0]).real.reshape(-1), atol=1e-6 * scale)
    assert np.allclose(X[:, 6], (2 * e[0]).real.reshape(-1), atol=1e-6 * scale)


def test_cubic_extrapolation_is_exact_for_cubic_histories():
    """With N^{n-3} given: a cubic history leaves no cubic-residual target
    and a quartic one leaves its fourth difference (24 e per unit step,
    weighted); mode 4 extrapolates a cubic history exactly (dN = N^{n+1} -
    N^n), and mode 2 with a zero network equals mode 4."""
    flow = _flow()
    L = 2 * math.pi
    f = lambda s: P.get_rhs(periodic_field(8, 8, 8, L, L, L, 0.1, s), flow)  # noqa: E731
    a, b, e, g = f(3), f(9) * 0.25, f(11) * 0.125, f(13) * 0.0625
    modes = np.arange(a[0].size)
    cub = [a + k * b + k * k * e + k**3 * g for k in range(5)]  # N^{n-3} .. N^{n+1}
    _, T = M.features(cub[1], cub[2], cub[3], cub[4], modes, flow, 1.0, 1.0, 1.0, n_m3=cub[0])
    assert np.max(np.abs(T)) <= 1e-12 * np.max(np.abs(b))
    _, Tq = M.features(cub[1], cub[2], cub[3], cub[4], modes, flow, 1.0, 1.0, 1.0)
    assert np.max(np.abs(Tq)) > 1e-3 * np.max(np.abs(g))  # quadratic extrapolation misses g
    quart = [a + k**4 * e for k in range(5)]
    _, T4 = M.features(quart[1], quart[2], quart[3], quart[4], modes, flow, 1.0, 1.0, 2.0, n_m3=quart[0])
    want = M.mode_weight(flow).reshape(-1) * (24 * e[2]).real.reshape(-1) / 2.0
    assert np.allclose(T4[:, 4], want, rtol=1e-12, atol=1e-12 * np.max(np.abs(want)))
    # the guess itself: N(t) = N^n + t b + t^2 e + t^3 g around the state's N^n
    u = periodic_field(8, 8, 8, L, L, L, 0.1, 5)
    n0 = P.get_rhs(u, flow)
    hist = [n0 - k * b + k * k * e - k**3 * g for k in (1, 2, 3)]
    _, rhs4, _ = M.guess(u, *hist[:2], flow, 4, h3=hist[2])
    want_d = b + e + g
    want_d[:, 0, 0, 0] = 0.0
    assert np.max(np.abs((rhs4 - n0) - want_d)) <= 1e-12 * np.max(np.abs(want_d))
    params = M.init(16, 32, 6, seed=1)
    params[4][:] = 0.0
    params[5][:] = 0.0
    _, rhs2, _ = M.guess(u, *hist[:2], flow, 2, params, (1.0, 1.0, 1.0), h3=hist[2])
    assert np.array_equal(rhs2, rhs4)


def test_walls_guess_mode0_and_corrector_map():
    """Walls: mode 0 is the reference predictor; one corrector pass from any
    guess is the reference's correction S(u^n; u_g, N_ext, N(u_g))."""
    from oracle import walls as W
    from dnsjax_amd.fields import wall_field

    flow = W.WallFlow(W.WallParams(nx=8, ny=17, nz=8, fd_order=4, re=400.0, dt=0.01))
    u = wall_field(8, 17, 8, flow.p.lx, flow.p.lz, 0.05, seed=3)
    n0 = W.get_rhs(u, flow)
    z = np.zeros_like(u)
    pred, rhs, hist = M.guess_walls(u, z, z, flow, 0)
    assert np.array_equal(pred, W.predict(u, n0, flow)) and np.array_equal(rhs, n0)
    assert np.array_equal(hist, n0)
    h1 = _f32(W.get_rhs(wall_field(8, 17, 8, flow.p.lx, flow.p.lz, 0.05, seed=4), flow))
    h2 = _f32(W.get_rhs(wall_field(8, 17, 8, flow.p.lx, flow.p.lz, 0.05, seed=5), flow))
    for mode in (1, 3):
        g, ext, _ = M.guess_walls(u, h1, h2, flow, mode)
        d1 = n0 - h1
        want_ext = n0 + (d1 if mode == 1 else 2 * d1 - (h1 - h2))
        assert np.allclose(ext, want_ext, rtol=0, atol=1e-13 * np.max(np.abs(n0)))
        assert np.array_equal(g, W.predict(u, ext, flow))


def _wall_flow():
    from oracle import walls as W

    return W.WallFlow(W.WallParams(nx=8, ny=17, nz=8, fd_order=4, re=400.0, dt=0.01))


def test_wall_features_on_polynomial_histories():
    """Quadratic histories leave no residual; a cubic one leaves exactly its
    third difference, 6 e per unit step, weighted by w.  The coordinate
    features read back the point's (kx, y, kz), found here by index
    arithmetic over (kz, kx, y)."""
    from dnsjax_amd.fields import wall_field
    from oracle import walls as W

    flow = _wall_flow()
    p = flow.p
    f = lambda s: _f32(W.get_rhs(wall_field(8, 17, 8, p.lx, p.lz, 0.05, seed=s), flow))  # noqa: E731
    a, b, e = f(3), f(9) * 0.25, f(11) * 0.125
    pts = np.arange(a[0].size)
    quad = [a + k * b + k * k * e for k in range(4)]
    X, T = M.features_walls(*quad, pts, flow, 1.0, 1.0, 1.0)
    scale = np.max(np.abs(b))
    assert np.max(np.abs(T)) <= 1e-12 * scale
    assert np.allclose(X[:, 0], (b[0] + 3 * e[0]).real.reshape(-1), atol=1e-6 * scale)
    assert np.allclose(X[:, 7], (2 * e[0]).imag.reshape(-1), atol=1e-6 * scale)
    cub = [a + k * k * k * e for k in range(4)]
    _, T3 = M.features_walls(*cub, pts, flow, 1.0, 1.0, 2.0)
    iz, rest = np.divmod(pts, (p.nx // 2) * p.ny)
    ix, iy = np.divmod(rest, p.ny)
    kx = 2 * np.pi / p.lx * ix
    kz = 2 * np.pi / p.lz * np.where(iz < p.nz // 2, iz, iz - (p.nz - 1))
    w = 1.0 / (1.0 + p.implicitness * p.dt * (kx**2 + kz**2) / p.re)
    want = w * (6 * e[1]).imag.reshape(-1) / 2.0
    assert np.allclose(T3[:, 3], want, rtol=1e-12, atol=1e-12 * np.max(np.abs(want)))
    kmax = max(2 * np.pi / p.lx * (p.nx // 2), 2 * np.pi / p.lz * (p.nz // 2))
    ys = -np.cos(np.pi * iy / (p.ny - 1))
    assert np.allclose(X[:, 13], kx / kmax) and np.allclose(X[:, 14], ys) and np.allclose(X[:, 15], kz / kmax)
    assert np.allclose(X[:, 12], np.minimum(p.dt * (kx**2 + kz**2) / p.re, 8.0))


def test_wall_guess_mode2_adds_the_net_residual():
    """Mode 2 = mode 3 plus s_out * net / w, with the bf16 network on the
    point features; a zero output layer reproduces mode 3 exactly."""
    from dnsjax_amd.fields import wall_field
    from oracle import walls as W

    flow = _wall_flow()
    p = flow.p
    u = wall_field(8, 17, 8, p.lx, p.lz, 0.05, seed=3)
    h1 = _f32(W.get_rhs(wall_field(8, 17, 8, p.lx, p.lz, 0.05, seed=4), flow))
    h2 = _f32(W.get_rhs(wall_field(8, 17, 8, p.lx, p.lz, 0.05, seed=5), flow))
    params = M.init(16, 32, 6, seed=1)
    zero = [q.copy() for q in params]
    zero[4][:] = 0.0
    g3, ext3, _ = M.guess_walls(u, h1, h2, flow, 3)
    g0, ext0, _ = M.guess_walls(u, h1, h2, flow, 2, zero, (1.0, 1.0, 1.0))
    assert np.array_equal(ext0, ext3) and np.array_equal(g0, g3)
    s = (2.0, 3.0, 0.5)
    g2, ext2, _ = M.guess_walls(u, h1, h2, flow, 2, params, s)
    n0 = W.get_rhs(u, flow)
    X, _ = M.features_walls(h2, h1, n0, None, np.arange(n0[0].size), flow, *s)
    Y, _ = M.forward(params, X, bf16=True)
    w = M.mode_weight_walls(flow)
    for c in range(3):
        want = ext3[c] + s[2] * (Y[:, 2 * c] + 1j * Y[:, 2 * c + 1]).reshape(w.shape) / w
        assert np.allclose(ext2[c], want, rtol=1e-14, atol=1e-14)


def test_profile_features_on_polynomial_histories():
    """The per-(kz, kx) profile network's rows: a quadratic history leaves no
    target; a cubic one leaves its third difference (6 e per unit step)
    weighted by w, in the [component][real, imaginary][y] order; the D1 and
    D2 blocks and the mode features read back what index arithmetic over
    (kz, kx) and y gives."""
    from dnsjax_amd.fields import wall_field
    from oracle import walls as W

    flow = _wall_flow()
    p = flow.p
    ny, nkx = p.ny, p.nx // 2
    din = M.profile_din(ny)
    assert din % 32 == 0 and din >= 12 * ny + 3
    f = lambda s: _f32(W.get_rhs(wall_field(8, 17, 8, p.lx, p.lz, 0.05, seed=s), flow))  # noqa: E731
    a, b, e = f(3), f(9) * 0.25, f(11) * 0.125
    modes = np.arange((p.nz - 1) * nkx)
    quad = [a + k * b + k * k * e for k in range(4)]
    X, T = M.features_walls_profile(*quad, modes, flow, 1.0, 1.0, 1.0, din)
    scale = np.max(np.abs(b))
    assert T.shape == (len(modes), 6 * ny) and X.shape == (len(modes), din)
    assert np.max(np.abs(T)) <= 1e-12 * scale
    iz, ix = np.divmod(modes, nkx)
    # D1 = b + 3 e at k = 3 (component 1, imaginary part: block 2 * 1 + 1)
    assert np.allclose(X[:, 3 * ny:4 * ny], (b[1] + 3 * e[1])[iz, ix].imag, atol=1e-6 * scale)
    # D2 = 2 e (component 2, real part, in the D2 block)
    assert np.allclose(X[:, 6 * ny + 4 * ny:6 * ny + 5 * ny], (2 * e[2])[iz, ix].real, atol=1e-6 * scale)
    cub = [a + k * k * k * e for k in range(4)]
    _, T3 = M.features_walls_profile(*cub, modes, flow, 1.0, 1.0, 2.0, din)
    kx = 2 * np.pi / p.lx * ix
    kz = 2 * np.pi / p.lz * np.where(iz < p.nz // 2, iz, iz - (p.nz - 1))
    w = 1.0 / (1.0 + p.implicitness * p.dt * (kx**2 + kz**2) / p.re)
    want = w[:, None] * (6 * e[0])[iz, ix].real / 2.0
    assert np.allclose(T3[:, 0:ny], want, rtol=1e-12, atol=1e-12 * np.max(np.abs(want)))
    kmax = max(2 * np.pi / p.lx * nkx, 2 * np.pi / p.lz * (p.nz // 2))
    assert np.allclose(X[:, 12 * ny], np.minimum(p.dt * (kx**2 + kz**2) / p.re, 8.0))
    assert np.allclose(X[:, 12 * ny + 1], kx / kmax) and np.allclose(X[:, 12 * ny + 2], kz / kmax)
    assert not X[:, 12 * ny + 3:].any()


def test_profile_loss_scales_the_pinned_gradients():
    """The profile network's loss is the mean over samples and outputs: its
    gradients are the finite-difference-pinned ``loss_and_grads`` over dout,
    checked here by a central difference on one weight of each layer."""
    rng = np.random.default_rng(4)
    params = M.init(64, 32, 24, seed=5)
    x, t = rng.standard_normal((10, 64)), rng.standard_normal((10, 24))
    loss, g = M.profile_loss_and_grads(params, x, t, bf16=False)
    for li, idx in ((0, (3, 7)), (2, (5, 1)), (4, (30, 20)), (5, (2,))):
        h = 1e-6
        pp = [q.copy() for q in params]
        pp[li][idx] += h
        lp, _ = M.profile_loss_and_grads(pp, x, t, bf16=False)
        pp[li][idx] -= 2 * h
        lm, _ = M.profile_loss_and_grads(pp, x, t, bf16=False)
        assert abs((lp - lm) / (2 * h) - g[li][idx]) <= 1e-6 * max(1.0, abs(g[li][idx]))
    assert abs(loss - np.sum((M.forward(params, x)[0] - t) ** 2) / (10 * 24)) < 1e-12
