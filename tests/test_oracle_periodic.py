"""Pin the periodic oracle with closed forms and invariants (CPU only)."""

import math

import numpy as np
import pytest

from oracle import periodic as P
from oracle import stepper as S
from dnsjax_amd.fields import periodic_field


def _params(n=16, system="kolmogorov", re=40.0, L=2 * math.pi, **kw):
    return P.PeriodicParams(nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, re=re, system=system, **kw)


def _grid(p):
    y = np.arange(p.ny_padded) * p.ly / p.ny_padded
    z = np.arange(p.nz_padded) * p.lz / p.nz_padded
    x = np.arange(p.nx_padded) * p.lx / p.nx_padded
    return np.meshgrid(y, z, x, indexing="ij")


def test_harmonics():
    assert list(P.complex_harmonics(8)) == [0, 1, 2, 3, -3, -2, -1]
    assert list(P.real_harmonics(8)) == [0, 1, 2, 3]


def test_pad_truncate_errors():
    a = np.zeros((7, 7, 4), complex)
    with pytest.raises(ValueError):
        P.zeropad_fft(a, 4, 0)
    with pytest.raises(ValueError):
        P.zeropad_fft(a, 11, 0)
    with pytest.raises(ValueError):
        P.truncate_fft(np.zeros((12, 7, 4), complex), 9, 0)
    with pytest.raises(ValueError):
        P.zeropad_fft(a, 12, 2)


@pytest.mark.parametrize("h", [(1, 2, 3), (-3, 1, 0), (2, -5, 7), (7, 7, 7)])
def test_single_mode_inverse_closed_form(h):
    """One stored mode a at (hy,hz,hx) maps to 2 Re(a e^{i k.x}) (or its kx=0 form)."""
    p = _params(16)
    hy, hz, hx = h
    u = np.zeros((1,) + p.spec_shape, complex)
    a = 0.3 - 0.7j
    iy = hy % (p.ny - 1) if hy >= 0 else hy + (p.ny - 1)
    iz = hz % (p.nz - 1) if hz >= 0 else hz + (p.nz - 1)
    u[0, iy, iz, hx] = a
    Y, Z, X = _grid(p)
    ph = 2 * np.pi * (hx * X / p.lx + hy * Y / p.ly + hz * Z / p.lz)
    got = P.spec_to_phys(u, p)[0]
    if hx > 0:
        ref = 2 * np.real(a * np.exp(1j * ph))
    else:  # irfft keeps only the real part of the kx=0 plane
        ref = np.real(a * np.exp(1j * ph))
    assert np.max(np.abs(got - ref)) < 1e-13


def test_roundtrip_identity():
    p = _params(16)
    u = periodic_field(16, 16, 16, p.lx, p.ly, p.lz, 0.5, seed=3)
    back = P.phys_to_spec(P.spec_to_phys(u, p), p)
    assert np.max(np.abs(back - u)) < 1e-14 * np.max(np.abs(u)) * 10


def test_aliased_product_truncated_exactly():
    """cos(a x) cos(b x) with a+b > n/2: only the (a-b) mode survives, at 1/4."""
    p = _params(16)
    Y, Z, X = _grid(p)
    a, b = 6, 5  # a + b = 11 > 7 = n/2 - 1
    f = np.cos(2 * np.pi * a * X / p.lx) * np.cos(2 * np.pi * b * X / p.lx)
    s = P.phys_to_spec(f[None], p)[0]
    ref = np.zeros(p.spec_shape, complex)
    ref[0, 0, a - b] = 0.25
    assert np.max(np.abs(s - ref)) < 1e-15


def test_parseval():
    p = _params(16)
    f = P.Fourier(p)
    u = periodic_field(16, 16, 16, p.lx, p.ly, p.lz, 1.0, seed=5)
    phys = P.spec_to_phys(u, p)
    assert abs(P.get_norm2(u, f.k_metric) - np.mean(np.sum(phys**2, axis=0))) < 1e-13


def _abc(p, A=1.0, B=0.7, C=0.4):
    Y, Z, X = _grid(p)
    k = 2 * np.pi / p.lx
    u = np.array([
        A * np.sin(k * Z) + C * np.cos(k * Y),
        B * np.sin(k * X) + A * np.cos(k * Z),
        C * np.sin(k * Y) + B * np.cos(k * X),
    ])
    return u, k


def test_beltrami_nonlinear_zero():
    p = _params(16, system="decaying-box")
    flow = P.Flow(p)
    u_phys, _ = _abc(p)
    u = P.phys_to_spec(u_phys, p)
    rhs = P.get_rhs(u, flow)
    assert np.max(np.abs(rhs)) < 1e-14


def test_laminar_fixed_point():
    for system in ("kolmogorov", "waleffe"):
        for L in (4.0, 2 * math.pi):
            p = _params(16, system=system, L=L)
            flow = P.Flow(p)
            rhs = P.get_rhs(np.zeros((3,) + p.spec_shape, complex), flow)
            assert np.max(np.abs(rhs)) < 1e-15


def test_taylor_green_nonlinear_closed_form():
    """Rotational NL of Taylor-Green from analytic u and w, transformed directly."""
    p = _params(16, system="decaying-box")
    flow = P.Flow(p)
    Y, Z, X = _grid(p)
    k = 2 * np.pi / p.lx
    sx, cx, sy, cy, sz, cz = (np.sin(k * X), np.cos(k * X), np.sin(k * Y),
                              np.cos(k * Y), np.sin(k * Z), np.cos(k * Z))
    u = np.array([sx * cy * cz, -cx * sy * cz, 0 * X])
    w = np.array([-k * cx * sy * sz, -k * sx * cy * sz, 2 * k * sx * sy * cz])
    nl_exact = P.cross(u, w)
    us = P.phys_to_spec(u, p)
    got = P.get_nonlin(us, flow)
    ref = P.phys_to_spec(nl_exact, p)
    assert np.max(np.abs(got - ref)) < 1e-15
    # and the closed form itself: u x w has the single (2k) y/x structure
    assert abs(np.max(np.abs(ref)) - 0.125 * k) < 1e-14


def test_beltrami_decay_factor_and_zero_corrections():
    p = _params(16, system="decaying-box", re=10.0, dt=0.05, implicitness=0.5)
    flow = P.Flow(p)
    Y, Z, X = _grid(p)
    kz = 2 * 2 * np.pi / p.lz
    u = P.phys_to_spec(np.array([np.sin(kz * Z), np.cos(kz * Z), 0 * Z]), p)
    pc, ic = S.make_stepper(P.get_rhs, P.predict, P.correct, P.norm)
    pred, rhs, err = pc(u, flow)
    g = (1 / p.dt - (1 - p.implicitness) * kz**2 / p.re) / (1 / p.dt + p.implicitness * kz**2 / p.re)
    assert err < 1e-15
    assert np.max(np.abs(pred - g * u)) < 1e-15
    state, counts, errs = S.run(
        u, 3, 1e-10, 10,
        lambda s: pc(s, flow), lambda a, b, c: ic(a, b, c, flow),
        lambda s: P.correct_velocity(s, flow))
    assert counts == [0, 0, 0]
    assert np.max(np.abs(state - g**3 * u)) < 1e-14


def test_generator_solenoidal_rms_hermitian():
    n, L = 16, 2 * math.pi
    u = periodic_field(n, n, n, L, L, L, 0.1, seed=0)
    p = _params(n)
    f = P.Fourier(p)
    div = P.divergence(u, f.kx, f.ky, f.kz)
    assert np.max(np.abs(div)) < 1e-15
    ep = 0.5 * P.get_norm2(u, f.k_metric)
    assert abs(math.sqrt(2 * ep / 3) - 0.1) < 1e-14
    phys = P.spec_to_phys(u, p)
    assert abs(np.sqrt(np.mean(phys**2)) - 0.1) < 1e-12
    assert np.all(u[:, 0, 0, 0] == 0)
    # Hermitian at kx = 0 means the round trip is the identity
    assert np.max(np.abs(P.phys_to_spec(phys, p) - u)) < 1e-15


def test_projection_is_idempotent():
    p = _params(16)
    flow = P.Flow(p)
    u = periodic_field(16, 16, 16, p.lx, p.ly, p.lz, 0.1, seed=1)
    rhs = P.get_rhs(u, flow)
    f = flow.fourier
    assert np.max(np.abs(P.divergence(rhs, f.kx, f.ky, f.kz))) < 1e-15
    once = P.correct_velocity(u + 0.01 * rhs, flow)
    twice = P.correct_velocity(once, flow)
    assert np.max(np.abs(once - twice)) < 1e-17


@pytest.mark.parametrize("system,L,tilt", [("kolmogorov", 2 * math.pi, 0.0), ("waleffe", 2 * math.pi, 0.0),
                                           ("kolmogorov", 4.0, 30.0), ("waleffe", 4.0, 0.0)])
def test_stats_power_balance_closed_forms(system, L, tilt):
    """Diagnostics (flows/monochromatic.py:433-505) on states whose values the
    physics fixes: the zero perturbation is the laminar flow (E = 1/4,
    I = D = F/2, E' = 0); the perturbation minus the laminar profile is the
    fluid at rest (E = I = D = 0, E' = 1/4)."""
    p = P.PeriodicParams(nx=8, ny=8, nz=8, lx=L, ly=L, lz=L, re=40.0, system=system, tilt_degree=tilt)
    flow = P.Flow(p)
    F = flow.force_amplitude
    z = np.zeros((3,) + flow.fourier.lapl.shape, dtype=np.complex128)
    lam = P.get_stats(z, flow)
    assert abs(lam["E"] - 0.25) < 1e-14 and abs(lam["E'"]) < 1e-14
    assert abs(lam["I"] - F / 2) < 1e-14 * F and abs(lam["D"] - F / 2) < 1e-12 * F
    idx, uf = P.forced_modes(flow)
    rest = z.copy()
    rest[idx] = -uf
    st = P.get_stats(rest, flow)
    assert abs(st["E"]) < 1e-14 and abs(st["I"]) < 1e-14 * F and abs(st["D"]) < 1e-12 * F
    assert abs(st["E'"] - 0.25) < 1e-14


def test_fft_workers_switch_keeps_the_transforms():
    """The host-timing switch (scipy.fft with worker threads) gives the same
    round trip and forward transform as the default numpy path, to rounding,
    and switching back restores numpy.fft."""
    p = P.PeriodicParams(nx=16, ny=16, nz=16, lx=2 * np.pi, ly=2 * np.pi, lz=2 * np.pi)
    rng = np.random.default_rng(0)
    phys = rng.standard_normal(p.phys_shape)
    ref = P.rfft3d(phys, p)
    P.fft_workers(4)
    try:
        got = P.rfft3d(phys, p)
        back = P.irfft3d(got, p)
    finally:
        P.fft_workers(None)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-15)
    np.testing.assert_allclose(back, P.irfft3d(ref, p), rtol=0, atol=1e-13)
    assert P._FFT is np.fft and P._FFT_KW == {}
