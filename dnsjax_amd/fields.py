"""Seeded synthetic initial fields (the reference has no random initial state).

The reference starts from the laminar state or from a snapshot
(``flows/monochromatic.py:294-314``, ``flows/plane_couette.py:225-248``);
a random generator is listed as a to-do in its README.  These generators
replace ``init_state`` for the benchmark workloads.  They are host-side
set-up code (NumPy, deterministic for a given seed), not part of the
per-step path.

Periodic field, spectral layout ``(3, ny-1, nz-1, nx//2)`` [ky, kz, kx]:
  complex Gaussian 3-vectors per stored mode (real parts drawn first, then
  imaginary parts, each as one ``standard_normal`` call over the whole
  array), scaled by sqrt(E(k) / (4 pi k^2)) with E(k) = k^4 exp(-2 (k/4)^2)
  and k the integer-harmonic magnitude; Leray-projected with the physical
  wavenumbers; made Hermitian at kx = 0 by a round trip through an
  unpadded real 3D transform; mean mode zeroed; scaled so that
  sqrt(2 E' / 3) equals ``rms`` with E' = 1/2 sum k_metric |u|^2.

Wall-bounded field, spectral layout ``(3, nz-1, nx//2, ny)`` [kz, kx, y]:
  u = curl(psi), psi_c = (1 - y^2)^2 phi_c(y) per horizontal mode, with
  phi_c a Chebyshev series T_0..T_7 whose coefficients are complex
  Gaussians times exp(-m/4), the horizontal amplitude sqrt(E(k)/(2 pi k)),
  y-derivatives analytic; Hermitian at kx = 0 by a 2D round trip; scaled
  to ``rms`` with the y-quadrature norm.  u vanishes with its y-derivative
  at both walls.
"""

from __future__ import annotations

import math

import numpy as np


def _harm_complex(n: int) -> np.ndarray:
    q = (np.arange(n) + n // 2) % n - n // 2
    return np.concatenate([q[: n // 2], q[n // 2 + 1 :]])


def _spectrum(k: np.ndarray) -> np.ndarray:
    return k**4 * np.exp(-2.0 * (k / 4.0) ** 2)


def _pad_full_axis(a: np.ndarray, axis: int) -> np.ndarray:
    """Insert the (zero) Nyquist entry: (n-1) stored modes -> n FFT bins."""
    n1 = a.shape[axis]
    n = n1 + 1
    shape = list(a.shape)
    shape[axis] = n
    out = np.zeros(shape, dtype=a.dtype)
    src = [slice(None)] * a.ndim
    dst = [slice(None)] * a.ndim
    src[axis] = slice(0, n // 2)
    dst[axis] = slice(0, n // 2)
    out[tuple(dst)] = a[tuple(src)]
    src[axis] = slice(n // 2, None)
    dst[axis] = slice(n // 2 + 1, None)
    out[tuple(dst)] = a[tuple(src)]
    return out


def _drop_nyquist_axis(a: np.ndarray, axis: int) -> np.ndarray:
    n = a.shape[axis]
    keep = np.r_[0 : n // 2, n // 2 + 1 : n]
    return np.take(a, keep, axis=axis)


def hermitianize_periodic(u: np.ndarray, nx: int, ny: int, nz: int) -> np.ndarray:
    """Round trip through the unpadded real 3D transform (per component)."""
    out = np.empty_like(u)
    for c in range(u.shape[0]):
        full = _pad_full_axis(_pad_full_axis(u[c], 0), 1)
        full = np.concatenate([full, np.zeros(full.shape[:2] + (1,), full.dtype)], axis=2)
        phys = np.fft.irfftn(full, s=(ny, nz, nx), axes=(0, 1, 2))
        back = np.fft.rfftn(phys, axes=(0, 1, 2))[:, :, : nx // 2]
        out[c] = _drop_nyquist_axis(_drop_nyquist_axis(back, 0), 1)
    return out


def periodic_field(nx: int, ny: int, nz: int, lx: float, ly: float, lz: float,
                   rms: float, seed: int) -> np.ndarray:
    """Seeded solenoidal periodic perturbation, complex128 (3, ny-1, nz-1, nx//2)."""
    rng = np.random.default_rng(seed)
    shape = (3, ny - 1, nz - 1, nx // 2)
    u = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    hx = np.arange(nx // 2).reshape(1, 1, -1)
    hy = _harm_complex(ny).reshape(-1, 1, 1)
    hz = _harm_complex(nz).reshape(1, -1, 1)
    k = np.sqrt(hx**2 + hy**2 + hz**2 + 0.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.where(k > 0, np.sqrt(_spectrum(k) / (4 * np.pi * k**2)), 0.0)
    u = u * amp
    kx = hx * (2 * np.pi / lx)
    ky = hy * (2 * np.pi / ly)
    kz = hz * (2 * np.pi / lz)
    k2 = kx**2 + ky**2 + kz**2
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = np.where(k2 > 0, 1.0 / k2, 0.0)
    kdotu = kx * u[0] + ky * u[1] + kz * u[2]
    proj = kdotu * inv
    u = u - np.stack([kx * proj, ky * proj, kz * proj])
    u = hermitianize_periodic(u, nx, ny, nz)
    u[:, 0, 0, 0] = 0
    km = np.where(hx == 0, 1.0, 2.0)
    ep = 0.5 * float(np.sum(km * np.abs(u) ** 2))
    if ep > 0:
        u = u * (rms / math.sqrt(2 * ep / 3))
    return np.ascontiguousarray(u)


# --------------------------------------------------------------- walls
def _cheb_series(coef: np.ndarray, y: np.ndarray):
    """Values and y-derivatives of sum_m coef[..., m] T_m(y)."""
    from numpy.polynomial import chebyshev as C

    m = coef.shape[-1]
    flat = coef.reshape(-1, m)
    val = np.array([C.chebval(y, c) for c in flat]).reshape(coef.shape[:-1] + (len(y),))
    der = np.array([C.chebval(y, C.chebder(c)) for c in flat]).reshape(coef.shape[:-1] + (len(y),))
    return val, der


def wall_quadrature_weights(ys: np.ndarray) -> np.ndarray:
    """Linear weights of the y-quadrature used by the wall norm.

    Odd Ny: the non-uniform composite Simpson rule of the reference
    (``geometries/cartesian.py:104-140``).  Even Ny (no reference
    counterpart): Simpson panels over the first Ny-4 intervals plus the
    exact integral of the cubic through the last four points.
    """
    n = len(ys)
    w = np.zeros(n)
    m = n if n % 2 == 1 else n - 3
    h = np.diff(ys[:m])
    for i in range(0, m - 2, 2):
        h0, h1 = h[i], h[i + 1]
        hs, hp, r = h0 + h1, h0 * h1, h0 / h1
        w[i] += hs / 6 * (2 - 1 / r)
        w[i + 1] += hs / 6 * (hs**2 / hp)
        w[i + 2] += hs / 6 * (2 - r)
    if n % 2 == 0:
        x = ys[n - 4 :]
        a, b = x[0], x[3]
        V = np.vander(x - a, 4, increasing=True).T
        mom = np.array([(b - a) ** (j + 1) / (j + 1) for j in range(4)])
        w[n - 4 :] += np.linalg.solve(V, mom)
    return w


def wall_field(nx: int, ny: int, nz: int, lx: float, lz: float, rms: float,
               seed: int, n_cheb: int = 8) -> np.ndarray:
    """Seeded no-slip perturbation, complex128 (3, nz-1, nx//2, ny)."""
    rng = np.random.default_rng(seed)
    nkz, nkx = nz - 1, nx // 2
    ys = -np.cos(np.arange(ny) * np.pi / (ny - 1))
    shape = (3, nkz, nkx, n_cheb)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    a = a * np.exp(-np.arange(n_cheb) / 4.0)
    hx = np.arange(nkx).reshape(1, -1)
    hz = _harm_complex(nz).reshape(-1, 1)
    k = np.sqrt(hx**2 + hz**2 + 0.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.where(k > 0, np.sqrt(_spectrum(k) / (2 * np.pi * k)), 0.0)
    a = a * amp[None, :, :, None]
    phi, dphi = _cheb_series(a, ys)
    env = (1 - ys**2) ** 2
    denv = -4 * ys * (1 - ys**2)
    psi = env * phi
    dpsi = denv * phi + env * dphi
    kx = (hx * (2 * np.pi / lx))[:, :, None]
    kz = (hz * (2 * np.pi / lz))[:, :, None]
    u = np.empty((3, nkz, nkx, ny), dtype=np.complex128)
    u[0] = dpsi[2] - 1j * kz * psi[1]
    u[1] = 1j * kz * psi[0] - 1j * kx * psi[2]
    u[2] = 1j * kx * psi[1] - dpsi[0]
    # Hermitian at kx = 0: 2D round trip per (component, y).
    for c in range(3):
        full = _pad_full_axis(u[c], 0)  # (nz, nkx, ny)
        full = np.concatenate([full, np.zeros((nz, 1, ny), full.dtype)], axis=1)
        phys = np.fft.irfftn(full, s=(nz, nx), axes=(0, 1))
        back = np.fft.rfftn(phys, axes=(0, 1))[:, :nkx, :]
        u[c] = _drop_nyquist_axis(back, 0)
    u[:, :, :, 0] = 0
    u[:, :, :, -1] = 0
    w = wall_quadrature_weights(ys)
    km = np.where(hx == 0, 1.0, 2.0)[:, :, None]
    prof = np.sum(km * np.abs(u) ** 2, axis=(0, 1, 2))
    ep = 0.5 * float(np.dot(w, prof)) / 2.0
    if ep > 0:
        u = u * (rms / math.sqrt(2 * ep / 3))
    return np.ascontiguousarray(u)
