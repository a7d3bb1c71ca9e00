"""Oracle: plane Couette (wall-bounded) step, FP64.

Test infrastructure only (see ``oracle/__init__.py``).

Restates, in NumPy/SciPy:
  * Fornberg weights and the banded D1/D2 matrices (``fd.py:21-108``);
  * the Cartesian Fourier tables, y-integrated norm and non-uniform
    Simpson rule (``geometries/cartesian.py:17-140``), extended to even Ny
    (documented below; the reference exits there,
    ``geometries/cartesian.py:118-123``);
  * LAPACK ``ab`` packing, ``dgbtrf`` factorisation and the
    ``dgbtrs``-equivalent banded solve, split re/im
    (``geometries/cartesian.py:291-488``);
  * ``build_Lk_neumann``, ``build_Hk_dirichlet``, ``precompute_imm``
    (``geometries/cartesian.py:491-801``);
  * the plane Couette flow: base flow, FD curl, 2D dealiased transforms,
    matrix-free ``Lk``/``Hk^-`` matvecs and the six-stage influence-matrix
    iteration, predictor/corrector/norm (``flows/plane_couette.py:64-549``,
    ``operators.py:104-146``, ``fft.py:193-288``).

Pinned by (tests/test_oracle_walls.py): the reference's five known-answer
banded-solver tests restated with their seeds and tolerances
(``tests/test_banded_solver.py:68-197``), SciPy ``dgbtrs`` and
``np.linalg.solve``; Fornberg polynomial exactness and the uniform-grid
textbook stencils; IMM
invariants (no-slip walls, zero discrete wall divergence); laminar fixed
point (RHS == 0 at u' = 0 once projected by the IMM); quadrature exact for
quadratics on the CGL grid.

Even Ny quadrature (no reference counterpart, "parity unpinned" for the
even-Ny iteration counts): non-uniform Simpson panels over the first Ny-4
intervals, and the exact integral of the cubic interpolant through the last
four points over the last three intervals.  Both pieces are exact for
quadratics (the reference's odd-Ny rule also is, and no more on unequal
panels).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
from scipy.linalg.lapack import get_lapack_funcs

from .periodic import (
    complex_harmonics,
    cross,
    real_harmonics,
    truncate_fft,
    truncate_rfft,
    zeropad_fft,
    zeropad_rfft,
)


# ------------------------------------------------------------------ fd.py
def fornberg_weights(z: float, x: np.ndarray, m: int) -> np.ndarray:
    """``fd.py:21-65`` (Fornberg 1998, SIAM Rev. 40, 685-691)."""
    n = len(x) - 1
    C = np.zeros((n + 1, m + 1))
    C[0, 0] = 1.0
    c1 = 1.0
    c4 = x[0] - z
    for i in range(1, n + 1):
        mn = min(i, m)
        c2 = 1.0
        c5 = c4
        c4 = x[i] - z
        for j in range(i):
            c3 = x[i] - x[j]
            c2 *= c3
            if j == i - 1:
                for k in range(mn, 0, -1):
                    C[i, k] = c1 * (k * C[i - 1, k - 1] - c5 * C[i - 1, k]) / c2
                C[i, 0] = -c1 * c5 * C[i - 1, 0] / c2
            for k in range(mn, 0, -1):
                C[j, k] = (c4 * C[j, k] - k * C[j, k - 1]) / c3
            C[j, 0] = c4 * C[j, 0] / c3
        c1 = c2
    return C


def build_diff_matrices(y: np.ndarray, p: int):
    """``fd.py:68-108``: D1 with (p+1)-point, D2 with (p+2)-point stencils."""
    Ny = len(y)
    s1, s2 = p + 1, p + 2
    h1, h2 = s1 // 2, s2 // 2
    D1 = np.zeros((Ny, Ny))
    D2 = np.zeros((Ny, Ny))
    for i in range(Ny):
        j0 = max(0, min(i - h1, Ny - s1))
        D1[i, j0 : j0 + s1] = fornberg_weights(y[i], y[j0 : j0 + s1], 1)[:, 1]
        j0 = max(0, min(i - h2, Ny - s2))
        D2[i, j0 : j0 + s2] = fornberg_weights(y[i], y[j0 : j0 + s2], 2)[:, 2]
    return D1, D2


def cgl_grid(ny: int) -> np.ndarray:
    """``flows/plane_couette.py:73-77``."""
    return -np.cos(np.arange(ny, dtype=np.float64) * np.pi / (ny - 1))


# ----------------------------------------------------------- quadrature
def simpson_panels(f: np.ndarray, ys: np.ndarray) -> float:
    """``geometries/cartesian.py:104-140`` (odd number of points)."""
    if len(ys) % 2 == 0:
        raise ValueError("Simpson panels need an odd number of points.")
    h = np.diff(ys)
    h0, h1 = h[:-1:2], h[1::2]
    y0, y1, y2 = f[:-2:2], f[1:-1:2], f[2::2]
    hsum, hprod, r = h0 + h1, h0 * h1, h0 / h1
    panels = (hsum / 6) * (y0 * (2 - 1 / r) + y1 * (hsum**2 / hprod) + y2 * (2 - r))
    return float(np.sum(panels))


def cubic_end_weights(x: np.ndarray) -> np.ndarray:
    """Weights of the exact integral over [x0, x3] of the cubic through 4 points."""
    a, b = x[0], x[3]
    V = np.vander(x - a, 4, increasing=True).T  # V[m, i] = (x_i - a)^m
    mom = np.array([(b - a) ** (m + 1) / (m + 1) for m in range(4)])
    return np.linalg.solve(V, mom)


def integrate_scalar_in_y(f: np.ndarray, ys: np.ndarray) -> float:
    """Reference rule for odd Ny; the documented extension for even Ny."""
    if len(ys) % 2 == 1:
        return simpson_panels(f, ys)
    n = len(ys)
    head = simpson_panels(f[: n - 3], ys[: n - 3])
    w = cubic_end_weights(ys[n - 4 :])
    return head + float(np.dot(w, f[n - 4 :]))


def quadrature_weights(ys: np.ndarray) -> np.ndarray:
    """The linear weights w with integrate_scalar_in_y(f) == sum(w * f)."""
    n = len(ys)
    return np.array([integrate_scalar_in_y(np.eye(n)[i], ys) for i in range(n)])


# --------------------------------------------------------------- banded LU
def pack_banded(A: np.ndarray, kl: int, ku: int) -> np.ndarray:
    """``geometries/cartesian.py:291-320``."""
    Ny = A.shape[0]
    ab = np.zeros((2 * kl + ku + 1, Ny), dtype=A.dtype)
    for j in range(Ny):
        for i in range(max(0, j - ku), min(Ny - 1, j + kl) + 1):
            ab[kl + ku + i - j, j] = A[i, j]
    return ab


def banded_lu(A: np.ndarray, kl: int, ku: int):
    """``geometries/cartesian.py:323-357`` (SciPy dgbtrf, 0-based pivots)."""
    ab = pack_banded(A, kl, ku)
    (gbtrf,) = get_lapack_funcs(("gbtrf",), (ab,))
    lu, piv, info = gbtrf(ab, kl, ku, overwrite_ab=True)
    if info != 0:
        raise ValueError(f"dgbtrf failed with info={info}")
    return lu, piv.astype(np.int32)


def banded_solve(ab: np.ndarray, ipiv: np.ndarray, b: np.ndarray, kl: int, ku: int):
    """``geometries/cartesian.py:360-441``: dgbtrs('N') from packed factors."""
    Ny = ab.shape[-1]
    kuu = kl + ku
    w = np.concatenate([np.asarray(b, dtype=np.float64), np.zeros(kl)])
    for j in range(Ny - 1):
        pj = ipiv[j]
        bj, bpj = w[j], w[pj]
        w[j], w[pj] = bpj, bj
        w[j + 1 : j + 1 + kl] = w[j + 1 : j + 1 + kl] - ab[kl + ku + 1 : kl + ku + 1 + kl, j] * bpj
    w = np.concatenate([np.zeros(kuu), w[:Ny]])
    for j in range(Ny - 1, -1, -1):
        jp = j + kuu
        xj = w[jp] / ab[kl + ku, j]
        w[jp] = xj
        w[jp - kuu : jp] = w[jp - kuu : jp] - ab[0:kuu, j] * xj
    return w[kuu:]


def banded_solve_batch(ab: np.ndarray, ipiv: np.ndarray, b: np.ndarray, kl: int, ku: int):
    """``banded_solve`` for a batch of real right-hand sides at once.

    ab (..., 2kl+ku+1, Ny), ipiv (..., Ny), b (..., Ny) or (..., Ny, K) for K
    right-hand sides per matrix: the same operations per right-hand side, in
    the same order, as ``banded_solve`` (so the results are bit-identical to
    it); vectorised over the batch for speed."""
    Ny = ab.shape[-1]
    lead = ab.shape[:-2]
    M = int(np.prod(lead, dtype=np.int64))
    A = np.ascontiguousarray(ab, dtype=np.float64).reshape(M, ab.shape[-2], Ny)
    piv = np.asarray(ipiv).reshape(M, Ny)
    bb = np.asarray(b, dtype=np.float64)
    multi = bb.ndim == len(lead) + 2
    K = bb.shape[-1] if multi else 1
    kuu = kl + ku
    w = np.zeros((M, Ny + kl, K))
    w[:, :Ny] = bb.reshape(M, Ny, K)
    rows = np.arange(M)
    for j in range(Ny - 1):
        pj = piv[:, j]
        bj = w[:, j].copy()
        bpj = w[rows, pj].copy()
        w[:, j] = bpj
        w[rows, pj] = bj
        w[:, j + 1 : j + 1 + kl] = w[:, j + 1 : j + 1 + kl] - A[:, kl + ku + 1 : kl + ku + 1 + kl, j, None] * bpj[:, None]
    v = np.zeros((M, kuu + Ny, K))
    v[:, kuu:] = w[:, :Ny]
    for j in range(Ny - 1, -1, -1):
        jp = j + kuu
        xj = v[:, jp] / A[:, kl + ku, j, None]
        v[:, jp] = xj
        v[:, jp - kuu : jp] = v[:, jp - kuu : jp] - A[:, 0:kuu, j, None] * xj[:, None]
    out = v[:, kuu:]
    return out.reshape(lead + (Ny, K)) if multi else out.reshape(lead + (Ny,))


def banded_solve_modes(ab: np.ndarray, ipiv: np.ndarray, rhs: np.ndarray) -> np.ndarray:
    """``geometries/cartesian.py:463-488``: per-mode solve, complex split re/im."""
    p = (ab.shape[-2] - 1) // 3
    if np.iscomplexobj(rhs):
        x = banded_solve_batch(ab, ipiv, np.stack([rhs.real, rhs.imag], axis=-1), p, p)
        return x[..., 0] + 1j * x[..., 1]
    return banded_solve_batch(ab, ipiv, rhs, p, p)


def build_Lk_neumann(k2: float, D1: np.ndarray, D2: np.ndarray) -> np.ndarray:
    """``geometries/cartesian.py:491-526``."""
    Ny = D2.shape[0]
    L = D2 - k2 * np.eye(Ny)
    if k2 == 0.0:
        L[0, :] = 0.0
        L[0, 0] = 1.0
        L[-1, :] = D1[-1, :]
    else:
        L[0, :] = D1[0, :]
        L[-1, :] = D1[-1, :]
    return L


def build_Hk_dirichlet(k2: float, D2: np.ndarray, dt: float, c: float, nu: float):
    """``geometries/cartesian.py:529-572``."""
    Ny = D2.shape[0]
    Lraw = D2 - k2 * np.eye(Ny)
    Hk = (1.0 / dt) * np.eye(Ny) - c * nu * Lraw
    Hm = (1.0 / dt) * np.eye(Ny) + (1.0 - c) * nu * Lraw
    for H in (Hk, Hm):
        H[0, :] = 0.0
        H[0, 0] = 1.0
        H[-1, :] = 0.0
        H[-1, -1] = 1.0
    return Hk, Hm


def precompute_imm(y, kx_vals, kz_vals, p, dt, c, nu, D1, D2, backend="banded"):
    """``geometries/cartesian.py:575-801``."""
    Ny, Nkz, Nkx = len(y), len(kz_vals), len(kx_vals)
    out = {k: np.zeros((Nkz, Nkx, Ny)) for k in ("p1", "p2", "v1", "v2", "q1", "q2")}
    out["M_inv"] = np.zeros((Nkz, Nkx, 2, 2))
    kl = ku = p
    if backend == "banded":
        out["Lk_ab"] = np.zeros((Nkz, Nkx, 3 * p + 1, Ny))
        out["Hk_ab"] = np.zeros((Nkz, Nkx, 3 * p + 1, Ny))
        out["Lk_piv"] = np.zeros((Nkz, Nkx, Ny), dtype=np.int32)
        out["Hk_piv"] = np.zeros((Nkz, Nkx, Ny), dtype=np.int32)
    else:
        for k in ("Lk", "Hk", "Hk_minus"):
            out[k] = np.zeros((Nkz, Nkx, Ny, Ny))
    e1 = np.zeros(Ny)
    e1[0] = 1.0
    e2 = np.zeros(Ny)
    e2[-1] = 1.0
    # every per-mode table is a function of k^2 alone (for fixed D1, D2, dt,
    # c, nu): it is computed at the first mode of each k^2 and copied, which
    # gives the same arrays as the reference's loop, faster
    done: dict = {}
    for iz, kz in enumerate(kz_vals):
        for ix, kx in enumerate(kx_vals):
            k2 = float(kx**2 + kz**2)
            if k2 in done:
                jz, jx = done[k2]
                for k in out:
                    if k not in ("D1", "D2"):
                        out[k][iz, ix] = out[k][jz, jx]
                continue
            done[k2] = (iz, ix)
            Lk = build_Lk_neumann(k2, D1, D2)
            Hk, Hm = build_Hk_dirichlet(k2, D2, dt, c, nu)
            p1 = np.linalg.solve(Lk, e1)
            p2 = np.linalg.solve(Lk, e2)
            r1, r2 = -D1 @ p1, -D1 @ p2
            r1[0] = r1[-1] = r2[0] = r2[-1] = 0.0
            v1, v2 = np.linalg.solve(Hk, r1), np.linalg.solve(Hk, r2)
            q1r, q2r = p1.copy(), p2.copy()
            q1r[0] = q1r[-1] = q2r[0] = q2r[-1] = 0.0
            q1, q2 = np.linalg.solve(Hk, q1r), np.linalg.solve(Hk, q2r)
            M = np.array([[D1[0] @ v1, D1[0] @ v2], [D1[-1] @ v1, D1[-1] @ v2]])
            Minv = np.array([[0.0, 0.0], [0.0, 1.0 / M[1, 1]]]) if k2 == 0.0 else np.linalg.inv(M)
            for k, v in (("p1", p1), ("p2", p2), ("v1", v1), ("v2", v2), ("q1", q1), ("q2", q2), ("M_inv", Minv)):
                out[k][iz, ix] = v
            if backend == "banded":
                out["Lk_ab"][iz, ix], out["Lk_piv"][iz, ix] = banded_lu(Lk, kl, ku)
                out["Hk_ab"][iz, ix], out["Hk_piv"][iz, ix] = banded_lu(Hk, kl, ku)
            else:
                out["Lk"][iz, ix], out["Hk"][iz, ix], out["Hk_minus"][iz, ix] = Lk, Hk, Hm
    out["D1"], out["D2"] = D1, D2
    return out


# ------------------------------------------------------------ flow config
@dataclass
class WallParams:
    """The subset of ``parameters.py`` the plane Couette step reads."""

    nx: int = 8
    ny: int = 17
    nz: int = 8
    lx: float = 2 * math.pi
    lz: float = math.pi
    re: float = 400.0
    fd_order: int = 4
    oversampling_factor: int = 3
    dt: float = 0.01
    implicitness: float = 0.5
    corrector_tolerance: float = 1e-10
    max_corrector_iterations: int = 10
    ly: float = 2.0  # parameters.py:190-191

    @property
    def nx_padded(self) -> int:
        return self.oversampling_factor * self.nx // 2

    @property
    def nz_padded(self) -> int:
        return self.oversampling_factor * self.nz // 2

    @property
    def spec_shape(self) -> tuple:  # sharding.py:423-427
        return (self.nz - 1, self.nx // 2, self.ny)

    @property
    def phys_shape(self) -> tuple:  # sharding.py:430-434
        return (self.ny, self.nz_padded, self.nx_padded)


def irfft2d(x: np.ndarray, p: WallParams) -> np.ndarray:
    """``fft.py:241-288``: (ny, nkz, nkx) -> (ny, Pz, Px)."""
    y = zeropad_fft(x, p.nz_padded, 1)
    y = np.fft.ifft(y, axis=1, norm="forward")
    y = zeropad_rfft(y, p.nx_padded // 2 + 1)
    return np.fft.irfft(y, axis=2, norm="forward")


def rfft2d(x: np.ndarray, p: WallParams) -> np.ndarray:
    """``fft.py:193-238``: (ny, Pz, Px) -> (ny, nkz, nkx)."""
    y = truncate_rfft(np.fft.rfft(x, axis=2, norm="forward"), p.nx // 2)
    return truncate_fft(np.fft.fft(y, axis=1, norm="forward"), p.nz, 1)


def spec_to_phys_2d(v: np.ndarray, p: WallParams) -> np.ndarray:
    """``operators.py:127-146``."""
    return np.stack([irfft2d(np.transpose(v[i], (2, 0, 1)), p) for i in range(v.shape[0])])


def phys_to_spec_2d(v: np.ndarray, p: WallParams) -> np.ndarray:
    """``operators.py:104-124``."""
    return np.stack([np.transpose(rfft2d(v[i], p), (1, 2, 0)) for i in range(v.shape[0])])


@dataclass
class WallFlow:
    """``flows/plane_couette.py:35-219`` + ``geometries/cartesian.py:17-66``."""

    p: WallParams
    backend: str = "banded"
    ys: np.ndarray = field(init=False)

    def __post_init__(self) -> None:
        p = self.p
        self.ys = cgl_grid(p.ny)
        self.base_flow = np.zeros((3, p.ny))
        self.base_flow[0] = self.ys
        self.curl_base_flow = np.zeros((3, p.ny))
        self.curl_base_flow[2] = -np.ones(p.ny)
        self.nonlin_base_flow = np.zeros((3, p.ny))
        self.nonlin_base_flow[1] = self.ys * np.ones(p.ny)
        self.base_flow = self.base_flow[:, :, None, None]
        self.curl_base_flow = self.curl_base_flow[:, :, None, None]
        self.nonlin_base_flow = self.nonlin_base_flow[:, :, None, None]
        self.D1, self.D2 = build_diff_matrices(self.ys, p.fd_order)
        self.D1_bnd = self.D1[[0, -1], :]
        self.kx_global = real_harmonics(p.nx) * 2 * np.pi / p.lx
        self.kz_global = complex_harmonics(p.nz) * 2 * np.pi / p.lz
        self.kx = self.kx_global.reshape(1, -1, 1)
        self.kz = self.kz_global.reshape(-1, 1, 1)
        self.k_metric = np.where(self.kx == 0, 1, 2)
        self.k2 = self.kx**2 + self.kz**2
        self.k2_is_zero = self.k2 == 0.0
        self.imm = precompute_imm(
            self.ys, self.kx_global, self.kz_global, p.fd_order, p.dt,
            p.implicitness, 1.0 / p.re, self.D1, self.D2, self.backend,
        )

    # ---------------------------------------------------------------- ops
    def solve_L(self, rhs):
        if self.backend == "banded":
            return banded_solve_modes(self.imm["Lk_ab"], self.imm["Lk_piv"], rhs)
        return np.linalg.solve(self.imm["Lk"], rhs[..., None])[..., 0]

    def solve_H(self, rhs):
        if self.backend == "banded":
            return banded_solve_modes(self.imm["Hk_ab"], self.imm["Hk_piv"], rhs)
        return np.linalg.solve(self.imm["Hk"], rhs[..., None])[..., 0]


def dy(D: np.ndarray, u: np.ndarray) -> np.ndarray:
    """``einsum("ij, zxj -> zxi", D, u)`` (``flows/plane_couette.py:259-260``),
    as a matrix product over the real and imaginary parts (the reference's
    summation order is XLA's and is not reproduced by either form)."""
    DT = np.ascontiguousarray(D.T)
    n = u.shape[-1]
    if np.iscomplexobj(u):
        re = np.ascontiguousarray(u.real).reshape(-1, n) @ DT
        im = np.ascontiguousarray(u.imag).reshape(-1, n) @ DT
        return (re + 1j * im).reshape(u.shape)
    return (np.ascontiguousarray(u).reshape(-1, n) @ DT).reshape(u.shape)


def curl_walls(state: np.ndarray, flow: WallFlow) -> np.ndarray:
    """``flows/plane_couette.py:251-271``."""
    u, v, w = state
    return np.array(
        [
            dy(flow.D1, w) - 1j * flow.kz * v,
            1j * flow.kz * u - 1j * flow.kx * w,
            1j * flow.kx * v - dy(flow.D1, u),
        ]
    )


def get_rhs(state: np.ndarray, flow: WallFlow) -> np.ndarray:
    """``flows/plane_couette.py:274-289`` via ``rhs.py:34-92``."""
    p = flow.p
    vel = spec_to_phys_2d(state, p)
    vort = spec_to_phys_2d(curl_walls(state, flow), p)
    nl = cross(vel, vort)
    nl = nl + ((cross(vel, flow.curl_base_flow) + cross(flow.base_flow, vort)) + flow.nonlin_base_flow)
    return phys_to_spec_2d(nl, p)


def lk_matvec(u, D2, D1_bnd, k2, k2_is_zero):
    """``flows/plane_couette.py:299-338``."""
    out = dy(D2, u) - k2 * u
    bot = np.where(k2_is_zero[..., 0], u[..., 0], np.einsum("j,zxj->zx", D1_bnd[0], u))
    top = np.einsum("j,zxj->zx", D1_bnd[-1], u)
    out = out.copy()
    out[..., 0] = bot
    out[..., -1] = top
    return out


def hk_minus_matvec(u, D2, k2, dt, c, nu):
    """``flows/plane_couette.py:341-373``."""
    out = (1.0 / dt) * u + (1.0 - c) * nu * (dy(D2, u) - k2 * u)
    out = out.copy()
    out[..., 0] = u[..., 0]
    out[..., -1] = u[..., -1]
    return out


def imm_iteration(velocity_n, velocity_j, nonlin_n, nonlin_j, flow: WallFlow):
    """``flows/plane_couette.py:376-504``: Kleiser-Schumann influence matrix."""
    p = flow.p
    c, dt, nu = p.implicitness, p.dt, 1.0 / p.re
    kx, kz, k2 = flow.kx, flow.kz, flow.k2
    u_n, v_n, w_n = velocity_n
    Nu_n, Nv_n, Nw_n = nonlin_n
    Nu_j, Nv_j, Nw_j = nonlin_j
    d_hat_n = 1j * kx * u_n + dy(flow.D1, v_n) + 1j * kz * w_n
    div_Nj = 1j * kx * Nu_j + dy(flow.D1, Nv_j) + 1j * kz * Nw_j
    div_Nn = 1j * kx * Nu_n + dy(flow.D1, Nv_n) + 1j * kz * Nw_n
    if flow.backend == "banded":
        Lk_d = lk_matvec(d_hat_n, flow.D2, flow.D1_bnd, k2, flow.k2_is_zero)
    else:
        Lk_d = np.einsum("zxij,zxj->zxi", flow.imm["Lk"], d_hat_n)
    f_hat = d_hat_n / dt + c * div_Nj + (1 - c) * div_Nn + (1 - c) * nu * Lk_d
    f_hat = f_hat.copy()
    f_hat[..., 0] = 0.0
    f_hat[..., -1] = 0.0
    pP = flow.solve_L(f_hat)
    dx_pP, dy_pP, dz_pP = 1j * kx * pP, dy(flow.D1, pP), 1j * kz * pP
    if flow.backend == "banded":
        Hu, Hv, Hw = (hk_minus_matvec(a, flow.D2, k2, dt, c, nu) for a in (u_n, v_n, w_n))
    else:
        Hu, Hv, Hw = (np.einsum("zxij,zxj->zxi", flow.imm["Hk_minus"], a) for a in (u_n, v_n, w_n))
    R = [
        Hu - dx_pP + c * Nu_j + (1 - c) * Nu_n,
        Hv - dy_pP + c * Nv_j + (1 - c) * Nv_n,
        Hw - dz_pP + c * Nw_j + (1 - c) * Nw_n,
    ]
    for r in R:
        r[..., 0] = 0.0
        r[..., -1] = 0.0
    u_arb, v_arb, w_arb = (flow.solve_H(r) for r in R)
    d_wall = np.einsum("bj,zxj->zxb", flow.D1_bnd, v_arb)
    d_wall[..., 0] = np.where(flow.k2_is_zero[..., 0], 0.0, d_wall[..., 0])
    alpha = -np.einsum("zxab,zxb->zxa", flow.imm["M_inv"], d_wall)
    a1, a2 = alpha[..., 0][..., None], alpha[..., 1][..., None]
    v_new = v_arb + a1 * flow.imm["v1"] + a2 * flow.imm["v2"]
    q_new = a1 * flow.imm["q1"] + a2 * flow.imm["q2"]
    u_new = u_arb - 1j * kx * q_new
    w_new = w_arb - 1j * kz * q_new
    vel = np.array([u_new, v_new, w_new])
    return vel, vel - velocity_j


def predict(state, rhs, flow: WallFlow):
    """``flows/plane_couette.py:507-519``."""
    return imm_iteration(state, state, rhs, rhs, flow)[0]


def correct(state_prev, prediction, rhs_prev, rhs_next, flow: WallFlow):
    """``flows/plane_couette.py:522-540`` (N_n := rhs_prev, as in the reference)."""
    return imm_iteration(state_prev, prediction, rhs_prev, rhs_next, flow)


def get_norm2(a: np.ndarray, flow: WallFlow) -> float:
    """``geometries/cartesian.py:72-96``: Simpson in y of the mode sum, / Ly."""
    prof = np.sum((np.conj(a) * flow.k_metric * a).real, axis=(0, 1, 2))
    return integrate_scalar_in_y(prof, flow.ys) / flow.p.ly


def norm(correction: np.ndarray, flow: WallFlow) -> float:
    """``flows/plane_couette.py:543-549``."""
    return math.sqrt(get_norm2(correction, flow))


def correct_velocity(state, flow: WallFlow):
    """``flows/plane_couette.py:575-582``: pass-through."""
    return state


def get_stats(state, flow: WallFlow) -> dict:
    """``flows/plane_couette.py:591-613``: E' only."""
    return {"E'": get_norm2(state, flow) / 2}
