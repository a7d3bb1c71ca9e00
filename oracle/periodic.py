"""Oracle: triply periodic (Kolmogorov / Waleffe / decaying-box) step, FP64.

Test infrastructure only (see ``oracle/__init__.py``).

Restates, in NumPy:
  * pad/truncate rules and axis order of the dealiased 3D real FFT
    (``fft.py:41-190,291-408``), ``norm="forward"`` on every axis;
  * harmonics and spectral operators (``operators.py:18-56``,
    ``geometries/triply_periodic.py:16-148``);
  * base flow, ``ldt_1``/``ildt_2`` tables (``flows/monochromatic.py:113-285``);
  * the rotational nonlinear term (``rhs.py:34-92``, ``operators.py:149-173``);
  * RHS with Leray projection, predictor, corrector, norm
    (``flows/monochromatic.py:320-410``);
  * divergence correction + mean-mode zeroing
    (``flows/monochromatic.py:516-565``);
  * diagnostics E, I, D, E' (``flows/monochromatic.py:433-505``).

Pinned by (tests/test_oracle_periodic.py): single-mode closed forms of the
padded transforms, spec->phys->spec identity, exact truncation of aliased
products, Parseval, Beltrami (u x w = 0), laminar fixed point (RHS == 0), a
Taylor-Green closed-form nonlinear term, and the exact per-step decay factor
of a Beltrami mode in a decaying box.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

PERIODIC_SYSTEMS = ("decaying-box", "kolmogorov", "waleffe")
MONOCHROMATIC = ("kolmogorov", "waleffe")


# ---------------------------------------------------------------- harmonics
def real_harmonics(n: int) -> np.ndarray:
    """``operators.py:18-34``: [0, n/2), Nyquist omitted."""
    return np.arange(0, n // 2, dtype=np.int64)


def complex_harmonics(n: int) -> np.ndarray:
    """``operators.py:37-56``: [0..n/2-1, -n/2+1..-1], Nyquist omitted."""
    qs = (np.arange(n, dtype=np.int64) + n // 2) % n - n // 2
    out = np.zeros(n - 1, dtype=np.int64)
    out[: n // 2] = qs[: n // 2]
    out[n // 2 :] = qs[n // 2 + 1 :]
    return out


# ------------------------------------------------------------ pad/truncate
def zeropad_fft(a: np.ndarray, n: int, axis: int) -> np.ndarray:
    """``fft.py:41-88``."""
    if axis not in (0, 1):
        raise ValueError(f"axis must be 0 or 1; got {axis}.")
    N = a.shape[axis] + 1
    if n < N:
        raise ValueError(f"Target size {n} is smaller than input size {N}.")
    if (n - N) % 2 != 0:
        raise ValueError(f"Difference (n - N) = {n - N} cannot be odd.")
    shape = list(a.shape)
    shape[axis] = n
    out = np.zeros(shape, dtype=a.dtype)
    idx_in = [slice(None)] * a.ndim
    idx_out = [slice(None)] * a.ndim
    idx_in[axis] = slice(None, N // 2)
    idx_out[axis] = slice(None, N // 2)
    out[tuple(idx_out)] = a[tuple(idx_in)]
    idx_in[axis] = slice(N // 2, None)
    idx_out[axis] = slice(n - N // 2 + 1, None)
    out[tuple(idx_out)] = a[tuple(idx_in)]
    return out


def truncate_fft(a: np.ndarray, n: int, axis: int) -> np.ndarray:
    """``fft.py:91-137``."""
    if axis not in (0, 1):
        raise ValueError(f"axis must be 0 or 1; got {axis}.")
    N = a.shape[axis]
    if n > N:
        raise ValueError(f"Target size {n} is larger than input size {N}.")
    if (N - n) % 2 != 0:
        raise ValueError(f"Difference (N - n) = {N - n} cannot be odd.")
    shape = list(a.shape)
    shape[axis] = n - 1
    out = np.zeros(shape, dtype=a.dtype)
    idx_in = [slice(None)] * a.ndim
    idx_out = [slice(None)] * a.ndim
    idx_in[axis] = slice(None, n // 2)
    idx_out[axis] = slice(None, n // 2)
    out[tuple(idx_out)] = a[tuple(idx_in)]
    idx_in[axis] = slice(N - n // 2 + 1, None)
    idx_out[axis] = slice(n // 2, None)
    out[tuple(idx_out)] = a[tuple(idx_in)]
    return out


def zeropad_rfft(a: np.ndarray, n: int) -> np.ndarray:
    """``fft.py:140-163`` (axis 2 = kx)."""
    N = a.shape[2]
    if n < N:
        raise ValueError(f"Target mode count {n} is smaller than input {N}.")
    out = np.zeros(a.shape[:2] + (n,), dtype=a.dtype)
    out[:, :, :N] = a
    return out


def truncate_rfft(a: np.ndarray, n: int) -> np.ndarray:
    """``fft.py:166-190`` (axis 2 = kx)."""
    if n > a.shape[2]:
        raise ValueError(f"Target mode count {n} is larger than input {a.shape[2]}.")
    return a[:, :, :n].copy()


# ------------------------------------------------------------------ config
@dataclass
class PeriodicParams:
    """The subset of ``parameters.py`` the periodic step reads.

    ``ly`` is a parameter here; the reference fixes it to 4
    (``parameters.py:188-189``).
    """

    nx: int = 32
    ny: int = 32
    nz: int = 32
    lx: float = 4.0
    ly: float = 4.0
    lz: float = 4.0
    re: float = 40.0
    system: str = "kolmogorov"
    oversampling_factor: int = 3
    oversample_y: bool = True
    tilt_degree: float = 0.0
    dt: float = 0.01
    implicitness: float = 0.5
    corrector_tolerance: float = 1e-10
    max_corrector_iterations: int = 10
    correct_divergence: bool = True

    @property
    def nx_padded(self) -> int:  # parameters.py:241-243
        return self.oversampling_factor * self.nx // 2

    @property
    def ny_padded(self) -> int:  # parameters.py:244-249
        return self.oversampling_factor * self.ny // 2 if self.oversample_y else self.ny

    @property
    def nz_padded(self) -> int:  # parameters.py:252-254
        return self.oversampling_factor * self.nz // 2

    @property
    def spec_shape(self) -> tuple:  # sharding.py:400-404
        return (self.ny - 1, self.nz - 1, self.nx // 2)

    @property
    def phys_shape(self) -> tuple:  # sharding.py:407-411
        return (self.ny_padded, self.nz_padded, self.nx_padded)

    @property
    def tilt_rad(self) -> float:  # parameters.py:195
        return math.pi * self.tilt_degree / 180


# ------------------------------------------------------------ transforms
# The axis transforms are numpy.fft's (pocketfft, one thread).  For timing
# the oracle on several host cores, fft_workers(n) routes them through
# scipy.fft with n worker threads (the same algorithm family; the rounding
# may differ in the last place, so the tests keep the default).
_FFT = np.fft
_FFT_KW: dict = {}


def fft_workers(n: int | None) -> None:
    """n > 1: scipy.fft with n workers; None or 1: numpy.fft."""
    global _FFT, _FFT_KW
    if n and n > 1:
        import scipy.fft

        _FFT, _FFT_KW = scipy.fft, {"workers": int(n)}
    else:
        _FFT, _FFT_KW = np.fft, {}


def irfft3d(x: np.ndarray, p: PeriodicParams) -> np.ndarray:
    """``fft.py:351-408``: pad y, ifft y, pad z, ifft z, pad kx, irfft x."""
    y = zeropad_fft(x, p.ny_padded, 0)
    y = _FFT.ifft(y, axis=0, norm="forward", **_FFT_KW)
    y = zeropad_fft(y, p.nz_padded, 1)
    y = _FFT.ifft(y, axis=1, norm="forward", **_FFT_KW)
    y = zeropad_rfft(y, p.nx_padded // 2 + 1)
    return _FFT.irfft(y, axis=2, norm="forward", **_FFT_KW)


def rfft3d(x: np.ndarray, p: PeriodicParams) -> np.ndarray:
    """``fft.py:291-348``: rfft x, truncate, fft z, truncate, fft y, truncate."""
    y = truncate_rfft(_FFT.rfft(x, axis=2, norm="forward", **_FFT_KW), p.nx // 2)
    y = truncate_fft(_FFT.fft(y, axis=1, norm="forward", **_FFT_KW), p.nz, 1)
    y = truncate_fft(_FFT.fft(y, axis=0, norm="forward", **_FFT_KW), p.ny, 0)
    return y


def spec_to_phys(v: np.ndarray, p: PeriodicParams) -> np.ndarray:
    """``operators.py:82-101`` (vmapped over the leading component axis)."""
    return np.stack([irfft3d(v[i], p) for i in range(v.shape[0])])


def phys_to_spec(v: np.ndarray, p: PeriodicParams) -> np.ndarray:
    """``operators.py:59-79``."""
    return np.stack([rfft3d(v[i], p) for i in range(v.shape[0])])


def cross(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """``operators.py:149-173``."""
    return np.array(
        [
            a[1] * b[2] - a[2] * b[1],
            a[2] * b[0] - a[0] * b[2],
            a[0] * b[1] - a[1] * b[0],
        ]
    )


# ------------------------------------------------------------ geometry
@dataclass
class Fourier:
    """``geometries/triply_periodic.py:16-75``."""

    p: PeriodicParams
    kx: np.ndarray = field(init=False)
    ky: np.ndarray = field(init=False)
    kz: np.ndarray = field(init=False)
    k_metric: np.ndarray = field(init=False)
    lapl: np.ndarray = field(init=False)
    inv_lapl: np.ndarray = field(init=False)

    def __post_init__(self) -> None:
        p = self.p
        self.kx = real_harmonics(p.nx).reshape(1, 1, -1) * 2 * np.pi / p.lx
        self.kz = complex_harmonics(p.nz).reshape(1, -1, 1) * 2 * np.pi / p.lz
        self.ky = complex_harmonics(p.ny).reshape(-1, 1, 1) * 2 * np.pi / p.ly
        self.k_metric = np.where(self.kx == 0, 1, 2)
        self.lapl = -(self.kx**2 + self.ky**2 + self.kz**2)
        with np.errstate(divide="ignore"):
            self.inv_lapl = np.where(self.lapl < 0, 1 / self.lapl, 0.0)


def get_inprod(a: np.ndarray, b: np.ndarray, k_metric: np.ndarray) -> float:
    """``geometries/triply_periodic.py:81-91`` (real part of the sum)."""
    return float(np.sum(np.conj(a) * k_metric * b).real)


def get_norm2(a: np.ndarray, k_metric: np.ndarray) -> float:
    """``geometries/triply_periodic.py:94-96``."""
    return get_inprod(a, a, k_metric)


def get_norm(a: np.ndarray, k_metric: np.ndarray) -> float:
    """``geometries/triply_periodic.py:99-101``."""
    return math.sqrt(get_norm2(a, k_metric))


def derivative(d, kx, ky, kz, axis):
    """``geometries/triply_periodic.py:104-114``."""
    return (1j * (kx, ky, kz)[axis]) * d


def divergence(v, kx, ky, kz):
    """``geometries/triply_periodic.py:117-119``."""
    return sum(derivative(v[i], kx, ky, kz, i) for i in range(3))


def curl(v, kx, ky, kz):
    """``geometries/triply_periodic.py:122-132``."""
    return 1j * np.array(
        [
            ky * v[2] - kz * v[1],
            kz * v[0] - kx * v[2],
            kx * v[1] - ky * v[0],
        ]
    )


def gradient(d, kx, ky, kz):
    """``geometries/triply_periodic.py:135-137``."""
    return np.array([derivative(d, kx, ky, kz, i) for i in range(3)])


# ------------------------------------------------------------ flow setup
@dataclass
class Flow:
    """``flows/monochromatic.py:85-285``, with ``ly`` a parameter.

    The laminar constants use the Ly-general forms F = nu (2 pi qf / Ly)^2,
    I_lam = F / 2 (equal to the reference's ``pi^2/(4 Re)``, ``pi^2/(8 Re)``
    at Ly = 4, ``flows/monochromatic.py:134-137``).
    """

    p: PeriodicParams
    fourier: Fourier = field(init=False)

    def __post_init__(self) -> None:
        p = self.p
        self.fourier = f = Fourier(p)
        ny_p = p.ny_padded
        bfc = np.zeros(ny_p // 2 + 1, dtype=np.complex128)
        self.qf = 1
        if p.system in MONOCHROMATIC:
            if p.system == "kolmogorov":
                bfc[self.qf] += -0.5j
            else:
                bfc[self.qf] += 0.5
            self.force_amplitude = (1.0 / p.re) * (2 * np.pi * self.qf / p.ly) ** 2
            self.ekin_lam = 0.25
            self.input_lam = self.force_amplitude / 2
            self.dissip_lam = self.input_lam
        elif p.system == "decaying-box":
            self.force_amplitude = 0.0
            self.ekin_lam = 0.0
            self.input_lam = 0.0
            self.dissip_lam = 0.0
        else:
            raise NotImplementedError(p.system)
        dbfc = 1j * (2 * np.pi / p.ly) * bfc
        base = np.zeros((3, ny_p))
        dbase = np.zeros((3, ny_p))
        cbase = np.zeros((3, ny_p))
        nbase = np.zeros((3, ny_p))
        base[0] = np.fft.irfft(bfc, n=ny_p, norm="forward")
        dbase[0] = np.fft.irfft(dbfc, n=ny_p, norm="forward")
        cbase[2] = -dbase[0]
        nbase[1] = base[0] * dbase[0]
        t = p.tilt_rad
        if abs(t) != 0:  # flows/monochromatic.py:192-203
            base[2] = np.sin(t) * base[0]
            base[0] = base[0] * np.cos(t)
            dbase[2] = np.sin(t) * dbase[0]
            dbase[0] = dbase[0] * np.cos(t)
            cbase[0] = -np.sin(t) * cbase[2]
            cbase[2] = cbase[2] * np.cos(t)
        self.base_flow = base[:, :, None, None]
        self.dy_base_flow = dbase[:, :, None, None]
        self.curl_base_flow = cbase[:, :, None, None]
        self.nonlin_base_flow = nbase[:, :, None, None]
        # flows/monochromatic.py:270-285
        ldt_1 = 1 / p.dt + (1 - p.implicitness) * f.lapl / p.re
        ildt_2 = 1 / (1 / p.dt - p.implicitness * f.lapl / p.re)
        ldt_1 = np.array(ldt_1, dtype=np.float64)
        ildt_2 = np.array(ildt_2, dtype=np.float64)
        ldt_1[0, 0, 0] = 0.0
        ildt_2[0, 0, 0] = 0.0
        self.ldt_1 = ldt_1
        self.ildt_2 = ildt_2


# ------------------------------------------------------------ callables
def get_nonlin(state: np.ndarray, flow: Flow) -> np.ndarray:
    """``rhs.py:34-92`` with the periodic transforms and curl."""
    p, f = flow.p, flow.fourier
    vel = spec_to_phys(state, p)
    vort = spec_to_phys(curl(state, f.kx, f.ky, f.kz), p)
    nl = cross(vel, vort)
    nl = nl + (
        (cross(vel, flow.curl_base_flow) + cross(flow.base_flow, vort))
        + flow.nonlin_base_flow
    )
    return phys_to_spec(nl, p)


def get_rhs(state: np.ndarray, flow: Flow) -> np.ndarray:
    """``flows/monochromatic.py:366-386``: NL minus grad of inv-lapl of div NL."""
    f = flow.fourier
    nonlin = get_nonlin(state, flow)
    lapl_pressure = divergence(nonlin, f.kx, f.ky, f.kz)
    return nonlin - gradient(f.inv_lapl * lapl_pressure, f.kx, f.ky, f.kz)


def predict(state: np.ndarray, rhs: np.ndarray, flow: Flow) -> np.ndarray:
    """``flows/monochromatic.py:320-333,389-393``."""
    return (state * flow.ldt_1 + rhs) * flow.ildt_2


def correct(state_prev, prediction, rhs_prev, rhs_next, flow: Flow):
    """``flows/monochromatic.py:336-355,396-405`` (incremental CN correction)."""
    correction = flow.p.implicitness * (rhs_next - rhs_prev) * flow.ildt_2
    return prediction + correction, correction


def norm(correction: np.ndarray, flow: Flow) -> float:
    """``flows/monochromatic.py:408-410``."""
    return get_norm(correction, flow.fourier.k_metric)


def correct_velocity(state: np.ndarray, flow: Flow) -> np.ndarray:
    """``flows/monochromatic.py:516-565``: Leray projection + mean mode = 0."""
    f = flow.fourier
    out = state
    if flow.p.correct_divergence:
        corr = -gradient(f.inv_lapl * divergence(state, f.kx, f.ky, f.kz), f.kx, f.ky, f.kz)
        out = state + corr
    out = out.copy()
    out[:, 0, 0, 0] = 0
    return out


# ------------------------------------------------------------ diagnostics
def forced_modes(flow: Flow):
    """``flows/monochromatic.py:210-267``: (indices, unit coefficients)."""
    p, qf = flow.p, flow.qf
    t = p.tilt_rad
    tilt = not (abs(p.tilt_degree) == 0 or abs(p.tilt_degree - 180) == 0)
    tilt90 = abs(p.tilt_degree + 90) == 0 or abs(p.tilt_degree - 90) == 0
    kolm = p.system == "kolmogorov"
    if not tilt:
        idx = ((0, 0), (qf, -qf), (0, 0), (0, 0))
    elif tilt90:
        idx = ((2, 2), (qf, -qf), (0, 0), (0, 0))
    else:
        idx = ((0, 0, 2, 2), (qf, -qf, qf, -qf), (0, 0, 0, 0), (0, 0, 0, 0))
    if not tilt or tilt90:
        uf = np.array([-0.5j, 0.5j] if kolm else [0.5, 0.5], dtype=np.complex128)
    elif kolm:
        uf = np.array([-0.5j * np.cos(t), 0.5j * np.cos(t), -0.5j * np.sin(t), 0.5j * np.sin(t)])
    else:
        uf = np.array([0.5 * np.cos(t), 0.5 * np.cos(t), -0.5j * np.sin(t), 0.5j * np.sin(t)])
    return idx, uf


def get_stats(state: np.ndarray, flow: Flow) -> dict:
    """``flows/monochromatic.py:433-505``: E, I, D, E'."""
    p, f = flow.p, flow.fourier
    ep = get_norm2(state, f.k_metric) / 2
    if p.system in MONOCHROMATIC:
        idx, uf = forced_modes(flow)
        inp = float(np.sum(np.conj(uf * flow.force_amplitude) * state[idx]).real) + flow.input_lam
        energy = ep - flow.ekin_lam + inp / flow.force_amplitude
    else:
        inp = 0.0
        energy = ep
    enst = float(np.sum(-(f.lapl * (np.conj(state) * state))).real) + 2 * inp * p.re - flow.input_lam * p.re
    return {"E": energy, "I": inp, "D": enst / p.re, "E'": ep}
