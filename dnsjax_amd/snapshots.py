"""Snapshots in the reference's on-disk formats.

The reference starts a run from an ``.npz`` snapshot (``init_state``):

* triply periodic flows read ``velocity_phys``, the total velocity on the
  padded physical grid, shape (3, P_y, P_z, P_x) in [y, z, x], subtract the
  base flow and transform to spectral space (``flows/monochromatic.py:
  302-311``);
* plane Couette reads ``velocity_phys_nonexpanded``, shape (3, Ny, P_z, P_x)
  (y in grid points), and transforms x and z (``flows/plane_couette.py:
  233-242``).

The transforms follow the reference's dealiased conventions
(``fft.py:41-190``, ``operators.py:59-146``): ``norm="forward"`` (1/N on
each forward axis, none on the inverse), the x axis real (rfft, kx < nx/2
kept), the others complex with modes [0, n/2) and (P - n/2, P) kept and the
Nyquist mode dropped.  They run with torch's FFT on the tensor's device:
this is file I/O, not the time-step path.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from .params import Params


def _truncate(a: torch.Tensor, n: int, axis: int) -> torch.Tensor:
    """Keep modes [0, n/2) and (P - n/2, P) of a complex FFT axis (n - 1 modes)."""
    P = a.shape[axis]
    h = n // 2
    lo = a.narrow(axis, 0, h)
    hi = a.narrow(axis, P - h + 1, h - 1)
    return torch.cat([lo, hi], dim=axis)


def _zeropad(a: torch.Tensor, P: int, axis: int) -> torch.Tensor:
    """Inverse of ``_truncate``: n - 1 stored modes into a zero axis of length P."""
    m = a.shape[axis]
    h = (m + 1) // 2
    shape = list(a.shape)
    shape[axis] = P
    out = torch.zeros(shape, dtype=a.dtype, device=a.device)
    out.narrow(axis, 0, h).copy_(a.narrow(axis, 0, h))
    out.narrow(axis, P - (m - h), m - h).copy_(a.narrow(axis, h, m - h))
    return out


def phys_to_spec(phys: torch.Tensor, p: Params) -> torch.Tensor:
    """(3, P_y, P_z, P_x) real -> (3, ny-1, nz-1, nx/2) complex (``_rfft3d``)."""
    a = torch.fft.rfft(phys, dim=-1, norm="forward")[..., : p.nx // 2]
    a = _truncate(torch.fft.fft(a, dim=-2, norm="forward"), p.nz, -2)
    return _truncate(torch.fft.fft(a, dim=-3, norm="forward"), p.ny, -3).contiguous()


def spec_to_phys(spec: torch.Tensor, p: Params) -> torch.Tensor:
    """(3, ny-1, nz-1, nx/2) complex -> (3, P_y, P_z, P_x) real (``_irfft3d``)."""
    py, pz, px = p.padded
    a = torch.fft.ifft(_zeropad(spec, py, -3), dim=-3, norm="forward")
    a = torch.fft.ifft(_zeropad(a, pz, -2), dim=-2, norm="forward")
    full = torch.zeros(a.shape[:-1] + (px // 2 + 1,), dtype=a.dtype, device=a.device)
    full[..., : p.nx // 2] = a
    return torch.fft.irfft(full, n=px, dim=-1, norm="forward").contiguous()


def phys_to_spec_2d(phys: torch.Tensor, p: Params) -> torch.Tensor:
    """(3, Ny, P_z, P_x) real -> (3, nz-1, nx/2, Ny) complex (``_rfft2d``)."""
    a = torch.fft.rfft(phys, dim=-1, norm="forward")[..., : p.nx // 2]
    a = _truncate(torch.fft.fft(a, dim=-2, norm="forward"), p.nz, -2)
    return a.permute(0, 2, 3, 1).contiguous()


def spec_to_phys_2d(spec: torch.Tensor, p: Params) -> torch.Tensor:
    """(3, nz-1, nx/2, Ny) complex -> (3, Ny, P_z, P_x) real (``_irfft2d``)."""
    _, pz, px = p.padded
    a = spec.permute(0, 3, 1, 2)
    a = torch.fft.ifft(_zeropad(a, pz, -2), dim=-2, norm="forward")
    full = torch.zeros(a.shape[:-1] + (px // 2 + 1,), dtype=a.dtype, device=a.device)
    full[..., : p.nx // 2] = a
    return torch.fft.irfft(full, n=px, dim=-1, norm="forward").contiguous()


def base_flow(p: Params, device=None) -> torch.Tensor:
    """The laminar profile (3, P_y, 1, 1) of a triply periodic flow
    (``flows/monochromatic.py:117-203``): u_x = sin(2 pi y / Ly) (Kolmogorov)
    or cos (Waleffe), zero for the decaying box, tilted by ``tilt_degree``
    into (x, z)."""
    py = p.padded[0]
    y = torch.arange(py, dtype=torch.float64, device=device) * (p.ly / py)
    out = torch.zeros((3, py, 1, 1), dtype=torch.float64, device=device)
    if p.system == "decaying-box":
        return out
    arg = 2 * math.pi * y / p.ly
    prof = torch.sin(arg) if p.system == "kolmogorov" else torch.cos(arg)
    t = math.radians(p.tilt_degree)
    out[0, :, 0, 0] = prof * math.cos(t)
    if t != 0.0:
        out[2, :, 0, 0] = prof * math.sin(t)
    return out


def load_snapshot(path: str, p: Params, device=None) -> torch.Tensor:
    """The spectral state (contiguous complex128, the reference layout) of a
    snapshot."""
    with np.load(path, allow_pickle=False) as f:
        key = "velocity_phys_nonexpanded" if p.walls else "velocity_phys"
        phys = torch.from_numpy(np.asarray(f[key], dtype=np.float64)).to(device)
    if p.walls:
        return phys_to_spec_2d(phys, p)
    return phys_to_spec(phys - base_flow(p, phys.device), p)


def init_state(p: Params, snapshot: str | None = None, start_from_laminar: bool = False,
               device=None) -> torch.Tensor:
    """The reference's ``init_state`` (``flows/monochromatic.py:294-314``,
    ``flows/plane_couette.py:225-248``): zeros when starting from the laminar
    state, else the snapshot's spectral state; neither is an error."""
    if start_from_laminar:
        return torch.zeros((3, *p.spec_shape), dtype=torch.complex128, device=device)
    if snapshot is None:
        raise ValueError("Provide an initial condition.")
    return load_snapshot(snapshot, p, device)


def save_snapshot(path: str, state: torch.Tensor, p: Params) -> None:
    """Write a state in the reference's snapshot format (total velocity for
    periodic flows)."""
    s = state.detach().to(torch.complex128)
    if p.walls:
        phys = spec_to_phys_2d(s, p)
        np.savez(path, velocity_phys_nonexpanded=phys.cpu().numpy())
    else:
        phys = spec_to_phys(s, p) + base_flow(p, s.device)
        np.savez(path, velocity_phys=phys.cpu().numpy())
