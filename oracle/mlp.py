"""Oracle: the per-mode corrector-guess network, FP64 NumPy -- TEST INFRASTRUCTURE ONLY.

The reference has no learned component; this restates the network the
library trains and runs, so that the GPU kernels' forward values and
parameter gradients can be checked:

    x (D_IN) -> z1 = W1^T x + b1 -> h1 = silu(z1)
             -> z2 = W2^T h1 + b2 -> h2 = silu(z2)
             -> y  = W3^T h2 + b3            (D_OUT)
    loss = (1/B) sum_samples sum_outputs (y - t)^2

with W1 (D_IN, H), W2 (H, H), W3 (H, D_OUT).  ``bf16=True`` rounds the
weights and the layer inputs to bfloat16 (round to nearest even) as the
GPU inference kernel feeds its matrix cores.

``features`` and ``guess`` restate the corrector-guess step built on the
network for the periodic stepper (no reference counterpart; it reuses the
reference's predictor ``flows/monochromatic.py:320-333`` and the Leray
projection ``flows/monochromatic.py:516-565``); ``features_walls`` and
``guess_walls`` the same for plane Couette, per point (kz, kx, y) (it reuses
the predictor's solve ``flows/plane_couette.py:507-519``).

Pinned by: central finite differences of the loss (tests/test_oracle_mlp.py),
which fix every gradient independently of the backward formulas; for
``guess``: mode 0 reduces to the reference predictor exactly, and the
fixed-point identity (the corrector pass from the guess equals the plain
iteration map, tests/test_oracle_mlp.py) holds for any dN; for the wall
features: exact targets on polynomial histories (the residual of quadratic
extrapolation vanishes for quadratic and equals the third difference for
cubic histories) and coordinates read back from a field built from them.
"""

from __future__ import annotations

import numpy as np


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def to_bf16(a: np.ndarray) -> np.ndarray:
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    u = ((u + r) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(params, x, bf16=False):
    W1, b1, W2, b2, W3, b3 = params
    q = to_bf16 if bf16 else (lambda a: a)
    z1 = q(x) @ q(W1) + b1
    h1 = silu(z1)
    z2 = q(h1) @ q(W2) + b2
    h2 = silu(z2)
    y = q(h2) @ q(W3) + b3
    return y, (x, z1, h1, z2, h2)


def loss_and_grads(params, x, t, bf16=False):
    """Loss and parameter gradients.  ``bf16`` rounds weights and layer
    inputs as ``forward`` does and differentiates straight through the
    rounding (the rounded values are used in the backward pass)."""
    W1, b1, W2, b2, W3, b3 = params
    q = to_bf16 if bf16 else (lambda a: a)
    y, (x, z1, h1, z2, h2) = forward(params, x, bf16)
    xq, h1q, h2q = q(x), q(h1), q(h2)
    W2q, W3q = q(W2), q(W3)
    B = x.shape[0]
    loss = float(np.sum((y - t) ** 2) / B)
    dy = 2.0 * (y - t) / B
    gW3 = h2q.T @ dy
    gb3 = dy.sum(0)
    dz2 = (dy @ W3q.T) * dsilu(z2)
    gW2 = h1q.T @ dz2
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2q.T) * dsilu(z1)
    gW1 = xq.T @ dz1
    gb1 = dz1.sum(0)
    return loss, (gW1, gb1, gW2, gb2, gW3, gb3)


def init(d_in, h, d_out, seed):
    rng = np.random.default_rng(seed)
    W1 = rng.standard_normal((d_in, h)) / np.sqrt(d_in)
    W2 = rng.standard_normal((h, h)) / np.sqrt(h)
    W3 = rng.standard_normal((h, d_out)) / np.sqrt(h) * 0.1
    return [W1, np.zeros(h), W2, np.zeros(h), W3, np.zeros(d_out)]


def adam(params, grads, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015), in place."""
    for i in range(len(params)):
        m[i] = b1 * m[i] + (1 - b1) * grads[i]
        v[i] = b2 * v[i] + (1 - b2) * grads[i] ** 2
        mh = m[i] / (1 - b1**step)
        vh = v[i] / (1 - b2**step)
        params[i] = params[i] - lr * mh / (np.sqrt(vh) + eps)
    return params


# ------------------------------------------------------------ corrector guess
def mode_weight(flow):
    """w = ildt_2 / dt per mode, 0 for the mean mode (held fixed)."""
    from .periodic import Fourier  # noqa: F401  (flow.fourier is one)

    f, p = flow.fourier, flow.p
    w = 1.0 / (1.0 - p.dt * p.implicitness * f.lapl / p.re)
    w = np.broadcast_to(w, f.lapl.shape).copy()
    w[0, 0, 0] = 0.0
    return w


def k_features(flow):
    """Features 12..15 per mode: min(dt |k|^2/Re, 8), kx, ky, kz over k_max."""
    f, p = flow.fourier, flow.p
    kmax = max(2 * np.pi / p.lx * (p.nx // 2), 2 * np.pi / p.ly * (p.ny // 2), 2 * np.pi / p.lz * (p.nz // 2))
    shape = f.lapl.shape
    return [np.minimum(p.dt * (-f.lapl) / p.re, 8.0),
            np.broadcast_to(f.kx / kmax, shape), np.broadcast_to(f.ky / kmax, shape),
            np.broadcast_to(f.kz / kmax, shape)]


def features(n_m2, n_m1, n_cur, n_next, modes, flow, s_1, s_2, s_out, n_m3=None):
    """X (S,16), T (S,6) for flat mode indices over (ky, kz, kx): D1, D2
    (older terms rounded to FP32), k features; T the weighted residual of
    quadratic extrapolation, or of cubic extrapolation when n_m3 is given."""
    f32 = lambda a: a.astype(np.float32).astype(np.float64)  # noqa: E731
    X = np.zeros((len(modes), 16))
    for q in range(6):
        comp = (lambda a: a.real) if q % 2 == 0 else (lambda a: a.imag)
        cur = comp(n_cur[q // 2].reshape(-1)[modes])
        p1 = f32(comp(n_m1[q // 2].reshape(-1)[modes]))
        p2 = f32(comp(n_m2[q // 2].reshape(-1)[modes]))
        X[:, q] = (cur - p1) * s_1
        X[:, 6 + q] = ((cur - 2 * p1) + p2) * s_2
    for j, kf in enumerate(k_features(flow)):
        X[:, 12 + j] = kf.reshape(-1)[modes]
    T = None
    if n_next is not None:
        w = mode_weight(flow).reshape(-1)[modes]
        T = np.zeros((len(modes), 6))
        for q in range(6):
            c = q // 2
            r = n_next[c] - 3 * n_cur[c] + 3 * n_m1[c] - n_m2[c]
            if n_m3 is not None:  # minus the fourth difference's extrapolated part
                r = n_next[c] - 4 * n_cur[c] + 6 * n_m1[c] - 4 * n_m2[c] + n_m3[c]
            r = r.reshape(-1)[modes]
            T[:, q] = w * (r.real if q % 2 == 0 else r.imag) / s_out
    return X, T


def guess(state, h1, h2, flow, mode, params=None, scales=(1.0, 1.0, 1.0), h3=None):
    """(pred, rhs_prev, N^n) of the guessed predictor.

    h1, h2, h3: N^{n-1}, N^{n-2}, N^{n-3} as complex values (rounded to FP32
    by the caller as the device history keeps them; h3 optional).  mode 0:
    plain predictor; 1: dN = D1; 3: dN = D1 + D2; 4: dN = D1 + D2 + D3
    (cubic, needs h3); 2: the quadratic (without h3) or cubic (with h3)
    extrapolation plus s_out * net / w (bf16 inference)."""
    from . import periodic as P

    s_1, s_2, s_out = scales
    f, p = flow.fourier, flow.p
    n0 = P.get_rhs(state, flow)
    u1 = P.predict(state, n0, flow)
    if mode == 0:
        return u1, n0, n0.copy()
    d1 = n0 - h1
    dn = d1 if mode == 1 else d1 + ((n0 - 2 * h1) + h2)
    if mode == 4 or (mode == 2 and h3 is not None):
        dn = dn + (((n0 - 3 * h1) + 3 * h2) - h3)
    w = mode_weight(flow)
    if mode == 2:
        modes = np.arange(w.size)
        X, _ = features(h2, h1, n0, None, modes, flow, s_1, s_2, s_out)
        Y, _ = forward(params, X, bf16=True)
        with np.errstate(divide="ignore", invalid="ignore"):
            for c in range(3):
                net = (Y[:, 2 * c] + 1j * Y[:, 2 * c + 1]).reshape(w.shape)
                dn[c] = dn[c] + np.where(w > 0, s_out * net / np.where(w > 0, w, 1.0), 0.0)
    kd = f.kx * dn[0] + f.ky * dn[1] + f.kz * dn[2]
    d = np.stack([dn[0] + f.kx * kd * f.inv_lapl, dn[1] + f.ky * kd * f.inv_lapl,
                  dn[2] + f.kz * kd * f.inv_lapl])
    d[:, 0, 0, 0] = 0.0
    pred = u1 + p.implicitness * p.dt * w * d
    return pred, n0 + d, n0.copy()


def mode_weight_walls(flow):
    """w = 1 / (1 + c dt (kx^2 + kz^2) / Re) per point (kz, kx, y)."""
    p = flow.p
    w = 1.0 / (1.0 - p.dt * p.implicitness * (-(flow.kx**2 + flow.kz**2)) / p.re)
    return np.broadcast_to(w, (p.nz - 1, p.nx // 2, p.ny)).copy()


def k_features_walls(flow):
    """Features 12..15 per point: min(dt (kx^2 + kz^2)/Re, 8), kx/kmax, y, kz/kmax."""
    p = flow.p
    kmax = max(2 * np.pi / p.lx * (p.nx // 2), 2 * np.pi / p.lz * (p.nz // 2))
    shape = (p.nz - 1, p.nx // 2, p.ny)
    k2 = flow.kx**2 + flow.kz**2
    return [np.broadcast_to(np.minimum(p.dt * k2 / p.re, 8.0), shape),
            np.broadcast_to(flow.kx / kmax, shape), np.broadcast_to(flow.ys.reshape(1, 1, -1), shape),
            np.broadcast_to(flow.kz / kmax, shape)]


def features_walls(n_m2, n_m1, n_cur, n_next, points, flow, s_1, s_2, s_out):
    """X (S,16), T (S,6) for flat point indices over (kz, kx, y), as
    ``features`` with the wall-bounded coordinates and weight."""
    f32 = lambda a: a.astype(np.float32).astype(np.float64)  # noqa: E731
    X = np.zeros((len(points), 16))
    for q in range(6):
        comp = (lambda a: a.real) if q % 2 == 0 else (lambda a: a.imag)
        cur = comp(n_cur[q // 2].reshape(-1)[points])
        p1 = f32(comp(n_m1[q // 2].reshape(-1)[points]))
        p2 = f32(comp(n_m2[q // 2].reshape(-1)[points]))
        X[:, q] = (cur - p1) * s_1
        X[:, 6 + q] = ((cur - 2 * p1) + p2) * s_2
    for j, kf in enumerate(k_features_walls(flow)):
        X[:, 12 + j] = kf.reshape(-1)[points]
    T = None
    if n_next is not None:
        w = mode_weight_walls(flow).reshape(-1)[points]
        T = np.zeros((len(points), 6))
        for q in range(6):
            d3 = (n_next[q // 2] - 3 * n_cur[q // 2] + 3 * n_m1[q // 2] - n_m2[q // 2]).reshape(-1)[points]
            T[:, q] = w * (d3.real if q % 2 == 0 else d3.imag) / s_out
    return X, T


def guess_walls(state, h1, h2, flow, mode, params=None, scales=(1.0, 1.0, 1.0)):
    """(pred, rhs_prev, N^n) of the wall-bounded guess: the predictor's solve
    (``flows/plane_couette.py:507-519``) fed with N_ext = N^n + dN, dN = 0
    (mode 0), D1 (mode 1), D1 + D2 (mode 3) or D1 + D2 + s_out net / w
    (mode 2, bf16 inference); rhs_prev = N_ext.  h1, h2 are N^{n-1}, N^{n-2}
    rounded to FP32 by the caller."""
    from . import walls as W

    n0 = W.get_rhs(state, flow)
    if mode == 0:
        ext = n0
    else:
        d1 = n0 - h1
        ext = n0 + (d1 if mode == 1 else d1 + ((n0 - 2 * h1) + h2))
    if mode == 2:
        s_1, s_2, s_out = scales
        w = mode_weight_walls(flow)
        pts = np.arange(w.size)
        X, _ = features_walls(h2, h1, n0, None, pts, flow, s_1, s_2, s_out)
        Y, _ = forward(params, X, bf16=True)
        for c in range(3):
            net = (Y[:, 2 * c] + 1j * Y[:, 2 * c + 1]).reshape(w.shape)
            ext[c] = ext[c] + s_out * net / w
    return W.predict(state, ext, flow), ext, n0.copy()


# ------------------------------------------------------------ profile network
def profile_din(ny: int) -> int:
    """Input width of the per-(kz, kx) profile network: 12 Ny profile values
    and three mode features, padded to a multiple of 32."""
    return -(-(12 * ny + 3) // 32) * 32


def features_walls_profile(n_m2, n_m1, n_cur, n_next, modes, flow, s_1, s_2, s_out, din):
    """X (S, din), T (S, 6 Ny) for flat mode indices over (kz, kx): the wall-
    normal profiles of D1 and D2 (older terms rounded to FP32) ordered
    [component][real, imaginary][y], D1 block then D2 block, then
    min(dt k^2/Re, 8), kx/k_max, kz/k_max, zeros; T the profile of the
    weighted residual of quadratic extrapolation in the D1 block's order."""
    p = flow.p
    ny = p.ny
    f32 = lambda a: a.astype(np.float32).astype(np.float64)  # noqa: E731
    nkx = p.nx // 2
    iz, ix = np.asarray(modes) // nkx, np.asarray(modes) % nkx
    X = np.zeros((len(modes), din))
    for c in range(3):
        cur = n_cur[c][iz, ix]        # (S, Ny)
        p1 = f32(n_m1[c][iz, ix].real) + 1j * f32(n_m1[c][iz, ix].imag)
        p2 = f32(n_m2[c][iz, ix].real) + 1j * f32(n_m2[c][iz, ix].imag)
        d1 = cur - p1
        d2 = (cur - 2 * p1) + p2
        o = 2 * c * ny
        X[:, o:o + ny] = d1.real * s_1
        X[:, o + ny:o + 2 * ny] = d1.imag * s_1
        X[:, 6 * ny + o:6 * ny + o + ny] = d2.real * s_2
        X[:, 6 * ny + o + ny:6 * ny + o + 2 * ny] = d2.imag * s_2
    kx, kz = flow.kx_global[ix], flow.kz_global[iz]
    kmax = max(2 * np.pi / p.lx * (p.nx // 2), 2 * np.pi / p.lz * (p.nz // 2))
    k2 = kx**2 + kz**2
    X[:, 12 * ny] = np.minimum(p.dt * k2 / p.re, 8.0)
    X[:, 12 * ny + 1] = kx / kmax
    X[:, 12 * ny + 2] = kz / kmax
    T = None
    if n_next is not None:
        w = 1.0 / (1.0 + p.dt * p.implicitness * k2 / p.re)
        T = np.zeros((len(modes), 6 * ny))
        for c in range(3):
            d3 = (n_next[c] - 3 * n_cur[c] + 3 * n_m1[c] - n_m2[c])[iz, ix]
            o = 2 * c * ny
            T[:, o:o + ny] = w[:, None] * d3.real / s_out
            T[:, o + ny:o + 2 * ny] = w[:, None] * d3.imag / s_out
    return X, T


def profile_loss_and_grads(params, x, t, bf16=True):
    """The profile network's loss (1/(B dout)) sum (y - t)^2 and gradients:
    ``loss_and_grads`` divided by dout."""
    loss, g = loss_and_grads(params, x, t, bf16)
    d = t.shape[1]
    return loss / d, tuple(gi / d for gi in g)


def guess_walls_profile(state, h1, h2, flow, params, scales, din):
    """(pred, rhs_prev, N^n) of the wall-bounded guess with the profile
    network: N_ext = N^n + D1 + D2 + s_out net / w per point (bf16
    inference), the predictor's solve fed with it (h1, h2 rounded to FP32 by
    the caller)."""
    from . import walls as W

    p = flow.p
    s_1, s_2, s_out = scales
    n0 = W.get_rhs(state, flow)
    ext = n0 + (n0 - h1) + ((n0 - 2 * h1) + h2)
    M = (p.nz - 1) * (p.nx // 2)
    X, _ = features_walls_profile(h2, h1, n0, None, np.arange(M), flow, s_1, s_2, s_out, din)
    Y, _ = forward(params, X, bf16=True)
    ny = p.ny
    w = mode_weight_walls(flow)
    for c in range(3):
        o = 2 * c * ny
        net = (Y[:, o:o + ny] + 1j * Y[:, o + ny:o + 2 * ny]).reshape(p.nz - 1, p.nx // 2, ny)
        ext[c] = ext[c] + s_out * net / w
    return W.predict(state, ext, flow), ext, n0.copy()
