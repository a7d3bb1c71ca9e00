"""The per-(kz, kx) profile network of the plane Couette guess against the
oracle: its matrix-core forward and backward (bf16 operands, FP32
accumulation) against the FP64 oracle with the same bf16 rounding of the
weights and layer inputs (the device also rounds the backward operands dY
and dZ to bf16, hence the 2e-2 bound on the gradients), its training rows
and guess, and runs started from its guess against the plain trajectory.
"""

import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dnsjax_amd.params import Params  # noqa: E402
from oracle import mlp as M  # noqa: E402


def rel(a, b):
    a = a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _bf16_words(x: np.ndarray) -> torch.Tensor:
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + (((u >> 16) & 1) + 0x7FFF)) >> 16
    return torch.from_numpy(u.astype(np.uint16).view(np.int16)).cuda()


def _net(din, H, dout, seed):
    """A network of the given shape holding the oracle's initial parameters."""
    from dnsjax_amd.profile_net import ProfileNet

    net = ProfileNet(0, hidden=H, seed=seed, shape=(din, dout))
    params = M.init(din, H, dout, seed)
    net.params.copy_(torch.from_numpy(np.concatenate([p.reshape(-1) for p in params]).astype(np.float32)))
    return net, [p.astype(np.float32).astype(np.float64) for p in params]


def test_matrix_core_forward_and_gradients_match_the_oracle():
    """Shapes off the 64-tiles (din 96, H 64, dout 40, B 100): forward to
    1e-3 and every parameter gradient to 2e-2 relative."""
    din, H, dout, B = 96, 64, 40, 100
    net, params = _net(din, H, dout, seed=3)
    rng = np.random.default_rng(7)
    x = M.to_bf16(rng.standard_normal((B, din)))
    t = rng.standard_normal((B, dout)) * 0.5
    X = _bf16_words(x)
    T = torch.from_numpy(t.astype(np.float32)).cuda()
    Y = net.forward(X)
    y_ref, _ = M.forward(params, x, bf16=True)
    assert rel(Y, y_ref) < 1e-3, rel(Y, y_ref)
    loss = net.compute_grad(X, T)
    want_loss, want = M.profile_loss_and_grads(params, x, t.astype(np.float32).astype(np.float64), bf16=True)
    assert abs(float(loss) - want_loss) <= 1e-3 * want_loss
    got = [g.detach().cpu().numpy().astype(np.float64) for g in _split(net.grad, din, H, dout)]
    errs = [rel(a, b) for a, b in zip(got, want)]
    print(f"MEASURED profile net gradient rel errors {['%.2e' % e for e in errs]}")
    assert max(errs) < 2e-2, errs


def _split(flat, din, H, dout):
    shapes = [(din, H), (H,), (H, H), (H,), (H, dout), (dout,)]
    out, o = [], 0
    for s in shapes:
        k = math.prod(s)
        out.append(flat[o:o + k].view(s))
        o += k
    return out


def test_training_reduces_the_loss():
    din, H, dout, B = 64, 64, 32, 512
    net, _ = _net(din, H, dout, seed=1)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((B, din))
    w = rng.standard_normal((din, dout)) / math.sqrt(din)
    t = np.tanh(x @ w)
    X, T = _bf16_words(x), torch.from_numpy(t.astype(np.float32)).cuda()
    first = float(net.compute_grad(X, T))
    for _ in range(200):
        net.compute_grad(X, T)
        net.adam(3e-3)
    last = float(net.compute_grad(X, T))
    assert last < 0.2 * first, (first, last)


def _walls(nx, ny, nz, precision="fp64", tol=1e-10):
    from dnsjax_amd.flows import PlaneCouetteFlow

    return PlaneCouetteFlow(Params(system="plane-couette", nx=nx, ny=ny, nz=nz, lx=2 * math.pi, lz=math.pi,
                                   re=400.0, dt=0.005, fd_order=4, corrector_tolerance=tol,
                                   max_corrector_iterations=20, precision=precision))


def test_training_rows_and_guess_match_the_oracle():
    """Rows of four successive nonlinear terms, and the guess (prediction
    and rhs_prev) with an untrained network, against the oracle."""
    from dnsjax_amd.fields import wall_field
    from dnsjax_amd.learned import Scales
    from dnsjax_amd.profile_net import ProfileNet, profile_din
    from oracle import walls as W

    nx, ny, nz = 16, 33, 16
    flow = _walls(nx, ny, nz)
    of = W.WallFlow(W.WallParams(nx=nx, ny=ny, nz=nz, fd_order=4, re=400.0, dt=0.005))
    ns = [flow.get_rhs(flow.to_device(wall_field(nx, ny, nz, 2 * math.pi, math.pi, 0.05, seed=s))) for s in (1, 2, 3, 4)]
    sc = Scales(2.0, 3.0, 0.5)
    din = profile_din(ny)
    modes = torch.arange(0, (nz - 1) * (nx // 2), 3, dtype=torch.int64, device="cuda")
    X = torch.empty((modes.numel(), din), dtype=torch.int16, device="cuda")
    T = torch.empty((modes.numel(), 6 * ny), dtype=torch.float32, device="cuda")
    nat = flow._native
    nat._chk(nat.lib.dnsmi_pnet_features(nat.handle, ns[0].data_ptr(), ns[1].data_ptr(), ns[2].data_ptr(),
                                         ns[3].data_ptr(), modes.data_ptr(), modes.numel(), din, 2.0, 3.0, 0.5,
                                         X.data_ptr(), T.data_ptr(), None))
    torch.cuda.synchronize()
    nn = [n.cpu().numpy() for n in ns]
    xr, tr = M.features_walls_profile(*nn, modes.cpu().numpy(), of, 2.0, 3.0, 0.5, din)
    got_x = X.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    got_x = got_x.view(np.float32).astype(np.float64)
    assert rel(got_x, M.to_bf16(xr)) < 1e-2
    assert rel(T, tr) < 1e-6
    # the guess
    net = ProfileNet(ny, hidden=64, seed=5)
    state = flow.to_device(wall_field(nx, ny, nz, 2 * math.pi, math.pi, 0.05, seed=9))
    h1 = torch.stack([ns[3].real, ns[3].imag], -1).float().contiguous()
    h2 = torch.stack([ns[2].real, ns[2].imag], -1).float().contiguous()
    pred, rhs_prev = nat.empty(), nat.empty()
    net.guess(nat, state, pred, rhs_prev, h1, h2, sc)
    torch.cuda.synchronize()
    f32 = lambda t: t.cpu().numpy().astype(np.float32).astype(np.float64)  # noqa: E731
    h1c = f32(ns[3].real) + 1j * f32(ns[3].imag)
    h2c = f32(ns[2].real) + 1j * f32(ns[2].imag)
    params = [p.detach().cpu().numpy().astype(np.float64) for p in net.unpack()]
    want_pred, want_rhs, _ = M.guess_walls_profile(state.cpu().numpy(), h1c, h2c, of, params, sc.astuple(), din)
    e_rhs, e_pred = rel(rhs_prev, want_rhs), rel(pred, want_pred)
    print(f"MEASURED profile guess rhs_prev rel {e_rhs:.2e}, pred rel {e_pred:.2e}")
    # the untrained network's output is of the size of N itself here, so the
    # bf16 rounding of its inputs (taken from slightly different FP32
    # products on the two sides) shows at a few 1e-6 of N
    assert e_rhs < 2e-5 and e_pred < 2e-6


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_profile_guessed_run_stays_on_the_plain_trajectory(precision):
    """16 x 33 x 16 with a profile network trained on the run's own history:
    every step converges and the state stays within the tolerance's reach
    of the plain run."""
    from dnsjax_amd.driver import run
    from dnsjax_amd.fields import wall_field
    from dnsjax_amd.learned import collect, run_guessed
    from dnsjax_amd.profile_net import ProfileNet, fit_scales, make_dataset, mean_loss, train

    nx, ny, nz = 16, 33, 16
    flow = _walls(nx, ny, nz, precision)
    u0 = flow.to_device(wall_field(nx, ny, nz, 2 * math.pi, math.pi, 0.05, seed=2))
    snaps, u_mid = collect(flow, u0.clone(), 12)
    sc = fit_scales(snaps)
    net = ProfileNet(ny, hidden=128, seed=0)
    X, T = make_dataset(flow, snaps, sc, 2000, net.din, seed=1)
    train(net, X, T, iters=200, batch=256, lr=1e-3)
    loss, zero = mean_loss(net, X, T)
    n = 8
    plain = run(flow, u_mid.clone(), n)
    g = run_guessed(flow, u_mid.clone(), n, mode=2, net=net, scales=sc)
    print(f"MEASURED profile net train loss {loss:.3g} (zero {zero:.3g}); passes {g.passes}, "
          f"plain counts {plain.counts}")
    assert loss < zero
    assert plain.converged and g.converged and len(g.passes) == n
    assert rel(g.state.cpu().numpy(), plain.state.cpu().numpy()) <= (1e-8 if precision == "fp64" else 1e-6)
