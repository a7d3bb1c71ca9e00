"""Learned corrector guess on the GPU against the oracle.

Training kernels (FP32 gradients, Adam) against the FP64 oracle, the
matrix-core bf16 inference against the oracle's bf16 emulation, the guess
kernel against ``oracle.mlp.guess``, and whole runs: a guessed run must stay
on the plain run's trajectory to the corrector tolerance.
"""

import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from oracle import mlp as M  # noqa: E402
from oracle import periodic as P  # noqa: E402
from dnsjax_amd.params import Params  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _net_with(H, params_np):
    from dnsjax_amd.learned import Net

    net = Net(H, seed=0)
    net.set_params(torch.from_numpy(np.concatenate([p.reshape(-1) for p in params_np])).float())
    return net


def _random_params(H, seed):
    rng = np.random.default_rng(seed)
    p = M.init(16, H, 6, seed=seed)
    p[1] = rng.standard_normal(H) * 0.1
    p[3] = rng.standard_normal(H) * 0.1
    p[5] = rng.standard_normal(6) * 0.1
    p[4] = p[4] * 10
    return [a.astype(np.float32).astype(np.float64) for a in p]


@pytest.mark.parametrize("H", [32, 64])
@pytest.mark.parametrize("bf16", [False, True])
def test_gradients_match_oracle(H, bf16):
    rng = np.random.default_rng(H + int(bf16))
    B = 77  # not a multiple of the 32-sample block
    params = _random_params(H, 7)
    x = rng.standard_normal((B, 16)).astype(np.float32).astype(np.float64)
    t = rng.standard_normal((B, 6)).astype(np.float32).astype(np.float64)
    loss, grads = M.loss_and_grads(params, x, t, bf16=bf16)
    net = _net_with(H, params)
    dev = net.device
    gl = net.compute_grad(torch.from_numpy(x).float().to(dev), torch.from_numpy(t).float().to(dev), bf16=bf16)
    got = net.grad.cpu().double().numpy()
    tol = 3e-3 if bf16 else 1e-4  # bf16: an FP32 difference may flip one rounding
    assert abs(float(gl) - loss) <= (1e-3 if bf16 else 1e-5) * loss
    o = 0
    for g in grads:
        k = g.size
        assert rel(got[o:o + k], g.reshape(-1)) <= tol, (H, bf16, o)
        o += k


def test_adam_matches_oracle():
    H = 32
    params = _random_params(H, 3)
    rng = np.random.default_rng(5)
    net = _net_with(H, params)
    m = [np.zeros_like(p) for p in params]
    v = [np.zeros_like(p) for p in params]
    for step in range(1, 4):
        grads = [rng.standard_normal(p.shape).astype(np.float32).astype(np.float64) for p in params]
        net.grad.copy_(torch.from_numpy(np.concatenate([g.reshape(-1) for g in grads])).float())
        net.adam(1e-2)
        M.adam(params, grads, m, v, step, 1e-2)
    got = net.params.cpu().double().numpy()
    want = np.concatenate([p.reshape(-1) for p in params])
    assert np.max(np.abs(got - want)) <= 2e-6 * np.max(np.abs(want))


@pytest.mark.parametrize("H", [32, 64])
def test_matrix_core_forward_matches_bf16_oracle(H):
    rng = np.random.default_rng(11)
    B = 1000
    params = _random_params(H, 9)
    x = rng.standard_normal((B, 16)).astype(np.float32).astype(np.float64)
    want, _ = M.forward(params, x, bf16=True)
    net = _net_with(H, params)
    net.pack()
    got = net.forward(torch.from_numpy(x).float().to(net.device)).cpu().double().numpy()
    assert rel(got, want) <= 2e-3
    full, _ = M.forward(params, x, bf16=False)
    assert rel(got, full) <= 5e-2  # bf16 inference stays near the FP64 network


def _periodic(n, re, precision="fp64", dt=0.01):
    from dnsjax_amd.flows import TriplyPeriodicFlow

    L = 2 * math.pi
    return TriplyPeriodicFlow(Params(system="kolmogorov", nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, re=re, dt=dt,
                                     corrector_tolerance=1e-10, max_corrector_iterations=10,
                                     precision=precision, tile_cols=16))


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_guess_kernel_matches_oracle(mode):
    from dnsjax_amd.fields import periodic_field

    n, L = 16, 2 * math.pi
    flow = _periodic(n, 40.0)
    of = P.Flow(P.PeriodicParams(nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, re=40.0, dt=0.01, system="kolmogorov"))
    u = periodic_field(n, n, n, L, L, L, 0.1, 3)
    h1 = P.get_rhs(periodic_field(n, n, n, L, L, L, 0.1, 4), of).astype(np.complex64)
    h2 = P.get_rhs(periodic_field(n, n, n, L, L, L, 0.1, 6), of).astype(np.complex64)
    params = _random_params(32, 2)
    scales = (1.0, 3.0, 0.05)
    want_pred, want_rhs, want_hist = M.guess(u, h1.astype(np.complex128), h2.astype(np.complex128), of, mode,
                                             params, scales)
    net = _net_with(32, params)
    net.pack()
    nat = flow._native
    dev = lambda h: torch.from_numpy(np.stack([h.real, h.imag], -1).astype(np.float32)).to(nat.device)  # noqa: E731
    d1, d2 = dev(h1).contiguous(), dev(h2).contiguous()
    pred, rhs = nat.empty(), nat.empty()
    nat.guess(flow.to_device(u), pred, rhs, d1, d2, mode, net.operands, scales)
    got_pred, got_rhs = pred.cpu().numpy(), rhs.cpu().numpy()
    u1 = P.predict(u, P.get_rhs(u, of), of)
    assert rel(got_rhs, want_rhs) <= (1e-4 if mode == 2 else 1e-12)
    if mode == 0:
        assert rel(got_pred, want_pred) <= 1e-12
    else:
        assert rel(got_pred - u1, want_pred - u1) <= (5e-3 if mode == 2 else 1e-10)
    h = d2.cpu().numpy().astype(np.float64)  # hist2 now holds N^n
    w = np.stack([want_hist.real, want_hist.imag], -1)
    assert np.max(np.abs(h - w)) <= 1e-7 * np.max(np.abs(w))  # kept as FP32
    assert torch.equal(d1, dev(h1))  # hist1 is read only


@pytest.mark.parametrize("mode", [0, 3, 2])
def test_guess_carry_output(mode):
    """With carry, the guess writes B = ildt_2 (ldt_1 u^n + (1 - c) N^n) (the
    carry form's corrector input) in place of N^n + dN, and the same
    prediction; one carry-form pass from it equals the incremental pass
    from N^n + dN."""
    from dnsjax_amd.fields import periodic_field

    n, L = 16, 2 * math.pi
    flow = _periodic(n, 40.0)
    of = P.Flow(P.PeriodicParams(nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, re=40.0, dt=0.01, system="kolmogorov"))
    u = periodic_field(n, n, n, L, L, L, 0.1, 3)
    hs = [P.get_rhs(periodic_field(n, n, n, L, L, L, 0.1, s), of).astype(np.complex64) for s in (4, 6)]
    net = _net_with(32, _random_params(32, 2))
    net.pack()
    nat = flow._native
    dev = lambda h: torch.from_numpy(np.stack([h.real, h.imag], -1).astype(np.float32)).to(nat.device)  # noqa: E731
    ud = flow.to_device(u)
    outs = []
    for carry in (False, True):
        d = [dev(h).contiguous() for h in hs]
        pred, rhs = nat.empty(), nat.empty()
        nat.guess(ud, pred, rhs, d[0], d[1], mode, net.operands, (1.0, 3.0, 0.05), carry=carry)
        outs.append((pred, rhs))
    assert torch.equal(outs[0][0], outs[1][0])
    n0 = P.get_rhs(u, of)
    want_b = of.ildt_2 * (of.ldt_1 * u + (1 - of.p.implicitness) * n0)
    assert rel(outs[1][1].cpu().numpy(), want_b) <= 1e-12
    # one pass each way from the same guess
    p_inc, p_car = outs[0][0].clone(), outs[1][0].clone()
    err = nat.scalar()
    nat.iterate_correction(ud, p_inc, outs[0][1], nat.empty(), err)
    nat.iterate_correction_carry(p_car, outs[1][1], err)
    assert rel(p_car.cpu().numpy(), p_inc.cpu().numpy()) <= 1e-13


@pytest.mark.parametrize("mode", [4, 2])
def test_cubic_guess_kernel_matches_oracle(mode):
    """With a third history: cubic extrapolation (mode 4) and the network on
    top of it (mode 2) against the oracle; N^n overwrites the oldest history, hist3, and the other two are
    read only."""
    from dnsjax_amd.fields import periodic_field

    n, L = 16, 2 * math.pi
    flow = _periodic(n, 40.0)
    of = P.Flow(P.PeriodicParams(nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, re=40.0, dt=0.01, system="kolmogorov"))
    u = periodic_field(n, n, n, L, L, L, 0.1, 3)
    hs = [P.get_rhs(periodic_field(n, n, n, L, L, L, 0.1, s), of).astype(np.complex64) for s in (4, 6, 8)]
    params = _random_params(32, 2)
    scales = (1.0, 3.0, 0.05)
    want_pred, want_rhs, want_hist = M.guess(u, *[h.astype(np.complex128) for h in hs[:2]], of, mode, params, scales,
                                             h3=hs[2].astype(np.complex128))
    net = _net_with(32, params)
    net.pack()
    nat = flow._native
    dev = lambda h: torch.from_numpy(np.stack([h.real, h.imag], -1).astype(np.float32)).to(nat.device)  # noqa: E731
    d = [dev(h).contiguous() for h in hs]
    pred, rhs = nat.empty(), nat.empty()
    nat.guess(flow.to_device(u), pred, rhs, d[0], d[1], mode, net.operands, scales, hist3=d[2])
    u1 = P.predict(u, P.get_rhs(u, of), of)
    got_pred, got_rhs = pred.cpu().numpy(), rhs.cpu().numpy()
    assert rel(got_rhs, want_rhs) <= (1e-4 if mode == 2 else 1e-12)
    assert rel(got_pred - u1, want_pred - u1) <= (5e-3 if mode == 2 else 1e-10)
    if mode == 2:  # the network did contribute beyond cubic extrapolation
        _, cub, _ = M.guess(u, *[h.astype(np.complex128) for h in hs[:2]], of, 4, h3=hs[2].astype(np.complex128))
        assert rel(want_rhs, cub) > 1e-4
    h = d[2].cpu().numpy().astype(np.float64)
    w = np.stack([want_hist.real, want_hist.imag], -1)
    assert np.max(np.abs(h - w)) <= 1e-7 * np.max(np.abs(w))
    assert torch.equal(d[0], dev(hs[0])) and torch.equal(d[1], dev(hs[1]))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_guessed_runs_stay_on_the_plain_trajectory(precision):
    """Extrapolated and learned guesses: every step converges and the states
    stay within the corrector tolerance's reach of the plain run."""
    from dnsjax_amd.driver import run
    from dnsjax_amd.learned import Net, collect, fit_scales, make_dataset, mean_loss, run_guessed, train

    flow = _periodic(32, 60.0, precision, dt=0.01)
    u0 = flow.init_state(seed=5, rms=0.1)
    snaps, u_mid = collect(flow, u0, 12)
    scales = fit_scales(snaps)
    X, T = make_dataset(flow, snaps, scales, 20000, seed=1)
    assert X.shape == (10 * 20000, 16)  # 13 terms: 10 windows of four
    net = Net(32, seed=0)
    train(net, X, T, iters=300, batch=4096, lr=3e-3)
    loss, base = mean_loss(net, X, T)
    assert loss < base  # better than plain extrapolation on its own data

    n = 10
    plain = run(flow, u_mid.clone(), n)
    assert plain.converged
    evals = {}
    for mode in (1, 3, 4, 2):
        g = run_guessed(flow, u_mid.clone(), n, mode=mode, net=net, scales=scales)
        assert g.converged and len(g.passes) == n, (mode, g.passes)
        d = rel(g.state.cpu().numpy(), plain.state.cpu().numpy())
        assert d <= (1e-8 if precision == "fp64" else 1e-7), (mode, d)
        evals[mode] = g.rhs_evals
        if mode in (1, 3, 4):
            assert g.rhs_evals <= plain.rhs_evals, (mode, g.passes, plain.counts)
    print(f"MEASURED guessed RHS evaluations over {n} steps ({precision}): plain {plain.rhs_evals}, "
          f"linear {evals[1]}, quadratic {evals[3]}, cubic {evals[4]}, learned {evals[2]}")
    # cubic extrapolation strictly saves evaluations over the plain predictor,
    # and none of the higher orders costs more than the lower
    assert evals[4] < plain.rhs_evals and evals[4] <= evals[3] <= evals[1]
    # the network trained on the residual of cubic extrapolation
    sc3 = fit_scales(snaps, cubic=True)
    X3, T3 = make_dataset(flow, snaps, sc3, 20000, seed=1)
    assert X3.shape == (9 * 20000, 16)  # 13 terms: 9 windows of five
    net3 = Net(32, seed=0)
    train(net3, X3, T3, iters=300, batch=4096, lr=3e-3)
    loss3, base3 = mean_loss(net3, X3, T3)
    g = run_guessed(flow, u_mid.clone(), n, mode=2, net=net3, scales=sc3)
    d = rel(g.state.cpu().numpy(), plain.state.cpu().numpy())
    print(f"MEASURED cubic-residual net loss {loss3:.3g} (zero {base3:.3g}); passes {g.passes}")
    assert g.converged and len(g.passes) == n and d <= (1e-8 if precision == "fp64" else 1e-7), (g.passes, d)


@pytest.mark.parametrize("precision,mode,cubic", [("fp64", 3, False), ("mixed", 2, False), ("mixed", 4, False),
                                                  ("mixed", 2, True)])
def test_guessed_graph_steps_match_eager(precision, mode, cubic):
    """Device-decided guessed steps replayed from captured graphs equal the
    eager host loop bit for bit, with the same pass counts (quadratic and
    cubic bases; three histories shifted in place for the cubic ones)."""
    from dnsjax_amd.learned import GuessedGraphStepper, Net, collect, fit_scales, make_dataset, run_guessed, train

    flow = _periodic(32, 60.0, precision, dt=0.01)
    u0 = flow.init_state(seed=7, rms=0.1)
    snaps, u_mid = collect(flow, u0, 8)
    sc = fit_scales(snaps, cubic=cubic)
    X, T = make_dataset(flow, snaps, sc, 5000, seed=2)
    net = Net(32, seed=1)
    train(net, X, T, iters=50, batch=2048)
    n = 7
    ref = []
    eager = run_guessed(flow, u_mid.clone(), n, mode=mode, net=net, scales=sc,
                        callback=lambda i, s: ref.append(s.clone()))
    gs = GuessedGraphStepper(flow, u_mid, mode=mode, net=net, scales=sc)
    torch.cuda.synchronize()
    start = 3 if (mode == 4 or cubic) else 2
    assert gs.cubic == (start == 3)
    assert torch.equal(gs.state, ref[start - 1])  # the eager start-up steps
    for i in range(start, n):
        gs.step(1)
        torch.cuda.synchronize()
        _ = flow.get_rhs(ref[i])  # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(gs.state, ref[i]), f"step {i} differs"
    assert gs.passes() == eager.passes
    st = gs.step_stats()
    assert st["steps"] == n and st["unconverged"] == 0 and st["nan"] == 0


def _walls(nx, ny, nz, precision="fp64", dt=0.005):
    from dnsjax_amd.flows import PlaneCouetteFlow

    return PlaneCouetteFlow(Params(system="plane-couette", nx=nx, ny=ny, nz=nz, lx=2 * math.pi, lz=math.pi,
                                   re=400.0, dt=dt, fd_order=4, corrector_tolerance=1e-10,
                                   max_corrector_iterations=20, precision=precision, tile_cols=16))


def _features_on_device(flow, seq, idx, scales):
    nat = flow._native
    X = torch.empty((len(idx), 16), dtype=torch.float32, device=nat.device)
    T = torch.empty((len(idx), 6), dtype=torch.float32, device=nat.device)
    nat.mlp_features(*[flow.to_device(a) for a in seq], torch.from_numpy(idx).to(nat.device), *scales, X, T)
    return X.cpu().numpy().astype(np.float64), T.cpu().numpy().astype(np.float64)


def test_features_match_oracle_periodic_and_walls():
    """Training samples (features and weighted third-difference targets) of
    both layouts against the oracle, at every mode / point."""
    from oracle import walls as W
    from dnsjax_amd.fields import periodic_field, wall_field

    n, L = 16, 2 * math.pi
    flow = _periodic(n, 40.0)
    of = P.Flow(P.PeriodicParams(nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, re=40.0, dt=0.01, system="kolmogorov"))
    seq = [P.get_rhs(periodic_field(n, n, n, L, L, L, 0.1, s), of) for s in (3, 4, 5, 6)]
    idx = np.arange(seq[0][0].size)
    sc = (2.0, 3.0, 0.25)
    X, T = _features_on_device(flow, seq, idx, sc)
    wX, wT = M.features(*seq, idx, of, *sc)
    assert np.allclose(X, wX, rtol=2e-6, atol=2e-6 * np.max(np.abs(wX)))
    assert np.allclose(T, wT, rtol=2e-6, atol=2e-6 * np.max(np.abs(wT)))
    # cubic-residual targets with a fifth, older term
    n_m3 = P.get_rhs(periodic_field(n, n, n, L, L, L, 0.1, 2), of)
    nat = flow._native
    X3 = torch.empty((len(idx), 16), dtype=torch.float32, device=nat.device)
    T3 = torch.empty((len(idx), 6), dtype=torch.float32, device=nat.device)
    nat.mlp_features(*[flow.to_device(a) for a in seq], torch.from_numpy(idx).to(nat.device), *sc, X3, T3,
                     n_m3=flow.to_device(n_m3))
    _, wT3 = M.features(*seq, idx, of, *sc, n_m3=n_m3)
    T3 = T3.cpu().numpy().astype(np.float64)
    assert np.allclose(T3, wT3, rtol=2e-6, atol=2e-6 * np.max(np.abs(wT3)))
    assert np.allclose(X3.cpu().numpy(), X, rtol=0, atol=0)  # the features do not change
    wflow = _walls(8, 17, 8, dt=0.01)
    wf = W.WallFlow(W.WallParams(nx=8, ny=17, nz=8, fd_order=4, re=400.0, dt=0.01))
    wseq = [W.get_rhs(wall_field(8, 17, 8, wf.p.lx, wf.p.lz, 0.05, seed=s), wf) for s in (3, 4, 5, 6)]
    widx = np.arange(wseq[0][0].size)
    X, T = _features_on_device(wflow, wseq, widx, sc)
    wX, wT = M.features_walls(*wseq, widx, wf, *sc)
    assert np.allclose(X, wX, rtol=2e-6, atol=2e-6 * np.max(np.abs(wX)))
    assert np.allclose(T, wT, rtol=2e-6, atol=2e-6 * np.max(np.abs(wT)))


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_walls_guess_matches_oracle(mode):
    from oracle import walls as W
    from dnsjax_amd.fields import wall_field

    flow = _walls(8, 17, 8, dt=0.01)
    of = W.WallFlow(W.WallParams(nx=8, ny=17, nz=8, fd_order=4, re=400.0, dt=0.01))
    u = wall_field(8, 17, 8, of.p.lx, of.p.lz, 0.05, seed=3)
    h1 = W.get_rhs(wall_field(8, 17, 8, of.p.lx, of.p.lz, 0.05, seed=4), of).astype(np.complex64)
    h2 = W.get_rhs(wall_field(8, 17, 8, of.p.lx, of.p.lz, 0.05, seed=5), of).astype(np.complex64)
    params = _random_params(32, 2)
    scales = (1.0, 3.0, 0.05)
    want_pred, want_rhs, want_hist = M.guess_walls(u, h1.astype(np.complex128), h2.astype(np.complex128), of,
                                                   mode, params, scales)
    nat = flow._native
    dev = lambda h: torch.from_numpy(np.stack([h.real, h.imag], -1).astype(np.float32)).to(nat.device)  # noqa: E731
    d1, d2 = dev(h1).contiguous(), dev(h2).contiguous()
    pred, rhs = nat.empty(), nat.empty()
    net = _net_with(32, params)
    net.pack()
    nat.guess(flow.to_device(u), pred, rhs, d1, d2, mode, net.operands, scales)
    if mode == 2:  # bf16 network: the residual it adds is held to bf16 accuracy
        n0 = W.get_rhs(u, of)
        _, ext3, _ = M.guess_walls(u, h1.astype(np.complex128), h2.astype(np.complex128), of, 3)
        assert rel(rhs.cpu().numpy() - ext3, want_rhs - ext3) <= 5e-3
        assert rel(rhs.cpu().numpy(), want_rhs) <= 1e-4
        assert rel(pred.cpu().numpy(), want_pred) <= 1e-4
        assert np.max(np.abs(want_rhs - ext3)) > 1e-3 * np.max(np.abs(n0))  # the net did contribute
    else:
        assert rel(rhs.cpu().numpy(), want_rhs) <= 1e-12
        assert rel(pred.cpu().numpy(), want_pred) <= 1e-11  # IMM tolerance
    h = d2.cpu().numpy().astype(np.float64)
    w = np.stack([want_hist.real, want_hist.imag], -1)
    assert np.max(np.abs(h - w)) <= 1e-7 * np.max(np.abs(w))


def test_walls_guessed_runs_and_graph_steps():
    """Plane Couette 16 x 33 x 16: mode 0 reproduces the plain driver bit for
    bit; quadratic extrapolation stays on its trajectory with no more RHS
    evaluations; device-decided guessed graph steps equal the eager loop."""
    from dnsjax_amd.driver import run
    from dnsjax_amd.fields import wall_field
    from dnsjax_amd.learned import GuessedGraphStepper, run_guessed

    flow = _walls(16, 33, 16)
    u0 = flow.to_device(wall_field(16, 33, 16, 2 * math.pi, math.pi, 0.05, seed=2))
    n = 8
    plain = run(flow, u0.clone(), n)
    assert plain.converged
    g0 = run_guessed(flow, u0.clone(), n, mode=0)
    assert torch.equal(g0.state, plain.state) and g0.passes == [c + 1 for c in plain.counts]
    ref = []
    g3 = run_guessed(flow, u0.clone(), n, mode=3, callback=lambda i, s: ref.append(s.clone()))
    assert g3.converged and len(g3.passes) == n
    assert rel(g3.state.cpu().numpy(), plain.state.cpu().numpy()) <= 1e-8
    assert g3.rhs_evals <= plain.rhs_evals, (g3.passes, plain.counts)
    gs = GuessedGraphStepper(flow, u0, mode=3)
    torch.cuda.synchronize()
    assert torch.equal(gs.state, ref[1])
    for i in range(2, n):
        gs.step(1)
        torch.cuda.synchronize()
        assert torch.equal(gs.state, ref[i]), f"step {i} differs"
    assert gs.passes() == g3.passes


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_walls_learned_runs_and_graph_steps(precision):
    """Plane Couette 16 x 33 x 16 with a network trained on its own history:
    the learned run stays on the plain trajectory to the tolerance and its
    device-decided graph steps equal the eager loop bit for bit."""
    from dnsjax_amd.driver import run
    from dnsjax_amd.fields import wall_field
    from dnsjax_amd.learned import (GuessedGraphStepper, Net, collect, fit_scales, make_dataset, mean_loss,
                                    run_guessed, train)

    flow = _walls(16, 33, 16, precision=precision)
    u0 = flow.to_device(wall_field(16, 33, 16, 2 * math.pi, math.pi, 0.05, seed=2))
    snaps, u_mid = collect(flow, u0.clone(), 10)
    sc = fit_scales(snaps)
    X, T = make_dataset(flow, snaps, sc, 8000, seed=1)
    net = Net(32, seed=0)
    train(net, X, T, iters=200, batch=2048, lr=3e-3)
    loss, zero = mean_loss(net, X, T)
    assert loss < zero  # the network explains part of the quadratic residual
    n = 8
    plain = run(flow, u_mid.clone(), n)
    ref = []
    g2 = run_guessed(flow, u_mid.clone(), n, mode=2, net=net, scales=sc, callback=lambda i, s: ref.append(s.clone()))
    assert plain.converged and g2.converged and len(g2.passes) == n
    assert rel(g2.state.cpu().numpy(), plain.state.cpu().numpy()) <= (1e-8 if precision == "fp64" else 1e-6)
    gs = GuessedGraphStepper(flow, u_mid, mode=2, net=net, scales=sc)
    torch.cuda.synchronize()
    assert torch.equal(gs.state, ref[1])
    for i in range(2, n):
        gs.step(1)
        torch.cuda.synchronize()
        assert torch.equal(gs.state, ref[i]), f"step {i} differs"
    assert gs.passes() == g2.passes



def test_walls_network_gain_over_extrapolation():
    """A/B on plane Couette 16 x 33 x 16 at the reference's default
    tolerance (1e-5): the network trained on the solver's own history
    explains most of the held-out residual of quadratic extrapolation, and
    the run started from its guess needs no more RHS + IMM evaluations than
    the run started from quadratic extrapolation over the same steps."""
    from dnsjax_amd.fields import wall_field
    from dnsjax_amd.flows import PlaneCouetteFlow
    from dnsjax_amd.learned import Net, collect, fit_scales, make_dataset, mean_loss, run_guessed, train

    flow = PlaneCouetteFlow(Params(system="plane-couette", nx=16, ny=33, nz=16, lx=2 * math.pi, lz=math.pi,
                                   re=400.0, dt=0.01, fd_order=4, corrector_tolerance=1e-5,
                                   max_corrector_iterations=10, precision="mixed", tile_cols=16))
    u0 = flow.to_device(wall_field(16, 33, 16, 2 * math.pi, math.pi, 0.05, seed=2))
    snaps, u_mid = collect(flow, u0, 24)
    sc = fit_scales(snaps)
    X, T = make_dataset(flow, snaps, sc, 20000, seed=1)
    perm = torch.randperm(X.shape[0], generator=torch.Generator().manual_seed(3)).to(X.device)
    nv = 40000
    Xv, Tv, X, T = X[perm[:nv]], T[perm[:nv]], X[perm[nv:]], T[perm[nv:]]
    net = Net(32, seed=0)
    train(net, X, T, iters=600, batch=4096, lr=3e-3)
    loss, base = mean_loss(net, Xv, Tv)
    n = 16
    quad = run_guessed(flow, u_mid.clone(), n, mode=3)
    learned = run_guessed(flow, u_mid.clone(), n, mode=2, net=net, scales=sc)
    print(f"held-out loss {loss:.4g} (without the net {base:.4g}); passes quadratic {quad.passes}, "
          f"learned {learned.passes}")
    assert loss < 0.5 * base, (loss, base)
    assert quad.converged and learned.converged
    assert learned.rhs_evals <= quad.rhs_evals, (learned.passes, quad.passes)
