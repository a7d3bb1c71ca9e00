"""The learned corrector guess strictly saves RHS evaluations over both the
plain predictor and quadratic extrapolation at the plane Couette 64 x 128 x 64
workload (p = 4, Re 400, dt 0.01, seeded perturbation RMS 0.05, seed 2,
corrector tolerance 1e-5, the reference's default, flows/plane_couette.py:
507-540 for the guess's place in the step): after 2 FP64 steps, a 30-step
mixed run records the history the per-point network is trained on, and the
three forms then run the same 20 steps from the state that run ends in."""

import math

import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def test_learned_guess_saves_evaluations_over_plain_and_quadratic():
    from dnsjax_amd.driver import run
    from dnsjax_amd.flows import PlaneCouetteFlow
    from dnsjax_amd.learned import Net, collect, fit_scales, make_dataset, mean_loss, run_guessed, train
    from dnsjax_amd.params import Params

    def mk(prec):
        return PlaneCouetteFlow(Params(system="plane-couette", nx=64, ny=128, nz=64, lx=2 * math.pi, lz=math.pi,
                                       re=400.0, dt=0.01, fd_order=4, corrector_tolerance=1e-5,
                                       max_corrector_iterations=10, precision=prec))

    flow, mixed = mk("fp64"), mk("mixed")
    state = run(flow, flow.init_state(seed=2, rms=0.05), 2).state
    snaps, state = collect(mixed, state, 30)
    sc = fit_scales(snaps)
    X, T = make_dataset(mixed, snaps, sc, 100000, seed=1)
    del snaps
    nv = 100000
    Xv, Tv, X, T = X[-nv:], T[-nv:], X[:-nv], T[:-nv]
    net = Net(32, seed=0)
    train(net, X, T, iters=1500, batch=16384, lr=3e-3)
    loss, base = mean_loss(net, Xv, Tv)
    del X, T, Xv, Tv
    assert loss < 0.1 * base  # held-out residual of quadratic extrapolation

    n = 20
    ref = run(flow, state.clone(), n)
    plain = run(mixed, state.clone(), n)
    quad = run_guessed(mixed, state.clone(), n, mode=3)
    learned = run_guessed(mixed, state.clone(), n, mode=2, net=net, scales=sc)
    for r in (ref, plain, quad, learned):
        assert r.converged
    e_ref = float(flow.get_stats(ref.state)["E'"])
    e_l = float(flow.get_stats(learned.state)["E'"])
    print(f"MEASURED channel 64x128x64 RHS evaluations over {n} steps: plain {plain.rhs_evals}, "
          f"quadratic {quad.rhs_evals}, learned {learned.rhs_evals}; held-out loss {loss:.3g} (without the "
          f"network {base:.3g}); learned end energy vs FP64 {abs(e_l - e_ref) / e_ref:.2e}")
    assert learned.rhs_evals < quad.rhs_evals < plain.rhs_evals
    # a step stops just under the tolerance 1e-5 in each form, so the end
    # states differ at the tolerance's order, not at FP32's
    assert abs(e_l - e_ref) <= 1e-3 * e_ref
