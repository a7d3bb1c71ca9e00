"""The mixed form's split-precision evaluations and its escalation to FP64.

In the mixed form the step's first corrector evaluation records its state
u_ref, and every later evaluation is computed as N(u_ref) + P[dN], with only
the increment dN = (U + u_ref) x w_d + d x (W_ref + w_d), d = u - u_ref,
going through the FP32 transforms.  The
FP32 rounding then scales with the increment, so the difference of two
evaluations, which is what the corrector's error measures, is accurate to
FP32 relative to that difference, and the corrector converges to tight
tolerances as the FP64 one does.  A step whose FP32 passes still stop
converging escalates to FP64 passes; the host loop and the device step do
this identically.
"""

import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dnsjax_amd.params import Params  # noqa: E402
from oracle import periodic as P  # noqa: E402


def flow_of(n, precision, tol=1e-10, plain_mixed=False, walls=False, **kw):
    """Fixed kernel variants (tile_cols 16, variants 0): two contexts then
    compute bit for bit alike."""
    if walls:
        from dnsjax_amd.flows import PlaneCouetteFlow

        return PlaneCouetteFlow(Params(system="plane-couette", nx=n, ny=2 * n + 1, nz=n, lx=2 * math.pi,
                                       lz=math.pi, re=400.0, dt=0.01, fd_order=4, corrector_tolerance=tol,
                                       max_corrector_iterations=30, precision=precision,
                                       plain_mixed=plain_mixed, tile_cols=16, **kw))
    from dnsjax_amd.flows import TriplyPeriodicFlow

    L = 2 * math.pi
    return TriplyPeriodicFlow(Params(system="kolmogorov", nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, re=100.0, dt=0.005,
                                     corrector_tolerance=tol, max_corrector_iterations=10, precision=precision,
                                     plain_mixed=plain_mixed, tile_cols=16, **kw))


def pass_errors(flow, u0, passes):
    """The corrector errors of one step's predict_and_correct and `passes`
    further passes, with no stopping rule."""
    nat = flow._native
    pred, rhs, err = nat.predict_and_correct(u0)
    out = [float(err)]
    nxt = nat.empty()
    for _ in range(passes):
        nat.iterate_correction(u0, pred, rhs, nxt, err)
        rhs, nxt = nxt, rhs
        out.append(float(err))
    return out


def _rhs_at(flow, ref, u):
    """N(u) as a corrector pass computes it after a first pass at `ref`:
    an increment over ref in the split form."""
    nat = flow._native
    hist = [torch.zeros(nat.shape + (2,), dtype=torch.float32, device=nat.device) for _ in range(2)]
    nat.guess(ref, nat.empty(), nat.empty(), hist[0], hist[1], 0)  # a new step: no reference yet
    first = ref.clone()
    nat.iterate_correction(ref, first, nat.empty().zero_(), nat.empty())  # records at ref
    pred = u.clone()
    rhs_next = nat.empty()
    nat.iterate_correction(ref, pred, nat.empty().zero_(), rhs_next)
    return rhs_next


@pytest.mark.parametrize("n", [32, 64])
def test_split_evaluations_resolve_small_differences(n):
    """Two states 1e-6 apart (1e-3 from the reference state): the difference
    of their split evaluations matches the FP64 oracle's difference to 1e-3
    relative (FP32 rounding of increments 1e3 times the difference), while
    two full FP32 evaluations miss it by over a hundred times more."""
    L = 2 * math.pi
    op = P.PeriodicParams(nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, re=100.0, dt=0.005)
    of = P.Flow(op)
    split, plain = flow_of(n, "mixed"), flow_of(n, "mixed", plain_mixed=True)
    ua = split.init_state(seed=1, rms=0.1)
    ub = ua + 1e-6 * split.init_state(seed=2, rms=0.1)
    ref = ua + 1e-3 * split.init_state(seed=3, rms=0.1)  # the step's first state, nearby
    want = P.get_rhs(ua.cpu().numpy(), of) - P.get_rhs(ub.cpu().numpy(), of)
    got = (_rhs_at(split, ref, ua) - _rhs_at(split, ref, ub)).cpu().numpy()
    plain_diff = (plain.get_rhs(ua) - plain.get_rhs(ub)).cpu().numpy()
    e_split = np.linalg.norm(got - want) / np.linalg.norm(want)
    e_plain = np.linalg.norm(plain_diff - want) / np.linalg.norm(want)
    # and a single split evaluation is as accurate as a full FP32 one
    n_split = _rhs_at(split, ref, ua).cpu().numpy()
    n_want = P.get_rhs(ua.cpu().numpy(), of)
    e_one = np.linalg.norm(n_split - n_want) / np.linalg.norm(n_want)
    print(f"MEASURED split difference rel {e_split:.2e}, full FP32 difference rel {e_plain:.2e}, "
          f"one split evaluation rel {e_one:.2e}")
    assert e_split < 1e-3
    assert e_plain > 100 * e_split
    assert e_one < 1e-5


@pytest.mark.parametrize("walls", [False, True])
def test_split_form_resolves_the_corrector_far_below_full_fp32(walls):
    """Corrector errors of one step with no stopping rule (32^3 periodic;
    16 x 33 x 16 plane Couette): the full FP32 evaluations stall at their
    resolution, the split ones keep converging to a floor at least 20 times
    lower; a tolerance between the two floors is then met by the split form
    without escalating, in the FP64 path's number of passes."""
    from dnsjax_amd.driver import run

    n = 16 if walls else 32
    passes = 24 if walls else 8
    f64, split, plain = (flow_of(n, p, walls=walls, plain_mixed=q)
                         for p, q in (("fp64", False), ("mixed", False), ("mixed", True)))
    u0 = f64.init_state(seed=2 if walls else 1, rms=0.05 if walls else 0.1)
    e64, esp, epl = (pass_errors(f, u0, passes) for f in (f64, split, plain))
    floor = lambda e: min(e[-4:])  # noqa: E731
    print(f"MEASURED corrector floors ({'walls' if walls else 'periodic'}): fp64 {floor(e64):.2e}, "
          f"split {floor(esp):.2e}, full FP32 {floor(epl):.2e}; split errors {['%.1e' % x for x in esp]}")
    assert floor(esp) <= floor(epl) / 20
    tol = math.sqrt(floor(esp) * floor(epl))
    a = run(flow_of(n, "fp64", tol=tol, walls=walls), u0.clone(), 4)
    b = run(flow_of(n, "mixed", tol=tol, walls=walls), u0.clone(), 4)
    print(f"MEASURED tolerance {tol:.2e}: fp64 counts {a.counts}, split counts {b.counts}")
    assert a.converged and b.converged and b.escalations == 0 and max(b.errors) <= tol
    assert sum(abs(x - y) for x, y in zip(a.counts, b.counts)) <= 2


def test_escalation_on_the_host_and_on_the_device_agree():
    """Full FP32 evaluations (plain_mixed) at tolerance 2e-13 stall above it
    (near 1.3e-10): the steps escalate to FP64 passes and still converge.
    The first FP64 pass still differences against the last FP32 evaluation,
    so this tolerance takes three; the device-decided steps with three FP64
    slots take the same passes and give the same states, bit for bit, and
    with two slots they stop one pass short and are counted unconverged."""
    from dnsjax_amd.driver import run
    from dnsjax_amd.graph import GraphStepper

    tol = 2e-13
    host, dev = flow_of(32, "mixed", tol=tol, plain_mixed=True), flow_of(32, "mixed", tol=tol, plain_mixed=True)
    u0 = host.init_state(seed=1, rms=0.1)
    n = 4
    states = []
    r = run(host, u0.clone(), n, callback=lambda i, s: states.append(s.clone()))
    print(f"MEASURED escalation host counts {r.counts}, escalations {r.escalations}, errors {r.errors}")
    assert r.converged and r.escalations >= 1 and max(r.errors) <= tol
    gs = GraphStepper(dev, u0, slots=10, esc_slots=3)
    for i in range(n):
        gs.step()
        torch.cuda.synchronize()
        assert torch.equal(gs.state, states[i]), i
    st = gs.step_stats()[0]
    assert gs.counts()[0] == r.counts
    assert st["unconverged"] == 0 and st["escalated"] == r.escalations
    short = GraphStepper(flow_of(32, "mixed", tol=tol, plain_mixed=True), u0, slots=10, esc_slots=2)
    short.step()
    torch.cuda.synchronize()
    assert short.step_stats()[0]["unconverged"] == 1 and short.counts()[0] == [r.counts[0] - 1]


@pytest.mark.parametrize("walls", [True, False], ids=["walls", "periodic"])
def test_recentering_on_the_host_and_on_the_device_agree(walls):
    """Re-centring of the split reference: a step whose predict-and-correct
    error exceeds recenter_far (here 2e3 tol) records a new reference at its
    first pass that starts from an error of at most recenter_near (1e3 tol),
    and its later passes are increments over that one.  The host loop and
    the device-decided steps re-centre at the same passes, giving the same
    states and counts bit for bit, and the run stays on the FP64 trajectory.
    Plane Couette keeps the FP64 run's counts.  The periodic corrector
    contracts about 70x per pass, so the rounding of the new reference's
    full FP32 evaluation (the plain FP32 floor) still shows near the
    tolerance: re-centring there costs up to one pass per step, which is
    why it is off by default for periodic flows."""
    from dnsjax_amd.driver import run
    from dnsjax_amd.fields import wall_field
    from dnsjax_amd.graph import GraphStepper

    kw = dict(recenter_far=2e3, recenter_near=1e3)
    host, dev = flow_of(16, "mixed", walls=walls, **kw), flow_of(16, "mixed", walls=walls, **kw)
    ref = flow_of(16, "fp64", walls=walls)
    if walls:
        u0 = host.to_device(wall_field(16, 33, 16, 2 * math.pi, math.pi, 0.05, seed=2))
    else:
        u0 = host.init_state(seed=1, rms=0.1)
    n = 5
    states = []
    r = run(host, u0.clone(), n, callback=lambda i, s: states.append(s.clone()))
    f = run(ref, u0.clone(), n)
    print(f"MEASURED recentering walls={walls}: counts {r.counts} (fp64 {f.counts}), re-recorded {r.rerecorded}, "
          f"escalations {r.escalations}, errors {['%.1e' % e for e in r.errors]}")
    assert r.converged and r.rerecorded >= 1 and r.escalations == 0
    if walls:
        assert sum(abs(a - b) for a, b in zip(r.counts, f.counts)) <= 1
    else:
        assert all(0 <= a - b <= 1 for a, b in zip(r.counts, f.counts))
    d = float(torch.linalg.vector_norm(r.state - f.state) / torch.linalg.vector_norm(f.state))
    assert d < 1e-6, d
    gs = GraphStepper(dev, u0, slots=host.params.max_corrector_iterations)
    for i in range(n):
        gs.step()
        torch.cuda.synchronize()
        assert torch.equal(gs.state, states[i]), i
    st = gs.step_stats()[0]
    assert gs.counts()[0] == r.counts
    assert st["unconverged"] == 0 and st["rerecorded"] == r.rerecorded and st["escalated"] == 0


def test_recentering_of_guessed_steps_on_the_host_and_on_the_device_agree():
    """Plane Couette steps started from quadratic extrapolation record their
    reference at the first pass; when that pass's error exceeds
    recenter_far, the reference is re-recorded as after predict-and-correct.
    The host loop and the device-decided guessed steps (captured graphs)
    agree bit for bit, with the same re-recordings."""
    from dnsjax_amd.fields import wall_field
    from dnsjax_amd.learned import GuessedGraphStepper, run_guessed

    kw = dict(recenter_far=2e3, recenter_near=1e3)
    host, dev = flow_of(16, "mixed", walls=True, **kw), flow_of(16, "mixed", walls=True, **kw)
    u0 = host.to_device(wall_field(16, 33, 16, 2 * math.pi, math.pi, 0.05, seed=2))
    n = 6
    ref = []
    g = run_guessed(host, u0.clone(), n, mode=3, callback=lambda i, s: ref.append(s.clone()))
    print(f"MEASURED guessed recentering: passes {g.passes}, re-recorded {g.rerecorded}, escalations {g.escalations}")
    assert g.converged and g.rerecorded >= 1 and g.escalations == 0
    gs = GuessedGraphStepper(dev, u0, mode=3)
    torch.cuda.synchronize()
    assert torch.equal(gs.state, ref[1])
    for i in range(2, n):
        gs.step(1)
        torch.cuda.synchronize()
        assert torch.equal(gs.state, ref[i]), i
    st = gs.step_stats()
    assert gs.passes() == g.passes and st["unconverged"] == 0 and st["rerecorded"] == g.rerecorded
