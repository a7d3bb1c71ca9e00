"""Graph-captured steps and ensembles against the eager driver (GPU).

Every replay is checked: the state after each graph launch must equal,
bit for bit, the eager driver's state after the same number of steps (same
kernels and variants), and the per-step corrector counts must match.
Eager work and synchronisations run between the replays.
"""

import math

import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dnsjax_amd.params import Params  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def periodic(n, re, precision="fp64", **kw):
    from dnsjax_amd.flows import TriplyPeriodicFlow

    L = 2 * math.pi
    return TriplyPeriodicFlow(Params(system="kolmogorov", nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, re=re, dt=0.01,
                                     corrector_tolerance=1e-10, max_corrector_iterations=10,
                                     precision=precision, tile_cols=16, **kw))


def walls(nx, ny, nz, precision="fp64"):
    from dnsjax_amd.flows import PlaneCouetteFlow

    return PlaneCouetteFlow(Params(system="plane-couette", nx=nx, ny=ny, nz=nz, lx=2 * math.pi, lz=math.pi,
                                   re=400.0, dt=0.005, fd_order=4, corrector_tolerance=1e-10,
                                   max_corrector_iterations=20, precision=precision, tile_cols=16))


def eager_trajectory(flow, u0, n):
    from dnsjax_amd.driver import run

    states = []
    r = run(flow, u0.clone(), n, callback=lambda i, s: states.append(s.clone()))
    return states, r.counts


@pytest.mark.parametrize("kind", ["periodic_fp64", "periodic_mixed", "walls_fp64"])
def test_graph_stepper_replays_match_eager(kind):
    from dnsjax_amd.graph import GraphStepper

    if kind.startswith("periodic"):
        prec = kind.split("_")[1]
        mk = lambda: periodic(32, 60.0, prec)  # noqa: E731
    else:
        mk = lambda: walls(16, 33, 16)  # noqa: E731
    f_eager, f_graph = mk(), mk()
    u0 = f_eager.init_state(seed=5, rms=0.05)
    n = 6
    ref_states, ref_counts = eager_trajectory(f_eager, u0, n)
    assert len(ref_states) == n, "eager run stopped early (corrector did not converge)"
    gs = GraphStepper(f_graph, u0)
    for i in range(n):
        gs.step(1)
        torch.cuda.synchronize()
        _ = f_eager.get_rhs(ref_states[i])  # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(gs.state, ref_states[i]), f"replay {i} differs"
    assert gs.counts()[0] == ref_counts
    st = gs.step_stats()[0]
    assert st["steps"] == n and st["nan"] == 0 and st["total_iterations"] == sum(ref_counts)


def test_ensemble_members_match_single_runs():
    """Uneven sizes, different Re, mixed precision: each member equals its own eager run."""
    from dnsjax_amd.graph import Ensemble

    specs = [(16, 25.0), (32, 40.0), (48, 70.0), (16, 55.0), (32, 20.0)]
    n = 4
    flows = [periodic(sz, re, "mixed") for sz, re in specs]
    states0 = [f.init_state(seed=300 + i, rms=0.05) for i, f in enumerate(flows)]
    refs = []
    for (sz, re), u0 in zip(specs, states0):
        refs.append(eager_trajectory(periodic(sz, re, "mixed"), u0, n))
    ens = Ensemble(flows, states0, n_streams=3)
    for i in range(n):
        ens.step(1)
        torch.cuda.synchronize()
        for m, (states, _) in enumerate(refs):
            assert torch.equal(ens.states()[m], states[i]), (m, i)
    for m, (_, counts) in enumerate(refs):
        assert ens.counts()[m] == counts


def test_batched_ensemble_members_match_single_runs():
    """Batched classes (three 16^3 members, two 32^3, one 48^3, plus an FP64
    16^3 member and a plane Couette member outside every class): each member
    equals its own eager run, and a class of k members issues the launches
    of one."""
    from dnsjax_amd.graph import Ensemble
    from dnsjax_amd.native import BatchPlan

    specs = [(16, 25.0, "mixed"), (32, 40.0, "mixed"), (16, 55.0, "mixed"), (48, 70.0, "mixed"),
             (16, 35.0, "mixed"), (32, 20.0, "mixed"), (16, 45.0, "fp64")]
    n = 4
    flows = [periodic(sz, re, prec) for sz, re, prec in specs] + [walls(16, 33, 16)]
    states0 = [f.init_state(seed=320 + i, rms=0.05) for i, f in enumerate(flows)]
    refs = [eager_trajectory(periodic(sz, re, prec), u0, n) for (sz, re, prec), u0 in zip(specs, states0)]
    refs.append(eager_trajectory(walls(16, 33, 16), states0[-1], n))
    ens = Ensemble(flows, states0, n_streams=3, batched=True)
    assert sorted(len(g) for g in ens.groups()) == [1, 1, 1, 2, 3]
    plans = [w for w in ens.plans[0] if isinstance(w, BatchPlan)]
    assert sorted(pl.members for pl in plans) == [1, 1, 2, 3]
    # one launch sequence per precision, whatever the class's size and members
    assert len({pl.launches for pl in plans}) == 2 and min(pl.launches for pl in plans) > 0
    for i in range(n):
        ens.step(1)
        torch.cuda.synchronize()
        for m, (states, _) in enumerate(refs):
            assert torch.equal(ens.states()[m], states[i]), (m, i)
    for m, (_, counts) in enumerate(refs):
        assert ens.counts()[m] == counts
    assert all(x["unconverged"] == 0 for x in ens.step_stats())


def test_ensemble_with_per_member_slots():
    """One slot count per member, each one more than its own eager run's
    largest count: every member still equals its eager run, with no step
    flagged unconverged."""
    from dnsjax_amd.graph import Ensemble

    specs = [(16, 25.0), (32, 40.0), (16, 70.0)]
    n = 4
    flows = [periodic(sz, re, "mixed") for sz, re in specs]
    states0 = [f.init_state(seed=310 + i, rms=0.05) for i, f in enumerate(flows)]
    refs = [eager_trajectory(periodic(sz, re, "mixed"), u0, n) for (sz, re), u0 in zip(specs, states0)]
    slots = [max(counts) + 1 for _, counts in refs]
    ens = Ensemble(flows, states0, n_streams=2, slots=slots)
    assert [mb.slots for mb in ens.members] == slots
    ens.step(n)
    torch.cuda.synchronize()
    for m, (states, counts) in enumerate(refs):
        assert torch.equal(ens.states()[m], states[-1]), m
        assert ens.counts()[m] == counts
    assert all(x["unconverged"] == 0 for x in ens.step_stats())
    with pytest.raises(ValueError):
        Ensemble(flows, states0, slots=[3, 3])
