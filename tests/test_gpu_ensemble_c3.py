"""The batched-ensemble configuration at its real workload: 64 Kolmogorov
trajectories, 16 each at 32^3, 48^3, 64^3 and 96^3, Re drawn uniformly in
[20, 80] (seed 3), dt 0.01, seeded spectra RMS 0.05 (seeds 300 + i), mixed
form, tolerance 1e-10, advanced by one captured graph launch per step.
Every member equals its own eager run bit for bit (states and counts), and
the first step of a 64^3 and a 96^3 member matches the FP64 oracle's to the
mixed form's bound."""

import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from oracle import periodic as P  # noqa: E402
from oracle import stepper as S  # noqa: E402

SIZES = [32] * 16 + [48] * 16 + [64] * 16 + [96] * 16


def _flows(res, pinned=None):
    """One flow per member; with `pinned` ({size: (tile_cols, variants)}),
    every member of a size runs that size's kernel variants."""
    from dnsjax_amd.flows import TriplyPeriodicFlow
    from dnsjax_amd.params import kolmogorov_2pi

    out = []
    for n, r in zip(SIZES, res):
        kw = {} if pinned is None else dict(zip(("tile_cols", "variants"), pinned[n]))
        out.append(TriplyPeriodicFlow(kolmogorov_2pi(n, float(r), 0.01, corrector_tolerance=1e-10,
                                                     max_corrector_iterations=10, precision="mixed", **kw)))
    return out


@pytest.mark.parametrize("batched", [False, True], ids=["branches", "batched"])
def test_64_member_ensemble_matches_eager_runs_and_the_oracle(batched):
    """`batched`: each size class advances by one launch per kernel over its
    16 members, instead of one branch per member.  In both forms every
    member of a size runs the autotuner's choice of kernel variants for
    that size: the variants agree to rounding, not bit for bit, and a
    second context left to its own autotuning can settle a near tie the
    other way."""
    from dnsjax_amd.driver import run
    from dnsjax_amd.flows import TriplyPeriodicFlow
    from dnsjax_amd.graph import Ensemble
    from dnsjax_amd.params import kolmogorov_2pi

    res = np.random.default_rng(3).uniform(20.0, 80.0, len(SIZES))
    n = 3
    pinned = {}
    for s in (32, 48, 64, 96):  # the autotuner's choice per size
        f = TriplyPeriodicFlow(kolmogorov_2pi(s, 50.0, 0.01, corrector_tolerance=1e-10,
                                              max_corrector_iterations=10, precision="mixed"))
        pinned[s] = f._native.variants()
        del f
    eager, graphed = _flows(res, pinned), _flows(res, pinned)
    states0 = [f.init_state(seed=300 + i, rms=0.05) for i, f in enumerate(eager)]
    refs = []
    for f, u0 in zip(eager, states0):
        traj = []
        r = run(f, u0.clone(), n, callback=lambda i, s: traj.append(s.clone()))
        assert r.converged and r.escalations == 0
        refs.append((traj, r.counts))
    slots = [min(max(c) + 1, 10) for _, c in refs]
    ens = Ensemble(graphed, states0, n_streams=8, slots=slots, capacity=n, batched=batched)
    if batched:
        assert sorted(len(g) for g in ens.groups()) == [16, 16, 16, 16]
        assert all(pl.members == 16 for pl in ens.plans[0])
    ens.step(n)
    torch.cuda.synchronize()
    for m, (traj, counts) in enumerate(refs):
        assert torch.equal(ens.states()[m], traj[-1]), m
        assert ens.counts()[m] == counts, m
    assert all(x["unconverged"] == 0 for x in ens.step_stats())
    # the first step of one 64^3 and one 96^3 member against the oracle
    for m in (32, 48):
        nn = SIZES[m]
        L = 2 * math.pi
        of = P.Flow(P.PeriodicParams(nx=nn, ny=nn, nz=nn, lx=L, ly=L, lz=L, re=float(res[m]), dt=0.01))
        pc, ic = S.make_stepper(P.get_rhs, P.predict, P.correct, P.norm)
        want, counts, _ = S.run(states0[m].cpu().numpy(), 1, 1e-10, 10, lambda s: pc(s, of),
                                lambda a, b, c: ic(a, b, c, of), lambda s: P.correct_velocity(s, of))
        got = refs[m][0][0].cpu().numpy()
        e = np.linalg.norm(got - want) / np.linalg.norm(want)
        print(f"MEASURED ensemble member {m} ({nn}^3, Re {res[m]:.1f}) first step vs oracle rel {e:.2e}, "
              f"counts {refs[m][1][0]} (oracle {counts[0]})")
        assert e < 1e-7
        assert abs(refs[m][1][0] - counts[0]) <= 1
