"""Captured time steps: one hipGraph launch per step, and ensembles.

``GraphStepper`` captures ``dnsmi_step_device`` (a whole time step with the
corrector loop decided on the device) into two hipGraphs that ping-pong
between two state buffers, so a run of n steps is n graph launches and one
synchronisation at the end.  Iteration counts are recorded on the device
and read once.

``Ensemble`` puts many independent trajectories (any mix of sizes, Reynolds
numbers and precisions, each with its own autotuned kernel variants) into
one captured graph per step parity: each member's step is a branch forked
from the capture stream onto one of a few side streams, so members run
concurrently and the whole ensemble advances by one graph launch.

The reference has no counterpart (it runs one trajectory and reads the
corrector error on the host every pass, ``__main__.py:161-191``); the
per-member physics and iteration rule are unchanged.
"""

from __future__ import annotations

import torch

from .params import PERIODIC_SYSTEMS


class _Solo:
    """A member outside every batch class, enqueued on its own."""

    def __init__(self, member, parity):
        self.member, self.parity = member, parity

    def enqueue(self):
        self.member.enqueue(self.parity)


class _Member:
    def __init__(self, flow, state, slots, capacity, esc_slots=None):
        nat = flow._native
        self.flow = flow
        self.nat = nat
        self.esc_slots = esc_slots
        p = flow.params
        self.slots = p.max_corrector_iterations if slots is None else min(slots, p.max_corrector_iterations)
        self.bufs = [state.detach().clone().contiguous(), nat.empty()]
        self.rhs = [nat.empty(), nat.empty()]
        self.counts = torch.zeros(capacity, dtype=torch.int32, device=nat.device)
        nat.step_reset()

    def enqueue(self, parity: int) -> None:
        self.nat.step_device(self.bufs[parity], self.bufs[1 - parity], self.rhs[0], self.rhs[1], self.slots,
                             self.counts, esc_slots=self.esc_slots)


class Ensemble:
    """Trajectories advanced together by one captured graph launch per step.

    `slots` is the corrector slots per captured step: one value for every
    member, or one per member (a member's skipped slots still launch their
    kernels, so each can hold just what its own trajectory needs).
    `esc_slots`: the mixed form's FP64 escalation slots per step (None: the
    library default).  `batched`: the periodic members of one size,
    precision and kernel variants (``groups()``) advance by one launch per
    kernel over all of them (``dnsmi_batch_plan``; their slot count is the
    class's largest) instead of one branch per member; every member's
    results are the same bit for bit."""

    def __init__(self, flows, states, n_streams: int = 8, slots=None, capacity: int = 4096, esc_slots=None,
                 batched: bool = False):
        if len(flows) != len(states) or not flows:
            raise ValueError("need one state per flow")
        if len({id(f._native) for f in flows}) != len(flows):
            # members run concurrently and each uses its context's work buffers
            # and control block: two members may not share one flow
            raise ValueError("every member needs its own flow (native context)")
        per = list(slots) if isinstance(slots, (list, tuple)) else [slots] * len(flows)
        if len(per) != len(flows):
            raise ValueError("need one slot count per flow")
        self.members = [_Member(f, s, k, capacity, esc_slots) for f, s, k in zip(flows, states, per)]
        self.capacity = capacity
        self.parity = 0
        self.steps = 0
        self.plans = None
        if batched:
            self.plans = [self._plan(parity, esc_slots) for parity in (0, 1)]
        self.graphs = [self._capture(0, n_streams), self._capture(1, n_streams)]

    def groups(self):
        """Member indices per batch class: periodic members of one size,
        precision, kernel variants and corrector limits (members of a class
        share one launch per kernel in the batched form)."""
        out = {}
        for i, m in enumerate(self.members):
            p = m.flow.params
            key = (m.nat.shape, p.precision, m.nat.variants(), p.corrector_tolerance, p.max_corrector_iterations,
                   p.system in PERIODIC_SYSTEMS)
            out.setdefault(key, []).append(i)
        return list(out.values())

    def _plan(self, parity: int, esc_slots):
        from .native import BatchPlan

        plans = []
        for idx in self.groups():
            ms = [self.members[i] for i in idx]
            if ms[0].flow.params.system not in PERIODIC_SYSTEMS:
                plans += [_Solo(m, parity) for m in ms]  # wall-bounded members keep their own launches
                continue
            slots = max(m.slots for m in ms)
            for m in ms:
                m.slots = slots
            plans.append(BatchPlan([m.nat for m in ms], [m.bufs[parity] for m in ms],
                                   [m.bufs[1 - parity] for m in ms], [m.rhs[0] for m in ms],
                                   [m.rhs[1] for m in ms], slots, counts=[m.counts for m in ms],
                                   esc_slots=esc_slots))
        return plans

    def _capture(self, parity: int, n_streams: int):
        g = torch.cuda.CUDAGraph()
        work = self.plans[parity] if self.plans is not None else self.members
        side = [torch.cuda.Stream() for _ in range(min(n_streams, len(work)))]
        torch.cuda.synchronize()
        with torch.cuda.graph(g):
            main = torch.cuda.current_stream()
            for s in side:
                s.wait_stream(main)
            for i, w in enumerate(work):
                with torch.cuda.stream(side[i % len(side)]):
                    if self.plans is not None:
                        w.enqueue()
                    else:
                        w.enqueue(parity)
            for s in side:
                main.wait_stream(s)
        return g

    def step(self, n: int = 1) -> None:
        if self.steps + n > self.capacity:
            raise ValueError("iteration-count capacity exceeded")
        for _ in range(n):
            self.graphs[self.parity].replay()
            self.parity ^= 1
            self.steps += 1

    def states(self):
        return [m.bufs[self.parity] for m in self.members]

    def counts(self):
        return [m.counts[: self.steps].tolist() for m in self.members]

    def step_stats(self):
        return [m.nat.step_stats() for m in self.members]


class GraphStepper(Ensemble):
    """A single trajectory stepped by captured graphs (one launch per step)."""

    def __init__(self, flow, state, slots: int | None = None, capacity: int = 4096, esc_slots=None):
        super().__init__([flow], [state], n_streams=1, slots=slots, capacity=capacity, esc_slots=esc_slots)

    @property
    def state(self):
        return self.states()[0]
