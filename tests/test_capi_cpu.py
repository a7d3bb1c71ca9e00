"""CPU-only checks of the boundary: the C-ABI library loads and exports every
symbol include/dnsmi.h declares; host-side logic (parameters, stepper
composition, driver loop semantics) without touching a GPU."""

import ctypes
import os
import re

import numpy as np
import math

import pytest

from dnsjax_amd import _lib
from dnsjax_amd.driver import run
from dnsjax_amd.params import Params
from dnsjax_amd.timestep import make_stepper

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "dnsmi.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dnsmi_\w+)\s*\(", txt)))


def test_header_declares_the_boundary():
    fns = _header_functions()
    for name in ("dnsmi_create", "dnsmi_destroy", "dnsmi_get_rhs", "dnsmi_predict", "dnsmi_correct",
                 "dnsmi_norm", "dnsmi_predict_and_correct", "dnsmi_iterate_correction",
                 "dnsmi_correct_velocity", "dnsmi_last_error"):
        assert name in fns


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libdnsmi.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [f for f in _header_functions() if not hasattr(lib, f)]
    assert not missing, missing
    # the ctypes table covers the same surface
    assert set(_header_functions()) <= set(_lib.SIGNATURES)


def test_config_struct_matches_the_header():
    """The ctypes DnsmiConfig has every field of dnsmi_config, in order and
    with its C type (the library reads the whole struct)."""
    txt = open(os.path.join(ROOT, "include", "dnsmi.h")).read()
    body = re.search(r"typedef struct dnsmi_config \{(.*?)\} dnsmi_config;", txt, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        ctype, names = decl.split(None, 1)
        fields += [(n.strip(), ctype) for n in names.split(",")]
    want = {"int": ctypes.c_int, "double": ctypes.c_double}
    got = [(n, t) for n, t in _lib.DnsmiConfig._fields_]
    assert [n for n, _ in got] == [n for n, _ in fields]
    assert all(t is want[c] for (_, t), (_, c) in zip(got, fields))


def test_create_rejects_null_and_reports():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libdnsmi.so not built")
    lib = _lib.load()
    rc = lib.dnsmi_create(None, None)
    assert rc == 1
    assert b"null" in lib.dnsmi_last_error(None)


def test_params_defaults_follow_reference():
    p = Params()
    assert p.ly == 4.0 and p.spec_shape == (127, 127, 64) and p.padded == (192, 192, 192)
    w = Params(system="plane-couette", nx=64, ny=128, nz=64)
    assert w.ly == 2.0 and w.spec_shape == (63, 32, 128) and w.padded == (128, 96, 96)
    with pytest.raises(ValueError):
        Params(system="pipe")
    with pytest.raises(ValueError):
        Params(precision="fp16")


def test_generic_stepper_composition_matches_reference_order():
    calls = []

    def get_rhs(s):
        calls.append("rhs")
        return s * 2.0

    def predict(s, r):
        calls.append("predict")
        return s + r

    def correct(sp, p, rp, rn):
        calls.append("correct")
        return p + 0.5 * (rn - rp), 0.5 * (rn - rp)

    def norm(c):
        calls.append("norm")
        return float(np.abs(c).sum())

    pc, ic = make_stepper(get_rhs, predict, correct, norm)
    assert pc.fused is False
    s = np.ones(3)
    pred, rhs, err = pc(s)
    assert calls == ["rhs", "predict", "rhs", "correct", "norm"]
    assert np.allclose(pred, 3 + 0.5 * (6 - 2)) and np.allclose(rhs, 6) and err == 6.0
    calls.clear()
    pred2, rhs2, err2 = ic(s, pred, rhs)
    assert calls == ["rhs", "correct", "norm"]


class _ScriptedFlow:
    """Errors follow a script, so the loop's count semantics can be checked."""

    def __init__(self, errs, tol=1e-3, max_it=3):
        self.params = Params(corrector_tolerance=tol, max_corrector_iterations=max_it)
        self.errs = list(errs)

    def predict_and_correct(self, s):
        return s, None, self.errs.pop(0)

    def iterate_correction(self, sp, s, r):
        return s, None, self.errs.pop(0)

    def correct_velocity(self, s):
        return s, None


def test_driver_count_semantics():
    # step 1: converges after 2 extra; step 2: converges at once; step 3 hits max -> stop
    f = _ScriptedFlow([1.0, 0.5, 1e-4, 1e-5, 1.0, 1.0, 1.0, 1.0, 1e-9], tol=1e-3, max_it=3)
    r = run(f, 0.0, 10)
    assert r.counts == [2, 0, 3]
    assert not r.converged
    assert f.errs == [1e-9]  # nothing ran after the non-converged step


def test_driver_detects_nan():
    f = _ScriptedFlow([float("nan")], tol=1e-3, max_it=3)
    r = run(f, 0.0, 5)
    assert r.nan and not r.converged and r.counts == [0]


class _EscalatingFlow(_ScriptedFlow):
    """Scripted errors that change once the flow escalates to FP64 passes:
    `fp32` before the switch, `fp64` after it (popped per pass)."""

    def __init__(self, fp32, fp64, tol=1e-10, max_it=10, **kw):
        super().__init__(fp32, tol=tol, max_it=max_it)
        self.params = Params(precision="mixed", corrector_tolerance=tol, max_corrector_iterations=max_it, **kw)
        self.fp64 = list(fp64)
        self.on = False
        self.switches = []

    def iterate_correction(self, sp, s, r):
        return s, None, (self.fp64 if self.on else self.errs).pop(0)

    def escalate(self, on):
        self.on = on
        self.switches.append(on)


def test_mixed_form_escalates_a_stalled_corrector_to_fp64():
    """A pass that is not below half the error two passes back, at an error
    below the window's cap, switches the step's remaining passes to FP64; the
    step still iterates to the tolerance, and the next step starts in FP32."""
    f = _EscalatingFlow([1e-5, 1e-7, 2e-9, 1.6e-10, 1.3e-10, 1.2e-10,
                         1e-5, 1e-7, 1e-9, 5e-11],
                        [2e-12])
    # step 1: each of 1e-7, 2e-9, 1.6e-10 and 1.3e-10 is below half the error
    # two passes back; 1.2e-10 > 0.5 * 1.6e-10 -> the next pass is FP64 and
    # ends the step at 2e-12 <= tol.  Step 2 converges in FP32.
    r = run(f, 0.0, 2)
    assert r.counts == [6, 3] and r.escalations == 1 and r.converged
    assert f.switches == [True, False]
    assert r.errors == [2e-12, 5e-11]
    assert not f.errs and not f.fp64


def test_escalation_needs_an_error_inside_the_window():
    """Above the cap (min(1e4 tol, 1e-8)) a slow pass is slow contraction,
    not FP32 resolution: no escalation, the reference rule alone."""
    f = _EscalatingFlow([1e-3, 6e-4, 3.5e-4, 2e-4, 1.1e-4, 6e-5, 3.2e-5, 1.7e-5, 9e-6], [], tol=1e-5)
    assert f.params.escalate_cap_value == 1e-8
    r = run(f, 0.0, 1)
    assert r.counts == [8] and r.escalations == 0 and r.converged and f.switches == []


def test_plain_form_never_escalates():
    g = _ScriptedFlow([1e-5, 1e-7, 2e-9, 1.6e-10, 1.3e-10, 1.2e-10, 1.1e-10, 1.05e-10, 1.01e-10, 1.0e-10],
                      tol=1e-10, max_it=10)
    g.params = Params(corrector_tolerance=1e-10, max_corrector_iterations=10)
    assert g.params.escalate_ratio is None
    r = run(g, 0.0, 1)
    assert r.counts == [9] and r.escalations == 0


def test_escalated_step_that_still_stalls_ends_unconverged():
    """Escalation does not relax the tolerance: a step whose FP64 passes do
    not reach it either stops the run after max_corrector_iterations, as the
    reference does."""
    f = _EscalatingFlow([1e-7, 2e-9, 1.6e-10, 1.5e-10, 1.45e-10], [1.4e-10] * 4, max_it=8)
    r = run(f, 0.0, 3)
    assert r.counts == [8] and not r.converged and r.escalations == 1


class _RecenterFlow(_EscalatingFlow):
    """Scripted errors of a mixed plane Couette flow; records before which
    pass (counted over the run) the split reference was re-recorded."""

    def __init__(self, errs, **kw):
        super().__init__(errs, [], system="plane-couette", **kw)
        self.passes = 0
        self.rerecords = []
        flow = self

        class _Native:
            def split_rerecord(self):
                flow.rerecords.append(flow.passes + 1)

        self._native = _Native()

    def iterate_correction(self, sp, s, r):
        self.passes += 1
        return super().iterate_correction(sp, s, r)


def test_recentering_rule_of_the_host_loop():
    """Wall-bounded mixed flows re-centre by default (recenter_far 1e6,
    recenter_near 5e5 times the tolerance): a step whose predict-and-correct
    error is above 1e-4 records a new reference at the first pass that
    starts from an error of at most 5e-5; a step starting below 1e-4 keeps
    its reference; periodic flows never re-centre by default."""
    f = _RecenterFlow([3.8e-4, 3.8e-4, 8.6e-7, 4.3e-7, 1.2e-8, 3.8e-9, 1.3e-10, 3.3e-11,
                       5e-5, 4e-5, 1e-7, 1e-9, 5e-11])
    assert f.params.recenter_far_value == pytest.approx(1e-4) and f.params.recenter_near_value == pytest.approx(5e-5)
    r = run(f, 0.0, 2)
    assert r.counts == [7, 4] and r.converged and r.escalations == 0
    assert f.rerecords == [3] and r.rerecorded == 1  # pass 3 starts from 8.6e-7
    assert Params(precision="mixed").recenter_far is None
    assert Params(system="plane-couette", precision="mixed", recenter_far=None).recenter_far_value == math.inf
    with pytest.raises(ValueError):
        Params(system="plane-couette", recenter_far=10.0, recenter_near=20.0)


class _EnergyFlow(_ScriptedFlow):
    """Scripted errors; the 'state' is a step counter and E' = 1 / (1 + state)."""

    def predict_and_correct(self, s):
        return s + 1, None, self.errs.pop(0)

    def get_stats(self, s):
        return {"E'": 1.0 / (1.0 + s)}


def test_stats_recorder_follows_the_reference_driver(tmp_path):
    from dnsjax_amd.driver import StatsRecorder

    f = _EnergyFlow([1e-9] * 7, tol=1e-6, max_it=3)
    f.params = Params(corrector_tolerance=1e-6, dt=0.1)
    rec = StatsRecorder(f, 0, it_stats=3)
    r = run(f, 0, 6, callback=rec)
    rec.finish(r.state, len(r.counts))
    # t0, the top of iteration 3, the final state (iteration 6 is never a loop top)
    t3 = 0.1 + 0.1 + 0.1
    t6 = t3 + 0.1 + 0.1 + 0.1
    assert rec.ts == [0.0, t3, t6] and rec.Eps == [1.0, 1 / 4, 1 / 7]
    path = str(tmp_path / "stats.npz")
    rec.save(path)
    with np.load(path) as z:
        assert sorted(z.files) == ["Eps", "ts"]
        np.testing.assert_array_equal(z["ts"], [0.0, t3, t6])
    # a one-step run writes nothing
    g = _EnergyFlow([1e-9], tol=1e-6, max_it=3)
    g.params = Params(corrector_tolerance=1e-6, dt=0.1)
    rec1 = StatsRecorder(g, 0, it_stats=1)
    r1 = run(g, 0, 1, callback=rec1)
    rec1.finish(r1.state, len(r1.counts))
    with pytest.raises(ValueError):
        rec1.save(str(tmp_path / "none.npz"))



def test_walled_systems_fix_ly_to_two():
    """parameters.py:190-191: the channel is y in [-1, 1]; any other ly would
    only rescale the y-quadrature norm, so it is rejected."""
    assert Params(system="plane-couette").ly == 2.0
    with pytest.raises(ValueError):
        Params(system="plane-couette", ly=4.0)


def test_ensemble_rejects_members_sharing_a_context():
    """Members run concurrently on their contexts' work buffers and control
    blocks, so one flow may not appear twice (checked before any GPU work)."""
    from dnsjax_amd.graph import Ensemble

    class _Flow:
        _native = object()

    f = _Flow()
    with pytest.raises(ValueError, match="own flow"):
        Ensemble([f, f], [None, None])
