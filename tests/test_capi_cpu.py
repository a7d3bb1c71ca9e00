"""CPU-only checks of the boundary: the C-ABI library loads and exports every
symbol include/dnsmi.h declares; host-side logic (parameters, stepper
composition, driver loop semantics) without touching a GPU."""

import ctypes
import os
import re

import numpy as np
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


def test_driver_noise_floor_rule_for_mixed():
    f = _ScriptedFlow([1e-5, 1e-7, 2e-9, 1.6e-10, 1.3e-10, 1e-5, 1e-11], tol=1e-10, max_it=10)
    f.params = Params(precision="mixed", corrector_tolerance=1e-10, max_corrector_iterations=10)
    r = run(f, 0.0, 2)
    assert r.counts == [4, 1] and r.floor_stops == 1 and r.converged
    # the plain form keeps the reference rule
    g = _ScriptedFlow([1e-5, 1e-7, 2e-9, 1.6e-10, 1.3e-10, 1.2e-10, 1.1e-10, 1.05e-10, 1.01e-10, 1.0e-10], tol=1e-10, max_it=10)
    g.params = Params(corrector_tolerance=1e-10, max_corrector_iterations=10)
    r = run(g, 0.0, 1)
    assert r.counts == [9] and r.floor_stops == 0


def test_noise_floor_margin_after_a_stall():
    """After a stall stop has shown the floor F, a pass within the margin
    (2 F for the mixed form) of it ends the corrector without the confirming
    pass; F then only decreases, until a new stall sets it again."""
    f = _ScriptedFlow([1e-5, 1e-7, 2e-9, 1.6e-10, 1.3e-10,  # stall: F = 1.3e-10
                       1e-5, 1e-7, 2e-9, 2.5e-10,  # <= 2 F: margin stop
                       1e-5, 1e-7, 3e-10, 1.2e-10,  # 3e-10 > 2 F goes on; then F = 1.2e-10
                       1e-5, 2.5e-10, 2.3e-10],  # > 2 F, then a stall: F = 2.3e-10
                      tol=1e-10, max_it=10)
    f.params = Params(precision="mixed", corrector_tolerance=1e-10, max_corrector_iterations=10)
    assert f.params.corrector_floor_margin == 2.0
    r = run(f, 0.0, 4)
    assert r.counts == [4, 3, 3, 2] and r.floor_stops == 4 and r.converged
    assert not f.errs
    # without the margin the second step needs the confirming pass
    g = _ScriptedFlow([1e-5, 1e-7, 2e-9, 1.6e-10, 1.3e-10, 1e-5, 1e-7, 2e-9, 2.5e-10, 1.3e-10],
                      tol=1e-10, max_it=10)
    g.params = Params(precision="mixed", corrector_tolerance=1e-10, max_corrector_iterations=10,
                      corrector_floor_margin=0.0)
    r = run(g, 0.0, 2)
    assert r.counts == [4, 4] and r.floor_stops == 2


def test_noise_floor_rule_ignores_slow_contraction():
    """A loose tolerance with a corrector contracting only ~2x per pass (the
    wall-bounded one) must iterate to the tolerance: a stall above the
    absolute ceiling (1e-8) is not the noise floor."""
    f = _ScriptedFlow([1e-3, 6e-4, 3.5e-4, 2e-4, 1.1e-4, 6e-5, 3.2e-5, 1.7e-5, 9e-6], tol=1e-5, max_it=10)
    f.params = Params(precision="mixed", corrector_tolerance=1e-5, max_corrector_iterations=10)
    assert f.params.floor_cap_value == 1e-8
    r = run(f, 0.0, 1)
    assert r.counts == [8] and r.floor_stops == 0 and r.converged


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
