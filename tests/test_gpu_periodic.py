"""GPU parity of the periodic path against the FP64 oracle (through the C ABI).

Bar: relative L2 <= 1e-12 for FP64 fields, corrector errors to 1e-9
relative, identical corrector iteration counts over a run; the mixed form
to the tolerances stated in each test.
"""

import functools
import math
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dnsjax_amd.params import Params  # noqa: E402
from oracle import periodic as P  # noqa: E402
from oracle import stepper as S  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

CASES = {
    "periodic_16_kolmogorov": dict(nx=16, ny=16, nz=16, L=2 * math.pi, system="kolmogorov", re=40.0, dt=0.01, tilt=0.0),
    "periodic_16x8x32_waleffe_ly4": dict(nx=16, ny=8, nz=32, L=4.0, system="waleffe", re=60.0, dt=0.02, tilt=0.0),
    "periodic_8_kolmogorov_tilt30": dict(nx=8, ny=8, nz=8, L=4.0, system="kolmogorov", re=30.0, dt=0.02, tilt=30.0),
}


def rel(a, b):
    a = a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def make(case, **kw):
    from dnsjax_amd.flows import TriplyPeriodicFlow

    c = CASES[case]
    p = Params(system=c["system"], nx=c["nx"], ny=c["ny"], nz=c["nz"], lx=c["L"], ly=c["L"], lz=c["L"],
               re=c["re"], dt=c["dt"], tilt_degree=c["tilt"], **kw)
    return TriplyPeriodicFlow(p)


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.mark.parametrize("case", list(CASES))
def test_golden_rhs_step_and_projection(case):
    g = np.load(os.path.join(GOLDEN, case + ".npz"))
    flow = make(case)
    u = flow.to_device(g["state"])
    rhs = flow.get_rhs(u)
    assert rel(rhs, g["rhs0"]) < 1e-12
    pred, rhs1, err1 = flow.predict_and_correct(u)
    assert rel(pred, g["pred1"]) < 1e-12
    assert rel(rhs1, g["rhs1"]) < 1e-12
    assert abs(float(err1) - float(g["err1"])) <= 1e-9 * float(g["err1"])
    assert rel(u, g["state"]) == 0.0  # predict_and_correct leaves state untouched
    pred2, rhs2, err2 = flow.iterate_correction(u, pred, rhs1)
    assert rel(pred2, g["pred2"]) < 1e-12
    assert rel(rhs2, g["rhs2"]) < 1e-12
    assert abs(float(err2) - float(g["err2"])) <= 1e-9 * float(g["err2"])
    cv, _ = flow.correct_velocity(pred2.clone())
    assert rel(cv, g["corrected"]) < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", list(CASES))
def test_carry_form_matches_golden_steps(case, precision):
    """The carry form (the device steps' and host loops' passes: B =
    ildt_2 (ldt_1 u^n + (1 - c) N^n) kept instead of N, each pass
    pred <- B + c ildt_2 N(pred)) against the golden predict-and-correct
    and corrector iteration: FP64 to 1e-12 with errors to 1e-9; mixed to
    the mixed form's bound against the incremental form's own mixed result;
    B itself against its definition from the golden state and N^n."""
    g = np.load(os.path.join(GOLDEN, case + ".npz"))
    flow = make(case, precision=precision)
    nat = flow._native
    u = flow.to_device(g["state"])
    pred, carry, err = nat.empty(), nat.empty(), nat.scalar()
    nat.predict_and_correct_carry(u, pred, carry, err)
    e1 = float(err)
    p1 = pred.clone()
    nat.iterate_correction_carry(pred, carry, err)
    e2 = float(err)
    if precision == "fp64":
        assert rel(p1, g["pred1"]) < 1e-12 and rel(pred, g["pred2"]) < 1e-12
        assert abs(e1 - float(g["err1"])) <= 1e-9 * float(g["err1"])
        assert abs(e2 - float(g["err2"])) <= 1e-9 * float(g["err2"])
        c = CASES[case]
        of = P.Flow(P.PeriodicParams(nx=c["nx"], ny=c["ny"], nz=c["nz"], lx=c["L"], ly=c["L"], lz=c["L"],
                                     re=c["re"], dt=c["dt"], system=c["system"], tilt_degree=c["tilt"]))
        want_b = of.ildt_2 * (of.ldt_1 * g["state"] + (1 - of.p.implicitness) * g["rhs0"])
        want_b[:, 0, 0, 0] = 0.0
        assert rel(carry, want_b) < 1e-12
    else:
        q1, _, f1 = flow.predict_and_correct(u)
        q2, _, f2 = flow.iterate_correction(u, q1.clone(), flow.predict_and_correct(u)[1])
        assert rel(p1, q1.cpu().numpy()) < 1e-9 and rel(pred, q2.cpu().numpy()) < 1e-9
        assert abs(e1 - float(f1)) <= 1e-6 * float(f1) and abs(e2 - float(f2)) <= 1e-3 * float(f2)


@pytest.mark.parametrize("case", list(CASES))
def test_generic_composition_matches_fused(case):
    from dnsjax_amd.timestep import make_stepper

    g = np.load(os.path.join(GOLDEN, case + ".npz"))
    flow = make(case)
    u = flow.to_device(g["state"])
    pc, ic = make_stepper(lambda s: flow.get_rhs(s), lambda s, r: flow.predict(s, r),
                          lambda a, b, c, d: flow.correct(a, b, c, d), lambda c: flow.norm(c))
    assert pc.fused is False and flow.predict_and_correct.fused is True
    pred, rhs1, err1 = pc(u)
    assert rel(pred, g["pred1"]) < 1e-12 and rel(rhs1, g["rhs1"]) < 1e-12
    assert abs(float(err1) - float(g["err1"])) <= 1e-9 * float(g["err1"])
    pred2, rhs2, err2 = ic(u, pred, rhs1)
    assert rel(pred2, g["pred2"]) < 1e-12
    assert abs(float(err2) - float(g["err2"])) <= 1e-9 * float(g["err2"])


def test_norm_and_measure_corrections():
    g = np.load(os.path.join(GOLDEN, "periodic_16_kolmogorov.npz"))
    flow = make("periodic_16_kolmogorov", measure_corrections=True)
    o = P.Flow(P.PeriodicParams(nx=16, ny=16, nz=16, lx=2 * math.pi, ly=2 * math.pi, lz=2 * math.pi, re=40.0, dt=0.01))
    a = g["rhs1"]
    n = flow.norm(flow.to_device(a))
    assert abs(float(n) - P.get_norm(a, o.fourier.k_metric)) < 1e-13 * P.get_norm(a, o.fourier.k_metric)
    # a gradient part for the projection to remove: u + grad(phi)
    rng = np.random.default_rng(9)
    phi = rng.standard_normal(g["state"].shape[1:]) + 1j * rng.standard_normal(g["state"].shape[1:])
    f = o.fourier
    noisy = P.correct_velocity(g["state"], o) + 0.01 * P.gradient(phi, f.kx, f.ky, f.kz)
    noisy = P.phys_to_spec(P.spec_to_phys(noisy, o.p), o.p)  # Hermitian at kx = 0
    st, meas = flow.correct_velocity(flow.to_device(noisy))
    ref = P.correct_velocity(noisy, o)
    assert rel(st, ref) < 1e-12
    corr = ref - noisy
    corr[:, 0, 0, 0] = 0
    assert abs(float(meas["div"]) - P.get_norm(corr, o.fourier.k_metric)) < 1e-10 * P.get_norm(corr, o.fourier.k_metric)


_C0 = dict(n=32, L=2 * math.pi, re=40.0, dt=0.01, tol=1e-10, steps=100, rms=1e-2, seed=0)


@functools.lru_cache(maxsize=1)
def _c0_oracle():
    """Config 0 (32^3, Re 40, dt 0.01, tol 1e-10, seed 0, RMS 1e-2) on the
    oracle: its 100 states, counts and errors."""
    from dnsjax_amd.fields import periodic_field

    n, L = _C0["n"], _C0["L"]
    u0 = periodic_field(n, n, n, L, L, L, _C0["rms"], seed=_C0["seed"])
    op = P.PeriodicParams(nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, re=_C0["re"], dt=_C0["dt"])
    of = P.Flow(op)
    pc, ic = S.make_stepper(P.get_rhs, P.predict, P.correct, P.norm)
    ref_states = []
    _, ref_counts, ref_errs = S.run(u0, _C0["steps"], _C0["tol"], 10, lambda s: pc(s, of),
                                    lambda a, b, c: ic(a, b, c, of), lambda s: P.correct_velocity(s, of),
                                    callback=lambda i, s: ref_states.append(s))
    return u0, of, ref_states, ref_counts, ref_errs


def _c0_flow(precision):
    from dnsjax_amd.flows import TriplyPeriodicFlow

    n, L = _C0["n"], _C0["L"]
    return TriplyPeriodicFlow(Params(system="kolmogorov", nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, re=_C0["re"],
                                     dt=_C0["dt"], corrector_tolerance=_C0["tol"], max_corrector_iterations=10,
                                     precision=precision))


def test_run_32cubed_counts_and_per_step_error():
    """Config 0: its full 100 steps against the oracle, identical counts,
    every step within 1e-12."""
    from dnsjax_amd.driver import run as gpu_run

    u0, _, ref_states, ref_counts, ref_errs = _c0_oracle()
    flow = _c0_flow("fp64")
    gpu_states = []
    r = gpu_run(flow, flow.to_device(u0), _C0["steps"], callback=lambda i, s: gpu_states.append(s.cpu().numpy()))
    assert len(r.counts) == 100 and r.converged
    assert r.counts == ref_counts
    worst = max(rel(a, b) for a, b in zip(gpu_states, ref_states))
    print(f"MEASURED c0_run worst per-step rel {worst:.3e}")
    assert worst < 1e-12, worst
    for a, b in zip(r.errors, ref_errs):
        assert abs(a - b) <= 1e-6 * b + 1e-300


def test_mixed_run_32cubed_against_the_oracle():
    """Config 0 in the mixed form (split-precision FP32 evaluations) over its
    100 steps: every step's corrector reaches the tolerance without
    escalating to FP64, and the end state's energy, dissipation and
    velocity stay within 1e-6 relative of the FP64 oracle's."""
    from dnsjax_amd.driver import run as gpu_run

    u0, of, ref_states, ref_counts, _ = _c0_oracle()
    flow = _c0_flow("mixed")
    r = gpu_run(flow, flow.to_device(u0), _C0["steps"])
    assert len(r.counts) == 100 and r.converged and r.escalations == 0
    assert max(r.errors) <= _C0["tol"]
    end = r.state.cpu().numpy()
    st, ref = flow.get_stats(r.state), P.get_stats(ref_states[-1], of)
    e_err = abs(float(st["E'"]) - ref["E'"]) / abs(ref["E'"])
    d_err = abs(float(st["D"]) - ref["D"]) / abs(ref["D"])
    v_err = rel(end, ref_states[-1])
    diff = sum(abs(a - b) for a, b in zip(r.counts, ref_counts))
    print(f"MEASURED c0_mixed energy {e_err:.2e} dissipation {d_err:.2e} velocity {v_err:.2e} "
          f"extra iterations {sum(r.counts)} (oracle {sum(ref_counts)}, per-step differences {diff})")
    assert e_err < 1e-6 and d_err < 1e-6 and v_err < 1e-6
    assert abs(sum(r.counts) - sum(ref_counts)) <= 5


def test_mixed_precision_close_to_fp64():
    g = np.load(os.path.join(GOLDEN, "periodic_16_kolmogorov.npz"))
    flow = make("periodic_16_kolmogorov", precision="mixed")
    u = flow.to_device(g["state"])
    rhs = flow.get_rhs(u)
    assert rel(rhs, g["rhs0"]) < 1e-5
    pred, rhs1, err1 = flow.predict_and_correct(u)
    assert rel(pred, g["pred1"]) < 1e-6


def test_rejects_bad_configs_and_aliasing():
    from dnsjax_amd._lib import DnsmiError
    from dnsjax_amd.flows import TriplyPeriodicFlow

    with pytest.raises(DnsmiError):
        TriplyPeriodicFlow(Params(nx=10, ny=16, nz=16))  # 15 is not a supported padded length
    flow = make("periodic_8_kolmogorov_tilt30")
    u = flow.init_state(seed=1, rms=0.1)
    with pytest.raises(DnsmiError):
        flow._native.predict_and_correct(u, pred=u)
    with pytest.raises(ValueError):
        flow.get_rhs(u.to(torch.complex64))


@pytest.mark.parametrize("variants", [0, 1, 2, 4, 7, 32, 34])
@pytest.mark.parametrize("case", ["periodic_16_kolmogorov", "periodic_16x8x32_waleffe_ly4"])
def test_every_kernel_variant_matches(case, variants):
    """Each exact kernel variant (as the autotuner may pick) on the golden step."""
    g = np.load(os.path.join(GOLDEN, case + ".npz"))
    for prec, tol in (("fp64", 1e-12), ("mixed", 1e-5)):
        flow = make(case, tile_cols=16, variants=variants, precision=prec)
        u = flow.to_device(g["state"])
        assert rel(flow.get_rhs(u), g["rhs0"]) < tol
        pred, rhs1, err1 = flow.predict_and_correct(u)
        assert rel(pred, g["pred1"]) < tol


def test_autotuner_reports_choices():
    flow = make("periodic_16_kolmogorov")
    t = flow._native.tuning
    assert "y_inverse_seq=" in t and "z_inverse_seq=" in t and "x_split=" in t and "z_inverse_stream=" in t


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("seq", [0, 2], ids=["paired", "sequential"])
def test_streaming_z_inverse_stores_are_bit_identical(seq, prec):
    """Variant bit 32 (non-temporal stores of the z-inverse output) changes
    where the lines are cached, not the values: at 64^3 the RHS and a
    predict-and-correct equal the plain stores' bit for bit."""
    from dnsjax_amd.flows import TriplyPeriodicFlow
    from dnsjax_amd.params import kolmogorov_2pi

    flows = [TriplyPeriodicFlow(kolmogorov_2pi(64, 60.0, 0.01, corrector_tolerance=1e-10, precision=prec,
                                               tile_cols=16, variants=seq | nt)) for nt in (0, 32)]
    assert [f._native.variants()[1] for f in flows] == [seq, seq | 32]
    u = flows[0].init_state(seed=3, rms=0.1)
    r = [f.get_rhs(u) for f in flows]
    assert torch.equal(r[0], r[1])
    p = [f.predict_and_correct(u) for f in flows]
    assert torch.equal(p[0][0], p[1][0]) and torch.equal(p[0][1], p[1][1])


@pytest.mark.parametrize("case", list(CASES))
def test_stats_match_oracle(case):
    g = np.load(os.path.join(GOLDEN, case + ".npz"))
    c = CASES[case]
    flow = make(case)
    st = flow.get_stats(flow.to_device(g["pred2"]))
    op = P.PeriodicParams(nx=c["nx"], ny=c["ny"], nz=c["nz"], lx=c["L"], ly=c["L"], lz=c["L"], re=c["re"],
                          system=c["system"], dt=c["dt"], tilt_degree=c["tilt"])
    ref = P.get_stats(g["pred2"], P.Flow(op))
    for k in ("E", "I", "D", "E'"):
        assert abs(float(st[k]) - ref[k]) <= 1e-12 * max(abs(ref[k]), 1e-30), (k, float(st[k]), ref[k])


def test_stats_at_full_size_agree_with_the_norm():
    """128^3: the statistics' partial sums fit their buffer, and E' equals
    half the squared corrector norm of the state (two device paths, each
    pinned against the oracle at small sizes)."""
    from dnsjax_amd.flows import TriplyPeriodicFlow
    from dnsjax_amd.params import kolmogorov_2pi

    flow = TriplyPeriodicFlow(kolmogorov_2pi(128, 100.0, 0.005, tile_cols=16))
    u = flow.init_state(seed=1, rms=0.1)
    st = flow.get_stats(u)
    n = float(flow.norm(u))
    assert abs(float(st["E'"]) - 0.5 * n * n) <= 1e-12 * 0.5 * n * n
    assert all(math.isfinite(float(v)) for v in st.values())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_unpadded_y_matches_oracle(precision):
    """oversample_y = False (y transformed on ny points, no 3/2 padding):
    RHS and one predictor-corrector pass against the oracle."""
    from dnsjax_amd.fields import periodic_field
    from dnsjax_amd.flows import TriplyPeriodicFlow

    n, L = 16, 2 * math.pi
    flow = TriplyPeriodicFlow(Params(system="kolmogorov", nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, re=40.0, dt=0.01,
                                     oversample_y=False, precision=precision))
    op = P.PeriodicParams(nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, re=40.0, dt=0.01, system="kolmogorov",
                          oversample_y=False)
    of = P.Flow(op)
    u = periodic_field(n, n, n, L, L, L, 0.1, 11)
    tol = 1e-12 if precision == "fp64" else 1e-5
    assert rel(flow.get_rhs(flow.to_device(u)), P.get_rhs(u, of)) < tol
    pc, _ = S.make_stepper(P.get_rhs, P.predict, P.correct, P.norm)
    want, _, _ = pc(u, of)
    got, _, _ = flow.predict_and_correct(flow.to_device(u))
    assert rel(got, want) < tol


def test_decaying_box_against_oracle_and_beltrami_decay():
    """The decaying box (no base flow, no forcing) on the GPU: a seeded
    field's RHS, step and corrector iteration against the oracle (1e-12);
    and the closed form: the Beltrami field (sin kz, cos kz, 0) has a zero
    nonlinear term, so a step multiplies it by exactly
    g = (1/dt - (1-c) k^2/Re) / (1/dt + c k^2/Re) with no corrector pass."""
    from dnsjax_amd.fields import periodic_field
    from dnsjax_amd.flows import TriplyPeriodicFlow
    from dnsjax_amd.driver import run

    n, L = 16, 2 * math.pi
    kw = dict(system="decaying-box", nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, re=10.0, dt=0.05)
    flow = TriplyPeriodicFlow(Params(**kw))
    of = P.Flow(P.PeriodicParams(implicitness=0.5, **kw))
    u0 = periodic_field(n, n, n, L, L, L, 0.2, 13)
    u = flow.to_device(u0)
    assert rel(flow.get_rhs(u), P.get_rhs(u0, of)) < 1e-12
    pc, ic = S.make_stepper(P.get_rhs, P.predict, P.correct, P.norm)
    want, want_rhs, want_err = pc(u0, of)
    got, got_rhs, got_err = flow.predict_and_correct(u)
    assert rel(got, want) < 1e-12
    assert abs(float(got_err) - want_err) <= 1e-9 * want_err
    want2, _, _ = ic(u0, want, want_rhs, of)
    got2, _, _ = flow.iterate_correction(u, got, got_rhs)
    assert rel(got2, want2) < 1e-12

    p = of.p
    z = np.arange(p.nz_padded) * p.lz / p.nz_padded
    Z = np.broadcast_to(z[None, :, None], (p.ny_padded, p.nz_padded, p.nx_padded))
    k = 2 * 2 * np.pi / p.lz
    b = P.phys_to_spec(np.array([np.sin(k * Z), np.cos(k * Z), 0 * Z]), p)
    g = (1 / p.dt - 0.5 * k**2 / p.re) / (1 / p.dt + 0.5 * k**2 / p.re)
    r = run(flow, flow.to_device(b), 3)
    assert r.counts == [0, 0, 0]
    np.testing.assert_allclose(r.state.cpu().numpy(), g**3 * b, rtol=0, atol=1e-14)


@pytest.mark.parametrize("c", [0.3, 1.0])
def test_implicitness_against_oracle(c):
    """The Crank-Nicolson weight c enters ldt_1, ildt_2 and the corrector's
    c (N_next - N_prev) term: a step and a corrector iteration at c != 1/2
    against the oracle (Kolmogorov 16^3, 1e-12)."""
    from dnsjax_amd.fields import periodic_field
    from dnsjax_amd.flows import TriplyPeriodicFlow

    n, L = 16, 2 * math.pi
    kw = dict(system="kolmogorov", nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, re=40.0, dt=0.01, implicitness=c)
    flow = TriplyPeriodicFlow(Params(**kw))
    of = P.Flow(P.PeriodicParams(**kw))
    u0 = periodic_field(n, n, n, L, L, L, 0.1, 17)
    u = flow.to_device(u0)
    pc, ic = S.make_stepper(P.get_rhs, P.predict, P.correct, P.norm)
    want, want_rhs, want_err = pc(u0, of)
    got, got_rhs, got_err = flow.predict_and_correct(u)
    assert rel(got, want) < 1e-12
    assert abs(float(got_err) - want_err) <= 1e-9 * want_err
    want2, _, _ = ic(u0, want, want_rhs, of)
    got2, _, _ = flow.iterate_correction(u, got, got_rhs)
    assert rel(got2, want2) < 1e-12
