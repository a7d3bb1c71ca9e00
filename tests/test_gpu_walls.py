"""GPU parity of the plane Couette path against the FP64 oracle (C ABI).

Bar: set-up tables equal to the oracle's (pivots exactly, factors to
1e-13, the homogeneous solutions to 1e-12: they come from banded solves
here and dense solves in the oracle, and agree to the conditioning of L_k),
RHS and IMM predictor/corrector states to 1e-12 relative L2 per step,
corrector errors to 1e-9 relative, identical corrector iteration counts
over a run; at config 2's shape with odd ny (64 x 129 x 64, where the
reference's quadrature is defined) against the oracle directly.
Odd Ny is where the reference is defined; even Ny uses the documented
quadrature extension in both the oracle and the GPU path.
"""

import math
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dnsjax_amd.params import Params  # noqa: E402
from oracle import stepper as S  # noqa: E402
from oracle import walls as W  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

CASES = {
    "walls_8x17x8_p4": dict(nx=8, ny=17, nz=8, fd_order=4, re=400.0, dt=0.01),
    "walls_16x33x8_p6": dict(nx=16, ny=33, nz=8, fd_order=6, re=400.0, dt=0.01),
}


def rel(a, b):
    a = a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def make(case=None, **kw):
    from dnsjax_amd.flows import PlaneCouetteFlow

    c = dict(CASES[case]) if case else {}
    c.update(kw)
    p = Params(system="plane-couette", lx=2 * math.pi, lz=math.pi, **c)
    return PlaneCouetteFlow(p)


@pytest.mark.parametrize("case", list(CASES))
def test_setup_tables_match_oracle(case):
    c = CASES[case]
    flow = make(case)
    of = W.WallFlow(W.WallParams(nx=c["nx"], ny=c["ny"], nz=c["nz"], fd_order=c["fd_order"], re=c["re"], dt=c["dt"]))
    nat = flow._native
    nkz, nkx, ny, p = c["nz"] - 1, c["nx"] // 2, c["ny"], c["fd_order"]
    ys = nat.walls_table(5, (ny,))
    assert np.max(np.abs(ys - of.ys)) < 1e-15
    for which, key, shp in ((6, "Lk_piv", (nkz, nkx, ny)), (7, "Hk_piv", (nkz, nkx, ny))):
        assert np.array_equal(nat.walls_table(which, shp).astype(np.int32), of.imm[key])
    for which, key in ((0, "Lk_ab"), (1, "Hk_ab")):
        got = nat.walls_table(which, (nkz, nkx, 3 * p + 1, ny))
        ref = of.imm[key]
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
    for which, key in ((2, "v1"), (3, "q1")):
        got = nat.walls_table(which, (nkz, nkx, ny))
        assert rel(got, of.imm[key]) < 1e-12
    minv = nat.walls_table(4, (nkz, nkx, 2, 2))
    assert rel(minv, of.imm["M_inv"]) < 1e-12


SETUP_SHAPES = [
    dict(nx=8, ny=17, nz=8, fd_order=4, re=400.0, dt=0.01),
    dict(nx=16, ny=33, nz=8, fd_order=6, re=400.0, dt=0.01),
    dict(nx=16, ny=24, nz=16, fd_order=2, re=100.0, dt=0.02),
    dict(nx=16, ny=40, nz=8, fd_order=8, re=1000.0, dt=0.005, implicitness=0.3),
    dict(nx=64, ny=129, nz=64, fd_order=4, re=400.0, dt=0.01),
]


@pytest.mark.parametrize("shape", SETUP_SHAPES, ids=lambda s: f"{s['nx']}x{s['ny']}x{s['nz']}_p{s['fd_order']}")
def test_gpu_setup_equals_host_setup(shape):
    """The influence-matrix tables built on the GPU (band LU, homogeneous
    solutions, M^-1, chunk maps) equal the host set-up's bit for bit."""
    g, h = make(None, **shape), make(None, setup_on_host=True, **shape)
    assert "imm_setup=gpu" in g._native.tuning and "imm_setup=host" in h._native.tuning
    G, H = g._native, h._native
    ny, p = shape["ny"], shape["fd_order"]
    M, rows = (shape["nz"] - 1) * (shape["nx"] // 2), 3 * p + 1
    for which, shp in ((0, (M, rows, ny)), (1, (M, rows, ny)), (2, (M, ny)), (3, (M, ny)), (4, (M, 4)),
                       (6, (M, ny)), (7, (M, ny))):
        assert np.array_equal(G.walls_table(which, shp), H.walls_table(which, shp)), which
    sizes = G.walls_table(13, (5,))
    assert np.array_equal(sizes, H.walls_table(13, (5,)))
    U = int(sizes[4])
    assert np.array_equal(G.walls_table(12, (U * ny * 4,)), H.walls_table(12, (U * ny * 4,)))
    for i, which in enumerate((8, 9, 10, 11)):
        n = int(sizes[i])
        assert n > 0
        assert np.array_equal(G.walls_table(which, (n,)), H.walls_table(which, (n,))), which


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("variants", [None, 0, 16, 8, 64], ids=["autotuned", "chunked", "fused", "sequential",
                                                                "curl64"])
def test_golden_rhs_and_imm_step(case, variants):
    """The band-solve forms (chunk maps in five launches or fused into two;
    sequential sweeps), the FD curl in 64-row workgroups (variant bit 64; the
    automatic choice at large shapes) and the autotuner's choice against the
    golden steps."""
    g = np.load(os.path.join(GOLDEN, case + ".npz"))
    flow = make(case) if variants is None else make(case, tile_cols=16, variants=variants)
    u = flow.to_device(g["state"])
    rhs = flow.get_rhs(u)
    assert rel(rhs, g["rhs0"]) < 1e-12
    pred, rhs1, err1 = flow.predict_and_correct(u)
    assert rel(u, g["state"]) == 0.0
    assert rel(rhs1, g["rhs1"]) < 1e-12
    assert rel(pred, g["pred1"]) < 1e-12
    assert abs(float(err1) - float(g["err1"])) <= 1e-9 * float(g["err1"])
    pred2, rhs2, err2 = flow.iterate_correction(u, pred, rhs1)
    assert rel(pred2, g["pred2"]) < 1e-12
    assert rel(rhs2, g["rhs2"]) < 1e-12
    assert abs(float(err2) - float(g["err2"])) <= 1e-8 * float(g["err2"]) + 1e-20


@pytest.mark.parametrize("case", list(CASES))
def test_generic_callables_match_fused(case):
    from dnsjax_amd.timestep import make_stepper

    g = np.load(os.path.join(GOLDEN, case + ".npz"))
    flow = make(case)
    u = flow.to_device(g["state"])
    pc, ic = make_stepper(lambda s: flow.get_rhs(s), lambda s, r: flow.predict(s, r),
                          lambda a, b, c, d: flow.correct(a, b, c, d), lambda c: flow.norm(c))
    pred, rhs1, err1 = pc(u)
    assert rel(pred, g["pred1"]) < 1e-12
    assert abs(float(err1) - float(g["err1"])) <= 1e-9 * float(g["err1"])
    pred2, rhs2, err2 = ic(u, pred, rhs1)
    assert rel(pred2, g["pred2"]) < 1e-12
    # in-place rhs update (rhs_next aliasing rhs_prev) gives the same result
    p3, r3, e3 = flow._native.iterate_correction(u, pred.clone(), rhs1.clone(), rhs_next=None)
    rp = rhs1.clone()
    p4, r4, e4 = flow._native.iterate_correction(u, pred.clone(), rp, rhs_next=rp)
    assert torch.equal(p3, p4) and torch.equal(r3, r4) and float(e3) == float(e4)


@pytest.mark.parametrize("ny", [33, 32])
@pytest.mark.parametrize("variants", [0, 16, 8], ids=["chunked", "fused", "sequential"])
def test_run_counts_and_per_step_error(ny, variants):
    """Plane Couette 16 x ny x 16, p = 4, Re 400, dt 0.005, tol 1e-10: 10 steps.

    The wall corrector contracts by about 2 per pass, so dt 0.005 with up
    to 20 passes keeps every step converged (8-12 passes each)."""
    from dnsjax_amd.driver import run as gpu_run
    from dnsjax_amd.fields import wall_field

    nx = nz = 16
    flow = make(nx=nx, ny=ny, nz=nz, fd_order=4, re=400.0, dt=0.005, corrector_tolerance=1e-10,
                max_corrector_iterations=20, tile_cols=16, variants=variants)
    op = W.WallParams(nx=nx, ny=ny, nz=nz, fd_order=4, re=400.0, dt=0.005)
    of = W.WallFlow(op)
    u0 = wall_field(nx, ny, nz, op.lx, op.lz, 0.05, seed=2)
    pc, ic = S.make_stepper(W.get_rhs, W.predict, W.correct, W.norm)
    ref = []
    _, ref_counts, _ = S.run(u0, 10, 1e-10, 20, lambda s: pc(s, of), lambda a, b, c: ic(a, b, c, of),
                             lambda s: W.correct_velocity(s, of), callback=lambda i, s: ref.append(s))
    got = []
    r = gpu_run(flow, flow.to_device(u0), 10, callback=lambda i, s: got.append(s.cpu().numpy()))
    assert len(ref_counts) == 10 and r.converged
    assert r.counts == ref_counts
    worst = max(rel(a, b) for a, b in zip(got, ref))
    print(f"MEASURED walls_run ny={ny} variants={variants} worst per-step rel {worst:.3e}")
    assert worst < 1e-12, worst


def test_full_size_imm_invariants():
    """Config 2's shape (64 x 128 x 64, p = 4): no-slip walls and zero
    discrete wall divergence after one predictor-corrector step."""
    from dnsjax_amd.fields import wall_field

    nx, ny, nz = 64, 128, 64
    flow = make(nx=nx, ny=ny, nz=nz, fd_order=4, re=400.0, dt=0.01)
    u0 = flow.to_device(wall_field(nx, ny, nz, 2 * math.pi, math.pi, 0.05, seed=2))
    pred, rhs, err = flow.predict_and_correct(u0)
    v = pred.cpu().numpy()
    scale0 = np.max(np.abs(v))
    assert np.max(np.abs(v[..., 0])) < 1e-13 * scale0 and np.max(np.abs(v[..., -1])) < 1e-13 * scale0
    ys = flow._native.walls_table(5, (ny,))
    D1, _ = W.build_diff_matrices(ys, 4)
    kx = (np.arange(nx // 2) * 2 * np.pi / (2 * np.pi)).reshape(1, -1)
    hz = np.concatenate([np.arange(nz // 2), np.arange(-nz // 2 + 1, 0)])
    kz = (hz * 2 * np.pi / np.pi).reshape(-1, 1)
    div = 1j * kx[..., None] * v[0] + np.einsum("ij,zxj->zxi", D1, v[1]) + 1j * kz[..., None] * v[2]
    scale = np.max(np.abs(v))
    assert np.max(np.abs(div[..., -1])) < 1e-9 * scale
    bot = div[..., 0].copy()
    bot[0, 0] = 0
    assert np.max(np.abs(bot)) < 1e-9 * scale
    assert math.isfinite(float(err)) and float(err) > 0


@pytest.mark.parametrize("case", list(CASES))
def test_walls_stats_match_oracle(case):
    g = np.load(os.path.join(GOLDEN, case + ".npz"))
    c = CASES[case]
    flow = make(case)
    st = flow.get_stats(flow.to_device(g["pred2"]))
    of = W.WallFlow(W.WallParams(nx=c["nx"], ny=c["ny"], nz=c["nz"], fd_order=c["fd_order"], re=c["re"], dt=c["dt"]))
    ref = W.get_stats(g["pred2"], of)["E'"]
    assert abs(float(st["E'"]) - ref) <= 1e-12 * ref


@pytest.mark.parametrize("p", [2, 8])
@pytest.mark.parametrize("variants", [0, 16, 8, 64], ids=["chunked", "fused", "sequential", "curl64"])
def test_fd_orders_2_and_8_against_oracle(p, variants):
    """The remaining stencil orders (the goldens hold 4 and 6): 8 x 33 x 8,
    Re 400, dt 0.01.  The RHS to 1e-12, one predictor-corrector step and
    one corrector iteration to 1e-12 relative, their errors to 1e-9, for
    both band-solve forms (the chunked one runs with two map slots and one
    wave per SIMD at order 8)."""
    from dnsjax_amd.fields import wall_field

    nx, ny, nz = 8, 33, 8
    flow = make(nx=nx, ny=ny, nz=nz, fd_order=p, re=400.0, dt=0.01, tile_cols=16, variants=variants)
    of = W.WallFlow(W.WallParams(nx=nx, ny=ny, nz=nz, fd_order=p, re=400.0, dt=0.01))
    u0 = wall_field(nx, ny, nz, 2 * math.pi, math.pi, 0.05, seed=7)
    u = flow.to_device(u0)
    assert rel(flow.get_rhs(u), W.get_rhs(u0, of)) < 1e-12
    pc, ic = S.make_stepper(W.get_rhs, W.predict, W.correct, W.norm)
    want, want_rhs, want_err = pc(u0, of)
    got, got_rhs, got_err = flow.predict_and_correct(u)
    assert rel(got, want) < 1e-12
    assert abs(float(got_err) - want_err) <= 1e-9 * want_err
    want2, _, want_err2 = ic(u0, want, want_rhs, of)
    got2, _, got_err2 = flow.iterate_correction(u, got, got_rhs)
    assert rel(got2, want2) < 1e-12
    assert abs(float(got_err2) - want_err2) <= 1e-9 * want_err2


@pytest.mark.parametrize("c", [0.3, 1.0])
def test_implicitness_against_oracle(c):
    """The Crank-Nicolson weight c enters H_k, the homogeneous solutions and
    the IMM right-hand sides: a step and a corrector iteration at c != 1/2
    against the oracle (8 x 17 x 8, p = 4, 1e-12)."""
    from dnsjax_amd.fields import wall_field

    nx, ny, nz = 8, 17, 8
    flow = make(nx=nx, ny=ny, nz=nz, fd_order=4, re=400.0, dt=0.01, implicitness=c)
    of = W.WallFlow(W.WallParams(nx=nx, ny=ny, nz=nz, fd_order=4, re=400.0, dt=0.01, implicitness=c))
    u0 = wall_field(nx, ny, nz, 2 * math.pi, math.pi, 0.05, seed=11)
    u = flow.to_device(u0)
    pc, ic = S.make_stepper(W.get_rhs, W.predict, W.correct, W.norm)
    want, want_rhs, want_err = pc(u0, of)
    got, got_rhs, got_err = flow.predict_and_correct(u)
    assert rel(got, want) < 1e-12
    assert abs(float(got_err) - want_err) <= 1e-9 * want_err
    want2, _, _ = ic(u0, want, want_rhs, of)
    got2, _, _ = flow.iterate_correction(u, got, got_rhs)
    assert rel(got2, want2) < 1e-12


@pytest.mark.parametrize("variants", [0, 16, 8], ids=["chunked", "fused", "sequential"])
def test_incommensurate_box_against_oracle(variants):
    """Lx = 3, Lz = 1.7: only the +-kz pairs share k^2 (the set-up computes
    the per-mode tables and chunk maps once per distinct k^2, bit for bit),
    against the oracle's per-mode tables: one step and one corrector
    iteration at 16 x 33 x 16, p = 4 (1e-12)."""
    from dnsjax_amd.fields import wall_field
    from dnsjax_amd.flows import PlaneCouetteFlow

    nx, ny, nz, lx, lz = 16, 33, 16, 3.0, 1.7
    flow = PlaneCouetteFlow(Params(system="plane-couette", nx=nx, ny=ny, nz=nz, lx=lx, lz=lz, fd_order=4,
                                   re=400.0, dt=0.01, tile_cols=16, variants=variants))
    of = W.WallFlow(W.WallParams(nx=nx, ny=ny, nz=nz, lx=lx, lz=lz, fd_order=4, re=400.0, dt=0.01))
    u0 = wall_field(nx, ny, nz, lx, lz, 0.05, seed=5)
    u = flow.to_device(u0)
    pc, ic = S.make_stepper(W.get_rhs, W.predict, W.correct, W.norm)
    want, want_rhs, want_err = pc(u0, of)
    got, got_rhs, got_err = flow.predict_and_correct(u)
    assert rel(got, want) < 1e-12
    assert abs(float(got_err) - want_err) <= 1e-9 * want_err
    want2, _, _ = ic(u0, want, want_rhs, of)
    got2, _, _ = flow.iterate_correction(u, got, got_rhs)
    assert rel(got2, want2) < 1e-12


@pytest.mark.parametrize("variants", [0, 16, 8], ids=["chunked", "fused", "sequential"])
def test_config2_shape_odd_ny_against_oracle(variants):
    """Config 2's shape with the odd ny where the reference's quadrature is
    defined (64 x 129 x 64, p = 4, Re 400, dt 0.01, seed 2, RMS 0.05) against
    the oracle: the RHS, one predictor-corrector step and one corrector
    iteration to 1e-12 with their errors to 1e-9, then 5 steps at tolerance
    1e-7 with identical corrector counts (4 per step) and every step's state
    within 1e-12."""
    from dnsjax_amd.driver import run as gpu_run
    from dnsjax_amd.fields import wall_field

    nx, ny, nz = 64, 129, 64
    flow = make(nx=nx, ny=ny, nz=nz, fd_order=4, re=400.0, dt=0.01, corrector_tolerance=1e-7,
                max_corrector_iterations=20, tile_cols=16, variants=variants)
    op = W.WallParams(nx=nx, ny=ny, nz=nz, fd_order=4, re=400.0, dt=0.01)
    of = W.WallFlow(op)
    u0 = wall_field(nx, ny, nz, op.lx, op.lz, 0.05, seed=2)
    u = flow.to_device(u0)
    e_rhs = rel(flow.get_rhs(u), W.get_rhs(u0, of))
    pc, ic = S.make_stepper(W.get_rhs, W.predict, W.correct, W.norm)
    want, want_rhs, want_err = pc(u0, of)
    got, got_rhs, got_err = flow.predict_and_correct(u)
    e1 = rel(got, want)
    want2, _, want_err2 = ic(u0, want, want_rhs, of)
    got2, _, got_err2 = flow.iterate_correction(u, got, got_rhs)
    e2 = rel(got2, want2)
    print(f"MEASURED c2_odd variants={variants} rhs {e_rhs:.3e} pc {e1:.3e} ic {e2:.3e} "
          f"err {abs(float(got_err) / want_err - 1):.2e} {abs(float(got_err2) / want_err2 - 1):.2e}")
    assert e_rhs < 1e-12 and e1 < 1e-12 and e2 < 1e-12
    assert abs(float(got_err) - want_err) <= 1e-9 * want_err
    assert abs(float(got_err2) - want_err2) <= 1e-9 * want_err2
    ref = []
    _, ref_counts, _ = S.run(u0, 5, 1e-7, 20, lambda s: pc(s, of), lambda a, b, c: ic(a, b, c, of),
                             lambda s: W.correct_velocity(s, of), callback=lambda i, s: ref.append(s))
    states = []
    r = gpu_run(flow, u, 5, callback=lambda i, s: states.append(s.cpu().numpy()))
    worst = max(rel(a, b) for a, b in zip(states, ref))
    print(f"MEASURED c2_odd variants={variants} counts {r.counts} worst per-step rel {worst:.3e}")
    assert r.counts == ref_counts == [4, 4, 4, 4, 4]
    assert worst < 1e-12, worst


@pytest.mark.parametrize("shape", [(16, 33, 16, 4), (64, 129, 64, 4), (8, 33, 8, 8), (16, 34, 8, 6)])
def test_fused_imm_equals_five_launch_form(shape):
    """The fused two-launch influence-matrix step computes the five-launch
    chunked form's values bit for bit (same per-row arithmetic and chunk
    chains): predictor, corrector and correction equal, the error norms
    (summed in another order) to 1e-14."""
    from dnsjax_amd.fields import wall_field

    nx, ny, nz, p = shape
    u0 = wall_field(nx, ny, nz, 2 * math.pi, math.pi, 0.05, seed=3)
    out = {}
    for v in (0, 16):
        flow = make(nx=nx, ny=ny, nz=nz, fd_order=p, re=400.0, dt=0.01, tile_cols=16, variants=v)
        u = flow.to_device(u0)
        pred, rhs1, err1 = flow.predict_and_correct(u)
        pred2, rhs2, err2 = flow.iterate_correction(u, pred, rhs1)
        out[v] = (pred, pred2, rhs2, float(err1), float(err2))
    a, b = out[0], out[16]
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1]) and torch.equal(a[2], b[2])
    assert abs(a[3] - b[3]) <= 1e-14 * a[3] and abs(a[4] - b[4]) <= 1e-14 * a[4]
