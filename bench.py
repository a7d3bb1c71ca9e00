"""Time-steps/s of the periodic hot path at 128^3 Kolmogorov on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload (one "step" = one time step of the driver loop: predict_and_correct,
iterate_correction until the corrector error is below tolerance,
correct_velocity): Kolmogorov flow, nx = ny = nz = 128 (192^3 physical on
the 3/2 grid), 2 pi box, Re = 100, dt = 0.005, implicitness 0.5,
corrector tolerance 1e-10, seeded solenoidal field (seed 1, RMS 0.1).
The state is resident in HBM before the timed region.  `value` is the
plain FP64 form; beside it: the mixed form (FP32 transforms / nonlinear
product), the mixed form with the learned corrector guess (A/B over one
trajectory segment, `mixed_learned_vs_fp64`), and the 64-member ensemble.
Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-steps/s + HBM roofline % at 128³ Kolmogorov; mixed+learned vs FP64, 1 GPU"
PEAK_GBS = 8000.0
N = 128
RE, DT, RMS, SEED, TOL = 100.0, 0.005, 0.1, 1, 1e-10
WALLS_SEED = 2  # the plane Couette 64x128x64 workload's seeded perturbation


def progress(what: str) -> None:
    """One line on stderr per stage (stdout carries only the result line)."""
    print(f"[bench] {what}", file=sys.stderr, flush=True)


def kernel_bytes(n: int, b: int) -> dict:
    """Algorithmic HBM bytes per launch of each periodic kernel.

    S = one spectral FP64 vector field; intermediates hold b-byte complex.
    The y-forward kernel's epilogue traffic depends on its mode and is
    given per mode.
    """
    Nx, Ny, Nz, P = n // 2, n - 1, n - 1, 3 * n // 2
    S = 3 * Ny * Nz * Nx * 16
    return {
        "y_inverse": S + 5 * P * Nz * Nx * b,
        "z_inverse": 5 * P * Nz * Nx * b + 6 * P * P * Nx * b,
        "x_nonlinear": 9 * P * P * Nx * b,
        "z_forward": 3 * P * P * Nx * b + 3 * P * Nz * Nx * b,
        "y_forward:rhs": 3 * P * Nz * Nx * b + S,
        # the stepping loop's epilogues (carry form): the predictor reads u^n
        # and writes u^1 and B; a corrector pass reads B and the iterate and
        # writes the iterate
        "y_forward:predict": 3 * P * Nz * Nx * b + 3 * S,
        "y_forward:correct": 3 * P * Nz * Nx * b + 3 * S,
        # the incremental form of the public iterate_correction (also reads
        # N_prev and writes N)
        "y_forward:correct_incremental": 3 * P * Nz * Nx * b + 4 * S,
        # the mixed form's split evaluations: the recording evaluation writes
        # u_ref and the six physical fields (FP32), each increment reads
        # them; the recording corrector epilogue also writes B + c ildt_2
        # N(u_ref), which each increment's epilogue reads in place of B
        "y_inverse:split": S + 5 * P * Nz * Nx * b + S,
        "x_nonlinear:split": 9 * P * P * Nx * b + 6 * P * P * P * 4,
        "y_forward:correct_record": 3 * P * Nz * Nx * b + 4 * S,
        "y_forward:correct_split": 3 * P * Nz * Nx * b + 3 * S,
        "S": S,
    }


def rhs_model_bytes(n: int, b: int) -> int:
    """The fixed 5-pass pencil model of one RHS: each pass (y inverse, z
    inverse, x pass with the product, z forward, y forward) reads its input
    and writes its output once; zero pads are not read.  The state read and
    the RHS written are FP64 spectral fields in both forms (16 bytes per
    complex value); the intermediates take b bytes: 1228.0 MB (b = 16) and
    663.6 MB (b = 8) at 128^3."""
    Nx, Ny, Nz, P = n // 2, n - 1, n - 1, 3 * n // 2
    return Nx * (6 * Ny * Nz * 16 + b * (18 * P * Nz + 18 * P * P))


def run_steps(flow, state, k):
    from dnsjax_amd.driver import run

    r = run(flow, state, k)
    return r.state, r.counts


def gpu_form(precision: str, steps: int, warmup: int):
    import torch

    from dnsjax_amd.flows import TriplyPeriodicFlow
    from dnsjax_amd.params import kolmogorov_2pi

    p = kolmogorov_2pi(N, RE, DT, corrector_tolerance=TOL, max_corrector_iterations=10,
                       precision=precision)
    flow = TriplyPeriodicFlow(p)
    state = flow.init_state(seed=SEED, rms=RMS)
    state, _ = run_steps(flow, state, warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    state, counts = run_steps(flow, state, steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    # second, identical region with per-kernel HIP events on the launch stream
    flow._native.profile_begin()
    state, counts2 = run_steps(flow, state, steps)
    torch.cuda.synchronize()
    prof = flow._native.profile_end()
    # third region: the RHS alone (the y forward with its RHS-only epilogue),
    # the "FFT/NL pass" the pass model describes
    rhs = flow._native.get_rhs(state)
    torch.cuda.synchronize()
    flow._native.profile_begin()
    for _ in range(20):
        flow._native.get_rhs(state, out=rhs)
    torch.cuda.synchronize()
    prof_rhs = flow._native.profile_end()
    return {
        "steps_per_s": steps / elapsed,
        "ms_per_step": 1e3 * elapsed / steps,
        "corrector_iters_per_step": sum(counts) / len(counts),
        "counts": counts,
        "prof": prof,
        "prof_counts": counts2,
        "prof_rhs": prof_rhs,
    }


def measured_traffic(kernel: str, b: int, mode: str = "2"):
    """HBM bytes per dispatch from the committed rocprofv3 PMC summary
    (FETCH_SIZE / WRITE_SIZE passes scaled by a calibration copy), or None."""
    path = os.path.join(ROOT, "profiles", "periodic_128_pmc_traffic.json")
    if not os.path.exists(path):
        return None
    t = "double" if b == 16 else "float"
    args = {"y_inverse": f"{t}, 192, 16, true, 0", "x_nonlinear": f"{t}, 192, 16, 0",
            "y_forward": f"{t}, 192, 16, {mode}, true"}.get(kernel, f"{t}, 192, 16")
    ent = json.load(open(path))["kernels"].get(f"k_{kernel}<{args}>")
    return int(ent["traffic_bytes"]) if ent else None


def roofline(form: dict, b: int, split: bool = False) -> dict:
    kb = kernel_bytes(N, b)
    prof = form["prof"]
    # y_forward launches: one predict per step, the rest correct; in the
    # split mixed form every y inverse and x pass moves the reference
    # fields too (recorded once per step, read by every increment)
    steps = len(form["prof_counts"])
    per = {}
    for name, (ms, cnt) in prof.items():
        if name == "y_forward":
            # per step one predictor and one recording corrector epilogue
            # (predict_and_correct), then the increments
            n_pred = steps
            n_rec = steps if split else 0
            n_corr = cnt - n_pred - n_rec
            corr = kb["y_forward:correct_split" if split else "y_forward:correct"]
            byts = (n_pred * kb["y_forward:predict"] + n_rec * kb["y_forward:correct_record"] + n_corr * corr) / cnt
        elif split and name in ("y_inverse", "x_nonlinear"):
            byts = kb[name + ":split"]
        elif name in kb:
            byts = kb[name]
        else:
            continue
        avg_s = ms / cnt / 1e3
        per[name] = {"avg_us": avg_s * 1e6, "launches": cnt, "bytes": byts,
                     "gbs": byts / avg_s / 1e9}
    dom = max(per, key=lambda k: prof[k][0])
    d = per[dom]
    rhs_names = ("y_inverse", "z_inverse", "x_nonlinear", "z_forward", "y_forward")
    # one pass of the stepping loop: the five kernels with the y forward's
    # predictor / corrector epilogue, against their algorithmic bytes
    t_pass = sum(per[k]["avg_us"] for k in rhs_names if k in per) * 1e-6
    b_pass = sum(per[k]["bytes"] for k in rhs_names if k in per)
    # the RHS alone against the pass model
    pr = form["prof_rhs"]
    t_rhs = sum(pr[k][0] / pr[k][1] for k in rhs_names if k in pr) * 1e-3
    rhs_gbs = rhs_model_bytes(N, b) / t_rhs / 1e9
    return {
        "dominant": {"bound": "hbm", "kernel": dom, "achieved": round(d["gbs"], 1), "peak": PEAK_GBS,
                     "unit": "GB/s", "frac": round(d["gbs"] / PEAK_GBS, 4),
                     "traffic": measured_traffic(dom, b),
                     "avg_us": round(d["avg_us"], 2), "bytes_per_launch": int(d["bytes"])},
        "rhs": {"model_bytes": rhs_model_bytes(N, b), "t_us": round(t_rhs * 1e6, 2),
                "achieved": round(rhs_gbs, 1), "frac": round(rhs_gbs / PEAK_GBS, 4),
                "kernels_us": {k: round(pr[k][0] / pr[k][1] * 1e3, 2) for k in rhs_names if k in pr}},
        "pass": {"bytes": int(b_pass), "t_us": round(t_pass * 1e6, 2),
                 "achieved": round(b_pass / t_pass / 1e9, 1), "frac": round(b_pass / t_pass / 1e9 / PEAK_GBS, 4)},
        "kernels": {k: {kk: (round(vv, 2) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in per.items()},
    }


def cpu_baseline(budget_s: float, workers: int = 1) -> dict:
    """The FP64 oracle on the host, a bounded sample of steps: one thread
    (the reference's own CPU mode), or its FFTs on `workers` threads through
    scipy.fft (the elementwise work stays on one thread)."""
    from dnsjax_amd.fields import periodic_field
    from oracle import periodic as P
    from oracle import stepper as S

    L = 2 * math.pi
    op = P.PeriodicParams(nx=N, ny=N, nz=N, lx=L, ly=L, lz=L, re=RE, dt=DT)
    of = P.Flow(op)
    pc, ic = S.make_stepper(P.get_rhs, P.predict, P.correct, P.norm)
    u = periodic_field(N, N, N, L, L, L, RMS, SEED)

    def step(s):
        pred, rhs, err = pc(s, of)
        c = 0
        while err > TOL and c < 10:
            pred, rhs, err = ic(s, pred, rhs, of)
            c += 1
        return P.correct_velocity(pred, of), c

    P.fft_workers(workers)
    try:
        u, _ = step(u)  # first step excluded (__main__.py:130-134)
        n, t0 = 0, time.perf_counter()
        while True:
            u, _ = step(u)
            n += 1
            if time.perf_counter() - t0 > budget_s or n >= 50:
                break
        el = time.perf_counter() - t0
    finally:
        P.fft_workers(None)
    how = "1 thread" if workers == 1 else f"FFTs on {workers} threads (scipy.fft), elementwise work on 1"
    return {"value": n / el, "unit": "time-steps/s", "cores": workers, "kind": "port",
            "sample": f"{n} steps of the same 128^3 workload after 1 warm-up step, FP64 oracle, {how}"}


def learned_form(steps: int, collect_steps: int = 30, hidden: int = 32, train_iters: int = 1500) -> dict:
    """Mixed form with the learned corrector guess, against the plain forms
    over the same trajectory segment.

    The network is trained on the GPU (the library's FP32 training kernels)
    on the nonlinear-term history of a `collect_steps` run of the same
    workload; that run and the training are set-up, timed separately.
    Then, from the state it ends in, `steps` steps of each form are timed:
    plain FP64 (the reference's method), plain mixed, mixed with quadratic
    extrapolation only, and mixed with the learned guess.  The end-state
    energy of each mixed form is compared with the plain FP64 run's.
    """
    import torch

    from dnsjax_amd.driver import run
    from dnsjax_amd.flows import TriplyPeriodicFlow
    from dnsjax_amd.learned import Net, collect, fit_scales, make_dataset, mean_loss, run_guessed, train
    from dnsjax_amd.params import kolmogorov_2pi

    def mk(prec):
        return TriplyPeriodicFlow(kolmogorov_2pi(N, RE, DT, corrector_tolerance=TOL,
                                                 max_corrector_iterations=10, precision=prec))

    mixed, fp64 = mk("mixed"), mk("fp64")
    u0 = mixed.init_state(seed=SEED, rms=RMS)
    t0 = time.perf_counter()
    snaps, u_mid = collect(mixed, u0, collect_steps)
    sc = fit_scales(snaps)
    X, T = make_dataset(mixed, snaps, sc, 100000, seed=1)
    nv = 100000
    Xv, Tv, X, T = X[-nv:], T[-nv:], X[:-nv], T[:-nv]
    # the same network on the residual of cubic extrapolation
    sc3 = fit_scales(snaps, cubic=True)
    X3, T3 = make_dataset(mixed, snaps, sc3, 100000, seed=1)
    X3v, T3v, X3, T3 = X3[-nv:], T3[-nv:], X3[:-nv], T3[:-nv]
    del snaps
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    net = Net(hidden, seed=0)
    t0 = time.perf_counter()
    train(net, X, T, iters=train_iters, batch=16384, lr=3e-3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    loss, base = mean_loss(net, Xv, Tv)
    net3 = Net(hidden, seed=0)
    train(net3, X3, T3, iters=train_iters, batch=16384, lr=3e-3)
    loss3, base3 = mean_loss(net3, X3v, T3v)
    guesses = {"mixed_quadratic": (3, None, None), "mixed_learned": (2, net, sc),
               "mixed_cubic": (4, None, None), "mixed_learned_cubic": (2, net3, sc3)}

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t

    # one untimed pass of each form first (allocator growth, module loads)
    run(fp64, u_mid.clone(), 2)
    run(mixed, u_mid.clone(), 2)
    for mode, nn, ss in guesses.values():
        run_guessed(mixed, u_mid.clone(), 4, mode=mode, net=nn, scales=ss)
    rows = {}
    ref, t_ref = timed(lambda: run(fp64, u_mid.clone(), steps))
    rows["fp64_plain"] = (t_ref, ref.rhs_evals, ref.converged, ref.state, steps)
    r, t = timed(lambda: run(mixed, u_mid.clone(), steps))
    rows["mixed_plain"] = (t, r.rhs_evals, r.converged, r.state, steps)
    for name, (mode, nn, ss) in guesses.items():
        g, t = timed(lambda: run_guessed(mixed, u_mid.clone(), steps, mode=mode, net=nn, scales=ss))
        rows[name] = (t, g.rhs_evals, g.converged, g.state, steps)
    # the same forms with the corrector decided on the device, one captured
    # graph launch per step (no host read of the error per pass)
    from dnsjax_amd.graph import GraphStepper
    from dnsjax_amd.learned import GuessedGraphStepper

    # Corrector slots per captured step: a step that needs more passes than
    # the graph holds is flagged unconverged by the device counters, so a
    # run whose counters show none is identical to the eager loop's.
    for tag, slots in (("", None), ("_4slots", 4)):
        gp = GraphStepper(fp64, u_mid, slots=slots, capacity=steps + 4)
        gp.step(2)
        torch.cuda.synchronize()
        _, t = timed(lambda: gp.step(steps))
        st = gp.step_stats()[0]
        rows["fp64_plain_graph" + tag] = (t, sum(c + 2 for c in gp.counts()[0][2:]), st["unconverged"] == 0,
                                          gp.state, gp.steps)
        for name in ("mixed_learned", "mixed_cubic", "mixed_learned_cubic"):
            mode, nn, ss = guesses[name]
            gg = GuessedGraphStepper(mixed, u_mid, mode=mode, net=nn, scales=ss, slots=slots, capacity=steps + 5)
            gg.step(2)
            torch.cuda.synchronize()
            k0 = gg.steps
            _, t = timed(lambda: gg.step(steps))
            rows[name + "_graph" + tag] = (t, sum(c + 1 for c in gg.passes()[k0:]),
                                           gg.step_stats()["unconverged"] == 0, gg.state, gg.steps)
    # The graph steppers' end states include their start-up and warm-up
    # steps: each is compared with an eager FP64 run of the same length
    # (energy E', dissipation D and the velocity field).
    refs = {steps: ref.state}

    def ref_at(n):
        if n not in refs:
            refs[n] = run(fp64, u_mid.clone(), n).state
        return refs[n]

    def errors(st, n):
        r = ref_at(n)
        a, b = fp64.get_stats(st), fp64.get_stats(r)
        rel = lambda k: abs(float(a[k]) - float(b[k])) / abs(float(b[k]))  # noqa: E731
        return {"energy_rel_err_vs_fp64": rel("E'"), "dissipation_rel_err_vs_fp64": rel("D"),
                "velocity_rel_l2_vs_fp64": float(torch.linalg.vector_norm(st - r) / torch.linalg.vector_norm(r))}

    out = {}
    for name, (t, rhs, conv, st, n) in rows.items():
        out[name] = {"steps_per_s": round(steps / t, 3), "ms_per_step": round(1e3 * t / steps, 4),
                     "rhs_evals_per_step": round(rhs / steps, 3), "converged": conv, **errors(st, n)}
    # The configuration itself: 500 steps of each form from the seeded
    # initial state (the network was trained on the first 30 steps' history),
    # every step's corrector iterated to the tolerance; end-state errors of
    # each mixed form against the FP64 run.
    long_n = 500
    long_ref, t_lr = timed(lambda: run(fp64, u0.clone(), long_n))
    long_forms = {}
    for name, fn in (("mixed", lambda: run(mixed, u0.clone(), long_n)),
                     ("mixed_quadratic", lambda: run_guessed(mixed, u0.clone(), long_n, mode=3)),
                     ("mixed_learned", lambda: run_guessed(mixed, u0.clone(), long_n, mode=2, net=net, scales=sc)),
                     ("mixed_cubic", lambda: run_guessed(mixed, u0.clone(), long_n, mode=4)),
                     ("mixed_learned_cubic",
                      lambda: run_guessed(mixed, u0.clone(), long_n, mode=2, net=net3, scales=sc3))):
        rr, tt = timed(fn)
        a, b = fp64.get_stats(rr.state), fp64.get_stats(long_ref.state)
        rel_ = lambda k: abs(float(a[k]) - float(b[k])) / abs(float(b[k]))  # noqa: E731
        long_forms[name] = {
            "steps_per_s": round(long_n / tt, 3), "speedup_vs_fp64": round(t_lr / tt, 3),
            "rhs_evals_per_step": round(rr.rhs_evals / long_n, 3), "converged": rr.converged,
            "max_error": max(rr.errors), "escalations": rr.escalations,
            "energy_rel_err_vs_fp64": rel_("E'"), "dissipation_rel_err_vs_fp64": rel_("D"),
            "velocity_rel_l2_vs_fp64": float(torch.linalg.vector_norm(rr.state - long_ref.state)
                                             / torch.linalg.vector_norm(long_ref.state))}
    long_run = {"steps": long_n, "from": "the seeded initial state", "fp64_steps_per_s": round(long_n / t_lr, 3),
                "fp64_rhs_evals_per_step": round(long_ref.rhs_evals / long_n, 3),
                "fp64_converged": long_ref.converged, **long_forms}
    rate = lambda k: out[k]["steps_per_s"]  # noqa: E731
    speedups = {}
    for name in guesses:
        speedups[name] = round(rate(name) / rate("fp64_plain"), 3)
        for tag in ("_graph", "_graph_4slots"):
            if name + tag in out:
                speedups[name + tag] = round(rate(name + tag) / rate("fp64_plain" + tag), 3)
    return {"forms": out, "long_run": long_run, "speedup_vs_fp64": speedups,
            "net": {"hidden": hidden, "train_iters": train_iters, "train_s": round(t_train, 3),
                    "data_s": round(t_data, 3), "samples": int(X.shape[0]),
                    "heldout_loss": round(loss, 4), "heldout_loss_without_net": round(base, 4),
                    "cubic_heldout_loss": round(loss3, 4), "cubic_heldout_loss_without_net": round(base3, 4)},
            "segment": f"{steps} steps after a {collect_steps}-step training run from the workload's initial state"}


def walls_model_bytes(nx: int, ny: int, nz: int, p: int, b: int = 16) -> dict:
    """Fixed algorithmic byte models of one plane Couette RHS (five-pass
    pencil model: b * Nx1 * Ny * (6 Nz1 + 18 Pz)) and one FP64 influence-
    matrix step (5 S + 2 ab + 2 piv + 4 vec + M_inv: the fields read and
    written, the two band factors and pivots, the homogeneous solutions and
    the 2x2 inverses), with Nx1 = nx/2, Nz1 = nz - 1, Pz = 3 nz / 2."""
    nx1, nz1, pz = nx // 2, nz - 1, 3 * nz // 2
    m = nx1 * nz1
    spec = 3 * m * ny * 16
    imm = 5 * spec + 2 * m * (3 * p + 1) * ny * 8 + 2 * m * ny * 4 + 4 * m * ny * 8 + m * 4 * 8
    return {"rhs": b * nx1 * ny * (6 * nz1 + 18 * pz), "imm": imm}


def walls_roofline(prof: dict, nx: int, ny: int, nz: int, p: int) -> dict:
    """Time per call of the RHS (curl and z-inverse, x pass, z-forward and
    transpose) and of the IMM step from the per-kernel event times, against
    the byte models and the 8 TB/s HBM peak."""
    us = {k: 1e3 * ms / n for k, (ms, n) in prof.items()}
    mb = walls_model_bytes(nx, ny, nz, p)
    t_rhs = us.get("walls_z_inverse", 0.0) + us.get("walls_x_nonlinear", 0.0) + us.get("walls_z_forward", 0.0)
    t_imm = us.get("walls_imm", 0.0)
    out = {}
    for name, t, by in (("rhs", t_rhs, mb["rhs"]), ("imm", t_imm, mb["imm"])):
        gbs = by / (t * 1e-6) / 1e9 if t > 0 else 0.0
        out[name] = {"model_bytes": by, "t_us": round(t, 2), "achieved": round(gbs, 1), "peak": 8000.0,
                     "unit": "GB/s", "frac": round(gbs / 8000.0, 4)}
    return out


def walls_form(steps: int, warmup: int = 2, collect_steps: int = 30, train_iters: int = 1500,
               tol: float = 1e-5) -> dict:
    """Plane Couette 64 x 128 x 64 (CGL points in y), Lx = 2 pi, Lz = pi,
    Re 400, fd_order 4, dt 0.01, seeded perturbation RMS 0.05 (seed 2),
    corrector tolerance `tol` (the reference's default is 1e-5) and limit 10.

    After `warmup` FP64 steps, a `collect_steps` mixed run records the
    nonlinear-term history on which the per-point network is trained (set-up,
    timed separately).  From the state that run ends in, time-steps/s of the
    driver loop are timed over the same segment for the plain FP64 method
    (with per-kernel event times), the mixed form, and the mixed form started
    from the extrapolated guess (mode 3) and from the learned guess (mode 2);
    energy of each end state against the FP64 run's."""
    import torch

    from dnsjax_amd.driver import run
    from dnsjax_amd.flows import PlaneCouetteFlow
    from dnsjax_amd.learned import Net, collect, fit_scales, make_dataset, mean_loss, run_guessed, train
    from dnsjax_amd.params import Params

    def mk(prec):
        return PlaneCouetteFlow(Params(system="plane-couette", nx=64, ny=128, nz=64, lx=2 * math.pi, lz=math.pi,
                                       re=400.0, dt=0.01, fd_order=4, corrector_tolerance=tol,
                                       max_corrector_iterations=10, precision=prec))

    flow, mixed = mk("fp64"), mk("mixed")
    u0 = flow.init_state(seed=WALLS_SEED, rms=0.05)
    state = run(flow, u0.clone(), warmup).state
    t0 = time.perf_counter()
    snaps, state = collect(mixed, state, collect_steps)
    sc = fit_scales(snaps)
    X, T = make_dataset(mixed, snaps, sc, 100000, seed=1)
    nv = 100000
    Xv, Tv, X, T = X[-nv:], T[-nv:], X[:-nv], T[:-nv]
    del snaps
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    net = Net(32, seed=0)
    t0 = time.perf_counter()
    train(net, X, T, iters=train_iters, batch=16384, lr=3e-3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    loss, base = mean_loss(net, Xv, Tv)
    del X, T, Xv, Tv

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t

    # one untimed pass of each form first (allocator growth, module loads)
    run(flow, state.clone(), 2)
    run(mixed, state.clone(), 2)
    for mode in (3, 2):
        run_guessed(mixed, state.clone(), 3, mode=mode, net=net, scales=sc)
    r, el = timed(lambda: run(flow, state.clone(), steps))
    flow._native.profile_begin()
    r2 = run(flow, state.clone(), steps)
    torch.cuda.synchronize()
    prof = flow._native.profile_end()
    e_ref = float(flow.get_stats(r.state)["E'"])
    rm, tm = timed(lambda: run(mixed, state.clone(), steps))
    rg, tg = timed(lambda: run_guessed(mixed, state.clone(), steps, mode=3))
    rl, tl = timed(lambda: run_guessed(mixed, state.clone(), steps, mode=2, net=net, scales=sc))
    # The plain FP64 and the mixed + learned forms again with the corrector
    # decided on the device, one captured graph launch per step.  Each graph
    # holds one corrector slot more than the largest count of its eager run
    # (a skipped slot still costs its launches); a step needing more would
    # be flagged unconverged by the device counters.
    from dnsjax_amd.graph import GraphStepper
    from dnsjax_amd.learned import GuessedGraphStepper

    gp = GraphStepper(flow, state, slots=max(r.counts) + 1, capacity=steps + 4)
    gp.step(2)
    _, tgp = timed(lambda: gp.step(steps))
    gg = GuessedGraphStepper(mixed, state, mode=2, net=net, scales=sc, slots=max(rl.passes) + 1,
                             capacity=steps + 4)
    gg.step(2)
    _, tgg = timed(lambda: gg.step(steps))
    # their end states include their start-up and warm-up steps (gg.steps of
    # them for the guessed stepper): compare with an eager FP64 run of the
    # same length
    e_graph_ref = float(flow.get_stats(run(flow, state.clone(), gg.steps).state)["E'"])

    def err(st):
        return abs(float(flow.get_stats(st)["E'"]) - e_ref) / abs(e_ref)

    def err_graph(st):
        return abs(float(flow.get_stats(st)["E'"]) - e_graph_ref) / abs(e_graph_ref)

    # the configuration itself: 200 steps of each form from the seeded state
    long_n = 200
    long_ref, t_lr = timed(lambda: run(flow, u0.clone(), long_n))
    e_long = float(flow.get_stats(long_ref.state)["E'"])
    long_run = {"steps": long_n, "from": "the seeded initial state", "fp64_steps_per_s": round(long_n / t_lr, 3),
                "fp64_converged": long_ref.converged,
                "fp64_rhs_evals_per_step": round(long_ref.rhs_evals / long_n, 3)}
    for name, fn in (("mixed", lambda: run(mixed, u0.clone(), long_n)),
                     ("mixed_learned", lambda: run_guessed(mixed, u0.clone(), long_n, mode=2, net=net, scales=sc))):
        rr, tt = timed(fn)
        long_run[name] = {"steps_per_s": round(long_n / tt, 3), "speedup_vs_fp64": round(t_lr / tt, 3),
                          "converged": rr.converged, "rhs_evals_per_step": round(rr.rhs_evals / long_n, 3),
                          "escalations": rr.escalations,
                          "energy_rel_err_vs_fp64": abs(float(flow.get_stats(rr.state)["E'"]) - e_long) / abs(e_long)}

    return {"workload": f"plane Couette 64x128x64, p=4, Re 400, dt 0.01, RMS 0.05 (seed {WALLS_SEED}), tol {tol:g}"
                        + (" (reference default)" if tol == 1e-5 else ""),
            "steps_per_s": round(steps / el, 3), "ms_per_step": round(1e3 * el / steps, 4),
            "corrector_iters_per_step": round(sum(r.counts) / max(len(r.counts), 1), 3),
            "converged": r.converged and r2.converged,
            "kernels_avg_us": {k: round(1e3 * ms / n, 2) for k, (ms, n) in prof.items()},
            "roofline": walls_roofline(prof, 64, 128, 64, 4),
            "mixed": {"steps_per_s": round(steps / tm, 3), "rhs_evals_per_step": round(rm.rhs_evals / steps, 3),
                      "converged": rm.converged, "energy_rel_err_vs_fp64": err(rm.state)},
            "mixed_guess": {"steps_per_s": round(steps / tg, 3), "rhs_evals_per_step": round(rg.rhs_evals / steps, 3),
                            "converged": rg.converged, "energy_rel_err_vs_fp64": err(rg.state),
                            "speedup_vs_fp64": round(el / tg, 3),
                            "guess": "quadratic extrapolation of N fed to the predictor's solve"},
            "mixed_learned": {"steps_per_s": round(steps / tl, 3),
                              "rhs_evals_per_step": round(rl.rhs_evals / steps, 3),
                              "converged": rl.converged, "energy_rel_err_vs_fp64": err(rl.state),
                              "speedup_vs_fp64": round(el / tl, 3),
                              "guess": "quadratic extrapolation plus the per-point network's residual",
                              "net": {"hidden": 32, "train_iters": train_iters, "train_s": round(t_train, 3),
                                      "data_s": round(t_data, 3), "heldout_loss": round(loss, 4),
                                      "heldout_loss_without_net": round(base, 4)}},
            "graph": {"fp64": {"steps_per_s": round(steps / tgp, 3), "slots": gp.members[0].slots,
                               "converged": gp.step_stats()[0]["unconverged"] == 0},
                      "mixed_learned": {"steps_per_s": round(steps / tgg, 3), "slots": gg.slots,
                                        "converged": gg.step_stats()["unconverged"] == 0,
                                        "energy_rel_err_vs_fp64": err_graph(gg.state)},
                      "speedup_vs_fp64": round(tgp / tgg, 3)},
            "long_run": long_run,
            "fp64_rhs_evals_per_step": round(r.rhs_evals / steps, 3),
            "segment": f"{steps} steps after a {collect_steps}-step mixed data run from the workload's state"}


def ensemble_form(steps: int, n_streams: int = 8) -> dict:
    """64 Kolmogorov trajectories (16 each at 32^3, 48^3, 64^3, 96^3; Re drawn
    uniformly in [20, 80] with seed 3; dt 0.01; seeds 300 + i, RMS 0.05;
    mixed form, tol 1e-10), advanced by one captured graph launch per step,
    against the same 64 trajectories run one after another with the eager
    driver.  Aggregate time-steps/s over all members.

    Two captured forms: every member's step a branch of its own (one launch
    per kernel and member), and batched (the 16 members of a size advance by
    one launch per kernel over all of them).  The members of a size run the
    kernel variants the autotuner chose for its first member.  Every
    captured step launches a fixed number of corrector slots (skipped slots
    still launch their kernels, which exit at once): each member takes one
    slot more than the largest iteration count of its own sequential run
    (batched: the largest of its size); the device counters flag any step
    that would need more (`unconverged`), and every member's counts must
    equal its sequential run's."""
    import numpy as np
    import torch

    from dnsjax_amd.driver import run
    from dnsjax_amd.flows import TriplyPeriodicFlow
    from dnsjax_amd.graph import Ensemble
    from dnsjax_amd.params import kolmogorov_2pi

    sizes = [32] * 16 + [48] * 16 + [64] * 16 + [96] * 16
    res = np.random.default_rng(3).uniform(20.0, 80.0, len(sizes))
    t0 = time.perf_counter()
    flows, pinned = [], {}
    for n, r in zip(sizes, res):
        kw = {} if n not in pinned else dict(zip(("tile_cols", "variants"), pinned[n]))
        f = TriplyPeriodicFlow(kolmogorov_2pi(n, float(r), 0.01, corrector_tolerance=TOL,
                                              max_corrector_iterations=10, precision="mixed", **kw))
        pinned.setdefault(n, f._native.variants())
        flows.append(f)
    states = [f.init_state(seed=300 + i, rms=0.05) for i, f in enumerate(flows)]
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    m = len(flows)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    seq_counts = [run(f, s.clone(), steps).counts for f, s in zip(flows, states)]
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    slots = [min(max(cs) + 1, 10) for cs in seq_counts]
    out = {}
    for label, batched in (("branches", False), ("batched", True)):
        ens = Ensemble(flows, states, n_streams=n_streams, slots=slots, capacity=steps, batched=batched)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ens.step(steps)
        torch.cuda.synchronize()
        t_ens = time.perf_counter() - t0
        st = ens.step_stats()
        out[label] = {"value": round(m * steps / t_ens, 2), "speedup_vs_sequential": round(t_seq / t_ens, 3),
                      "counts_match_sequential": ens.counts() == seq_counts,
                      "unconverged": sum(x["unconverged"] for x in st),
                      "slots_mean": round(sum(mb.slots for mb in ens.members) / m, 2)}
        if batched:
            out[label]["classes"] = len(ens.plans[0])
            out[label]["launches_per_class_step"] = max(getattr(pl, "launches", 0) for pl in ens.plans[0])
        del ens
    b = out["batched"]
    return {"members": m, "sizes": "16 x {32,48,64,96}^3", "steps": steps, "n_streams": n_streams,
            "slots": {"min": min(slots), "max": max(slots), "mean": round(sum(slots) / m, 2)},
            "value": b["value"], "unit": "member time-steps/s", "form": "batched",
            "sequential_value": round(m * steps / t_seq, 2),
            "speedup_vs_sequential": b["speedup_vs_sequential"],
            "counts_match_sequential": b["counts_match_sequential"] and out["branches"]["counts_match_sequential"],
            "unconverged": b["unconverged"] + out["branches"]["unconverged"],
            "forms": out, "setup_s": round(t_setup, 2),
            "note": "each graph timing includes the first launch of each of its two graphs"}


def profile_workload(steps: int) -> None:
    """The workload profiled by rocprofv3 (kernel trace and PMC passes).

    First a plain device-to-device copy of CAL_BYTES (the calibration
    dispatch for the FETCH_SIZE / WRITE_SIZE counters: it must read and
    write exactly CAL_BYTES), then `steps` steps of the FP64 and of the
    mixed form at 128^3, then a short training run of the corrector-guess
    network and `steps` steps of the mixed form with the learned guess, then
    `steps` steps of plane Couette 64 x 128 x 64 in both forms.  The kernel
    variants are pinned (16-column tiles, paired transforms, non-temporal
    z-inverse stores in FP64; the fused influence-matrix step), the
    autotuner's usual choices at these sizes, so that the separate counter
    passes dispatch the same kernels.
    """
    import torch

    cal = 256 * 1024 * 1024
    a = torch.ones(cal // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    print(f"calibration copy: {cal} bytes read, {cal} bytes written", flush=True)
    del a, b
    from dnsjax_amd.flows import TriplyPeriodicFlow
    from dnsjax_amd.params import kolmogorov_2pi

    for prec in ("fp64", "mixed"):
        p = kolmogorov_2pi(N, RE, DT, corrector_tolerance=TOL, max_corrector_iterations=10,
                           precision=prec, tile_cols=16, variants=32 if prec == "fp64" else 0)
        flow = TriplyPeriodicFlow(p)
        state = flow.init_state(seed=SEED, rms=RMS)
        state, counts = run_steps(flow, state, steps)
        torch.cuda.synchronize()
        print(f"{prec}: {steps} steps, corrector iterations {counts}", flush=True)
    # the learned guess: a short training run, then `steps` guessed steps
    from dnsjax_amd.learned import Net, collect, fit_scales, make_dataset, run_guessed, train

    snaps, state = collect(flow, state, 6)
    sc = fit_scales(snaps)
    X, T = make_dataset(flow, snaps, sc, 16384, seed=1)
    net = Net(32, seed=0)
    train(net, X, T, iters=20, batch=16384)
    g = run_guessed(flow, state, steps, mode=2, net=net, scales=sc)
    torch.cuda.synchronize()
    print(f"mixed + learned guess: {steps} steps, corrector passes {g.passes}", flush=True)
    del flow, state, snaps, X, T, g
    # plane Couette 64 x 128 x 64 (p = 4), both forms
    from dnsjax_amd.flows import PlaneCouetteFlow
    from dnsjax_amd.params import Params

    for prec in ("fp64", "mixed"):
        wf = PlaneCouetteFlow(Params(system="plane-couette", nx=64, ny=128, nz=64, lx=2 * math.pi, lz=math.pi,
                                     re=400.0, dt=0.01, fd_order=4, corrector_tolerance=1e-5,
                                     max_corrector_iterations=10, precision=prec, tile_cols=16, variants=16))
        ws = wf.init_state(seed=WALLS_SEED, rms=0.05)
        ws, counts = run_steps(wf, ws, steps)
        torch.cuda.synchronize()
        print(f"plane Couette {prec}: {steps} steps, corrector iterations {counts}", flush=True)


def profile_cubic(steps: int) -> None:
    """The rocprofv3 workload of the fastest periodic form: config 1's
    128^3 workload, mixed steps started from cubic extrapolation, then as
    many plain FP64 steps, each run after a short untimed warm-up."""
    import torch

    from dnsjax_amd.driver import run
    from dnsjax_amd.flows import TriplyPeriodicFlow
    from dnsjax_amd.learned import run_guessed
    from dnsjax_amd.params import kolmogorov_2pi

    def mk(prec):
        return TriplyPeriodicFlow(kolmogorov_2pi(N, RE, DT, corrector_tolerance=TOL, max_corrector_iterations=10,
                                                 precision=prec))

    mixed, fp64 = mk("mixed"), mk("fp64")
    u0 = mixed.init_state(seed=SEED, rms=RMS)
    run_guessed(mixed, u0.clone(), 6, mode=4)
    run(fp64, u0.clone(), 3)
    torch.cuda.synchronize()
    t = time.perf_counter()
    g = run_guessed(mixed, u0.clone(), steps, mode=4)
    torch.cuda.synchronize()
    print(f"mixed + cubic extrapolation: {steps} steps in {time.perf_counter() - t:.4f} s, passes {g.passes}",
          flush=True)
    t = time.perf_counter()
    r = run(fp64, u0.clone(), steps)
    torch.cuda.synchronize()
    print(f"fp64: {steps} steps in {time.perf_counter() - t:.4f} s, corrector iterations {r.counts}", flush=True)


def walls_large_form(steps: int, collect_steps: int = 8, train_iters: int = 300, pairs: int = 20000) -> dict:
    """The learned-guess configuration: plane Couette 128 x 256 x 128 CGL
    points, fd_order 6, Re 1000, dt 0.005, corrector tolerance 1e-10,
    seeded perturbation RMS 0.05 (seed 4), `steps` steps (1000).

    A `collect_steps` mixed run from the seeded state records the nonlinear-
    term history; `pairs` (predictor history -> converged corrector N) pairs
    drawn from it train the per-(kz, kx) profile network (two hidden layers
    of 256, bf16 matrix-core GEMMs in both passes; set-up, timed
    separately).  Then each form runs the configuration from the seeded
    state: plain FP64 (the reference's method), mixed, mixed started from
    quadratic extrapolation and mixed started from the network's guess;
    corrector evaluations (RHS + IMM) per step, steps/s and end-state
    energy against the FP64 run."""
    import torch

    from dnsjax_amd.driver import run
    from dnsjax_amd.fields import wall_field
    from dnsjax_amd.flows import PlaneCouetteFlow
    from dnsjax_amd.learned import collect, run_guessed
    from dnsjax_amd.params import Params
    from dnsjax_amd.profile_net import ProfileNet, fit_scales, make_dataset, mean_loss, train

    nx, ny, nz = 128, 256, 128

    def mk(prec):
        return PlaneCouetteFlow(Params(system="plane-couette", nx=nx, ny=ny, nz=nz, lx=2 * math.pi,
                                       lz=math.pi, re=1000.0, dt=0.005, fd_order=6,
                                       corrector_tolerance=1e-10, max_corrector_iterations=20,
                                       precision=prec))

    t0 = time.perf_counter()
    fp64, mixed = mk("fp64"), mk("mixed")
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    u0 = fp64.to_device(wall_field(nx, ny, nz, 2 * math.pi, math.pi, 0.05, seed=4))
    t0 = time.perf_counter()
    snaps, _ = collect(mixed, u0.clone(), collect_steps)
    sc = fit_scales(snaps)
    net = ProfileNet(ny, hidden=256, seed=0)
    X, T = make_dataset(mixed, snaps, sc, pairs + 2000, net.din, seed=1)
    del snaps
    Xv, Tv, X, T = X[pairs:], T[pairs:], X[:pairs], T[:pairs]
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    train(net, X, T, iters=train_iters, batch=1024, lr=1e-3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    loss, base = mean_loss(net, Xv, Tv)
    del X, T, Xv, Tv

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t

    plain, tp = timed(lambda: run(fp64, u0.clone(), steps))
    e_p = float(fp64.get_stats(plain.state)["E'"])
    forms = {}
    for name, fn in (("mixed", lambda: run(mixed, u0.clone(), steps)),
                     ("mixed_quadratic", lambda: run_guessed(mixed, u0.clone(), steps, mode=3)),
                     ("mixed_learned", lambda: run_guessed(mixed, u0.clone(), steps, mode=2, net=net, scales=sc))):
        r, t = timed(fn)
        e = float(fp64.get_stats(r.state)["E'"])
        forms[name] = {"steps_per_s": round(len(r.errors) / t, 3),
                       "speedup_vs_fp64": round((tp / len(plain.counts)) / (t / len(r.errors)), 3),
                       "rhs_evals_per_step": round(r.rhs_evals / len(r.errors), 3), "converged": r.converged,
                       "steps_run": len(r.errors), "escalations": r.escalations,
                       "energy_rel_err_vs_fp64": abs(e - e_p) / abs(e_p)}
    return {"workload": "plane Couette 128x256x128, p=6, Re 1000, dt 0.005, RMS 0.05 (seed 4), tol 1e-10",
            "steps": steps, "from": "the seeded initial state", "setup_s": round(t_setup, 2),
            "fp64": {"steps_per_s": round(len(plain.counts) / tp, 3),
                     "rhs_evals_per_step": round(plain.rhs_evals / len(plain.counts), 3),
                     "corrector_iters_per_step": round(sum(plain.counts) / len(plain.counts), 3),
                     "converged": plain.converged, "steps_run": len(plain.counts),
                     "tuning": fp64._native.tuning.strip()},
            **forms,
            "speedup_vs_fp64": forms["mixed_learned"]["speedup_vs_fp64"],
            "net": {"kind": "per-(kz, kx) profile MLP", "din": net.din, "hidden": [net.H, net.H], "dout": net.dout,
                    "weights": "bf16 on the matrix cores, FP32 master copy", "pairs": pairs,
                    "train_iters": train_iters, "data_s": round(t_data, 2), "train_s": round(t_train, 2),
                    "heldout_loss": round(loss, 5), "heldout_loss_without_net": round(base, 5)}}


def profile_imm(calls: int) -> None:
    """The plane Couette influence-matrix step (predict: one IMM on the
    state's RHS) in the three band-solve forms (chunk maps in five launches
    or fused into two, sequential sweeps),
    at 64x128x64 (p=4) and 128x256x128 (p=6), `calls` timed calls each after
    five warm-up calls; prints the time per call, a checksum of the result
    (equal for both forms up to rounding) and the autotuner's choice."""
    import math

    import torch

    from dnsjax_amd.flows import PlaneCouetteFlow
    from dnsjax_amd.params import Params

    # the PMC passes' calibration dispatch, as in profile_workload
    cal = 256 * 1024 * 1024
    a = torch.ones(cal // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    print(f"calibration copy: {cal} bytes read, {cal} bytes written", flush=True)
    del a, b
    for nx, ny, nz, po in ((64, 128, 64, 4), (128, 256, 128, 6)):
        base = dict(system="plane-couette", nx=nx, ny=ny, nz=nz, lx=2 * math.pi, lz=math.pi, re=400.0,
                    dt=0.01, fd_order=po, corrector_tolerance=1e-5)
        for form, variants in (("chunked", 0), ("fused", 16), ("sequential", 8)):
            flow = PlaneCouetteFlow(Params(tile_cols=16, variants=variants, **base))
            nat = flow._native
            state = flow.init_state(seed=1, rms=0.05)
            rhs = nat.get_rhs(state)
            out = nat.empty()
            for _ in range(5):
                nat.predict(state, rhs, out)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record()
            for _ in range(calls):
                nat.predict(state, rhs, out)
            e1.record()
            torch.cuda.synchronize()
            print(f"{nx}x{ny}x{nz} p={po} {form}: {e0.elapsed_time(e1) * 1e3 / calls:.1f} us per IMM step, "
                  f"checksum {float(out.abs().sum()):.15g}", flush=True)
            del flow, nat
        flow = PlaneCouetteFlow(Params(**base))
        print(f"{nx}x{ny}x{nz} p={po} autotuner: {flow._native.tuning}", flush=True)
        del flow


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--walls-steps", type=int, default=20,
                    help="steps of the plane Couette 64x128x64 form (0 skips it)")
    ap.add_argument("--learned-steps", type=int, default=40,
                    help="steps of each form in the mixed+learned A/B (0 skips it)")
    ap.add_argument("--ensemble-steps", type=int, default=50,
                    help="steps of the 64-member ensemble form (0 skips it)")
    ap.add_argument("--walls-large-steps", type=int, default=800,
                    help="steps of the plane Couette 128x256x128 learned-guess configuration (0 skips "
                         "it; its FP64 corrector stops converging within 20 iterations near step 835)")
    ap.add_argument("--profile-walls", action="store_true",
                    help="run a few steps of the plane Couette form for rocprofv3 and exit")
    ap.add_argument("--profile-workload", action="store_true",
                    help="run the rocprofv3 workload (calibration copy + a few steps) and exit")
    ap.add_argument("--profile-imm", action="store_true",
                    help="time the plane Couette influence-matrix step in its three forms and exit")
    ap.add_argument("--profile-cubic", action="store_true",
                    help="run --steps mixed steps from cubic extrapolation and FP64 steps at 128^3 and exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); the multi-GPU slab "
              "decomposition is described, not implemented.", flush=True)
        sys.exit(2)
    if a.profile_workload:
        profile_workload(a.steps)
        return
    if a.profile_walls:
        print(json.dumps(walls_form(a.steps, warmup=1)), flush=True)
        return
    if a.profile_imm:
        profile_imm(a.steps)
        return
    if a.profile_cubic:
        profile_cubic(a.steps)
        return
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    progress("periodic 128^3: fp64 and mixed")
    fp64 = gpu_form("fp64", a.steps, a.warmup)
    mixed = gpu_form("mixed", a.steps, a.warmup)
    rf64 = roofline(fp64, 16)
    rmix = roofline(mixed, 8, split=True)
    out = {
        "metric": METRIC,
        "value": round(fp64["steps_per_s"], 3),
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(fp64["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded solenoidal field, E(k) ~ k^4 exp(-2 (k/4)^2))",
        "config": {"workload": "kolmogorov 128^3 (192^3 padded), 2pi box, Re=100, dt=0.005, c=0.5, tol=1e-10, seed 1 RMS 0.1",
                   "form": "plain fp64", "corrector_iters_per_step": round(fp64["corrector_iters_per_step"], 3)},
        "roofline": rf64["dominant"],
        "rhs_roofline": rf64["rhs"],
        "pass_roofline": rf64["pass"],
        "kernels": rf64["kernels"],
        "mixed": {"value": round(mixed["steps_per_s"], 3), "ms_per_step": round(mixed["ms_per_step"], 4),
                  "speedup_vs_fp64": round(mixed["steps_per_s"] / fp64["steps_per_s"], 3),
                  "corrector_iters_per_step": round(mixed["corrector_iters_per_step"], 3),
                  "roofline": rmix["dominant"], "rhs_roofline": rmix["rhs"], "pass_roofline": rmix["pass"],
                  "kernels": rmix["kernels"]},
    }
    if a.walls_steps > 0:
        progress("plane Couette 64x128x64")
        out["walls"] = walls_form(a.walls_steps)
        out["walls_mixed_learned_vs_fp64"] = out["walls"]["mixed_learned"]["speedup_vs_fp64"]
        # the same forms where the tolerance no longer sets the energy
        # difference (the wall corrector contracts ~30x per pair of passes,
        # so at 1e-5 a step stopping at the first pass differs by O(tol) from
        # one that paid the next pair)
        wt = walls_form(a.walls_steps, tol=1e-8)
        out["walls_tol_1e-8"] = {
            "fp64_steps_per_s": wt["steps_per_s"], "fp64_rhs_evals_per_step": wt["fp64_rhs_evals_per_step"],
            **{f: {k: wt[f][k] for k in ("steps_per_s", "rhs_evals_per_step", "energy_rel_err_vs_fp64",
                                         "converged")} for f in ("mixed", "mixed_guess", "mixed_learned")},
            "mixed_learned_vs_fp64": wt["mixed_learned"]["speedup_vs_fp64"], "long_run": wt["long_run"]}
    if a.learned_steps > 0:
        progress("periodic 128^3: learned guess")
        lf = learned_form(a.learned_steps)
        out["mixed_learned"] = lf
        out["mixed_learned_vs_fp64"] = lf["speedup_vs_fp64"]
    if a.ensemble_steps > 0:
        progress("64-member ensemble")
        out["ensemble"] = ensemble_form(a.ensemble_steps)
    if a.walls_large_steps > 0:
        progress("plane Couette 128x256x128 with the profile network")
        out["walls_large"] = walls_large_form(a.walls_large_steps)
    if not a.no_cpu:
        progress("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(a.cpu_seconds)
        ncores = min(16, len(os.sched_getaffinity(0)))
        out["cpu_baseline_all_cores"] = cpu_baseline(a.cpu_seconds, workers=ncores) if ncores > 1 else None
    else:
        out["cpu_baseline"] = None
    # the headline figures once more, last on the line (a reader of the
    # line's tail sees them)
    summ = {"fp64_steps_per_s": out["value"], "mixed_steps_per_s": out["mixed"]["value"],
            "rhs_roofline_frac_fp64": out["rhs_roofline"]["frac"],
            "rhs_roofline_frac_mixed": out["mixed"]["rhs_roofline"]["frac"]}
    if "mixed_learned" in out:
        lr = out["mixed_learned"]["long_run"]
        summ["periodic_500_steps"] = {
            f: {k: lr[f][k] for k in ("speedup_vs_fp64", "rhs_evals_per_step", "converged", "escalations",
                                      "energy_rel_err_vs_fp64", "dissipation_rel_err_vs_fp64",
                                      "velocity_rel_l2_vs_fp64")}
            for f in ("mixed", "mixed_quadratic", "mixed_learned", "mixed_cubic", "mixed_learned_cubic")}
        summ["periodic_500_steps"]["fp64_rhs_evals_per_step"] = lr["fp64_rhs_evals_per_step"]
    if "walls" in out:
        summ["walls_200_steps"] = {t: {f: {k: w["long_run"][f][k] for k in ("speedup_vs_fp64", "converged",
                                                                               "energy_rel_err_vs_fp64")}
                                       for f in ("mixed", "mixed_learned")}
                                   for t, w in (("tol_1e-5", out["walls"]), ("tol_1e-8", out["walls_tol_1e-8"]))}
    if "walls_large" in out:
        wl = out["walls_large"]
        summ[f"walls_128x256x128_{wl['fp64']['steps_run']}_steps"] = {
            f: {k: wl[f][k] for k in ("speedup_vs_fp64", "rhs_evals_per_step", "converged",
                                      "energy_rel_err_vs_fp64")} for f in ("mixed", "mixed_quadratic", "mixed_learned")}
        summ[f"walls_128x256x128_{wl['fp64']['steps_run']}_steps"]["fp64_rhs_evals_per_step"] = wl["fp64"]["rhs_evals_per_step"]
    if "ensemble" in out:
        summ["ensemble_member_steps_per_s"] = out["ensemble"]["value"]
        summ["ensemble_speedup_vs_sequential"] = out["ensemble"]["speedup_vs_sequential"]
    out["summary"] = summ
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
