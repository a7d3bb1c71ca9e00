"""A device context of libdnsmi for one configuration, over torch tensors.

Tensors are the device buffers (complex128, contiguous, on the current HIP
device); every call is enqueued on torch's current stream.  Nothing here
computes on the CPU.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from .params import Params


def _ptr(t: torch.Tensor | None):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class NativeContext:
    """Owns one ``dnsmi_ctx`` (set-up tables, FFT work buffers)."""

    def __init__(self, params: Params):
        self.lib = _lib.load()
        if not torch.cuda.is_available():
            raise RuntimeError("dnsjax_amd needs a HIP device (torch.cuda.is_available() is False)")
        self.params = p = params
        cfg = _lib.DnsmiConfig(
            system=_lib.SYSTEMS[p.system], nx=p.nx, ny=p.ny, nz=p.nz,
            oversampling_factor=p.oversampling_factor, oversample_y=int(p.oversample_y),
            lx=p.lx, ly=p.ly, lz=p.lz, re=p.re, dt=p.dt, implicitness=p.implicitness,
            tilt_degree=p.tilt_degree, fd_order=p.fd_order,
            precision=_lib.DNSMI_MIXED if p.precision == "mixed" else _lib.DNSMI_FP64,
            correct_divergence=int(p.correct_divergence), tile_cols=p.tile_cols,
            variants=p.variants, plain_mixed=int(p.plain_mixed),
            recenter_far=p.recenter_far or 0.0, recenter_near=p.recenter_near,
            setup_on_host=int(p.setup_on_host),
        )
        handle = ctypes.c_void_p()
        torch.cuda.current_device()  # make the device current before any HIP call
        _lib.check(self.lib.dnsmi_create(ctypes.byref(cfg), ctypes.byref(handle)), None)
        self.handle = handle
        self.shape = (3,) + p.spec_shape
        self.device = torch.device("cuda", torch.cuda.current_device())

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                torch.cuda.synchronize()
            except Exception:  # interpreter teardown: the device is being released anyway
                return
            self.lib.dnsmi_destroy(h)
            self.handle = None

    # ---------------------------------------------------------------- helpers
    def _chk(self, rc: int) -> None:
        _lib.check(rc, self.handle)

    def check_field(self, t: torch.Tensor, name: str) -> None:
        if t.dtype != torch.complex128 or tuple(t.shape) != self.shape or not t.is_contiguous():
            raise ValueError(f"{name}: need contiguous complex128 {self.shape}, got {t.dtype} {tuple(t.shape)}")
        if t.device != self.device:
            raise ValueError(f"{name}: tensor on {t.device}, context on {self.device}")

    def empty(self) -> torch.Tensor:
        return torch.empty(self.shape, dtype=torch.complex128, device=self.device)

    def scalar(self) -> torch.Tensor:
        return torch.empty((), dtype=torch.float64, device=self.device)

    # ---------------------------------------------------------------- calls
    def get_rhs(self, state, out=None):
        self.check_field(state, "state")
        out = self.empty() if out is None else out
        self._chk(self.lib.dnsmi_get_rhs(self.handle, _ptr(state), _ptr(out), _stream()))
        return out

    def predict(self, state, rhs, out=None):
        self.check_field(state, "state")
        self.check_field(rhs, "rhs")
        out = self.empty() if out is None else out
        self._chk(self.lib.dnsmi_predict(self.handle, _ptr(state), _ptr(rhs), _ptr(out), _stream()))
        return out

    def correct(self, state_prev, pred, rhs_prev, rhs_next, out=None, correction=True):
        for t, n in ((state_prev, "state_prev"), (pred, "prediction"), (rhs_prev, "rhs_prev"), (rhs_next, "rhs_next")):
            self.check_field(t, n)
        out = self.empty() if out is None else out
        corr = self.empty() if correction else None
        self._chk(self.lib.dnsmi_correct(self.handle, _ptr(state_prev), _ptr(pred), _ptr(rhs_prev),
                                         _ptr(rhs_next), _ptr(out), _ptr(corr), _stream()))
        return out, corr

    def norm(self, correction, out=None):
        self.check_field(correction, "correction")
        out = self.scalar() if out is None else out
        self._chk(self.lib.dnsmi_norm(self.handle, _ptr(correction), _ptr(out), _stream()))
        return out

    def predict_and_correct(self, state, pred=None, rhs_next=None, err=None):
        self.check_field(state, "state")
        pred = self.empty() if pred is None else pred
        rhs_next = self.empty() if rhs_next is None else rhs_next
        err = self.scalar() if err is None else err
        self._chk(self.lib.dnsmi_predict_and_correct(self.handle, _ptr(state), _ptr(pred),
                                                     _ptr(rhs_next), _ptr(err), _stream()))
        return pred, rhs_next, err

    def iterate_correction(self, state_prev, pred, rhs_prev, rhs_next=None, err=None):
        self.check_field(state_prev, "state_prev")
        self.check_field(pred, "prediction")
        self.check_field(rhs_prev, "rhs_prev")
        rhs_next = self.empty() if rhs_next is None else rhs_next
        err = self.scalar() if err is None else err
        self._chk(self.lib.dnsmi_iterate_correction(self.handle, _ptr(state_prev), _ptr(pred),
                                                    _ptr(rhs_prev), _ptr(rhs_next), _ptr(err), _stream()))
        return pred, rhs_next, err

    def get_stats(self, state, out=None):
        """Device tensor (4,) float64: E, I, D, E'."""
        self.check_field(state, "state")
        out = torch.empty(4, dtype=torch.float64, device=self.device) if out is None else out
        self._chk(self.lib.dnsmi_get_stats(self.handle, _ptr(state), _ptr(out), _stream()))
        return out

    # ---------------------------------------------------------------- device-side steps
    def _esc_args(self, esc_slots):
        """(esc_slots, ratio, cap) of the mixed form's FP64 escalation slots."""
        p = self.params
        if p.precision != "mixed" or p.escalate_ratio is None:
            return 0, 0.0, 0.0
        # two: the first FP64 pass still differences against the last FP32
        # evaluation, so its error sits at the FP32 resolution
        n = 2 if esc_slots is None else esc_slots
        return n, p.escalate_ratio, p.escalate_cap_value

    def step_device(self, state, nxt, rhs_a, rhs_b, slots, counts=None, tol=None, max_iter=None,
                    esc_slots=None):
        """Enqueue one full time step with the corrector decided on the device
        (capturable; no host synchronisation).  The new state goes to `nxt`.
        Mixed form: `esc_slots` FP64 slots (default 2) follow the `slots`
        passes, for a step whose FP32 passes stop converging."""
        p = self.params
        tol = p.corrector_tolerance if tol is None else tol
        max_iter = p.max_corrector_iterations if max_iter is None else max_iter
        n_esc, ratio, cap = self._esc_args(esc_slots)
        ncap = 0 if counts is None else counts.numel()
        self._chk(self.lib.dnsmi_step_device(self.handle, _ptr(state), _ptr(nxt), _ptr(rhs_a), _ptr(rhs_b),
                                             slots, n_esc, _ptr(counts), ncap, tol, max_iter, ratio, cap,
                                             _stream()))

    def escalate(self, on: bool) -> None:
        """Mixed form: run the following corrector passes in FP64 (or back)."""
        self._chk(self.lib.dnsmi_escalate(self.handle, int(bool(on))))

    def mlp_features(self, n_m2, n_m1, n_cur, n_next, modes, s_1, s_2, s_out, X, T=None, n_m3=None):
        """Training features (and targets when n_next and T are given: the
        residual of quadratic extrapolation, of cubic with n_m3)."""
        self._chk(self.lib.dnsmi_mlp_features(self.handle, _ptr(n_m3), _ptr(n_m2), _ptr(n_m1), _ptr(n_cur),
                                              _ptr(n_next),
                                              _ptr(modes), modes.numel(), s_1, s_2, s_out, _ptr(X), _ptr(T),
                                              _stream()))

    def guess(self, state, pred, rhs_prev, hist1, hist2, mode, net=None, scales=(1.0, 1.0, 1.0), hist3=None,
              carry=False):
        """Predictor with a corrector guess (mode 0 plain, 1 linear, 3
        quadratic, 4 cubic extrapolation, 2 quadratic -- cubic with hist3 --
        plus the network `net` = (H, packed weights, bias)).  N^n overwrites
        the oldest history: hist2 (swap hist1, hist2 next) or hist3 (rotate
        (hist1, hist2, hist3) <- (hist3, hist1, hist2) next).
        carry: rhs_prev receives the carry form's B (periodic)."""
        H, wb, bias = (32, None, None) if net is None else net
        s_1, s_2, s_out = scales
        self._chk(self.lib.dnsmi_guess(self.handle, _ptr(state), _ptr(pred), _ptr(rhs_prev), _ptr(hist1),
                                       _ptr(hist2), _ptr(hist3), _ptr(wb), _ptr(bias), H, s_1, s_2, s_out, mode,
                                       int(bool(carry)), _stream()))

    def predict_and_correct_carry(self, state, pred, carry, err):
        """Carry form (periodic): pred <- the first corrector iterate,
        carry <- B = ildt_2 (ldt_1 u^n + (1 - c) N^n)."""
        self.check_field(state, "state")
        self._chk(self.lib.dnsmi_predict_and_correct_carry(self.handle, _ptr(state), _ptr(pred), _ptr(carry),
                                                           _ptr(err), _stream()))

    def iterate_correction_carry(self, pred, carry, err):
        """Carry form (periodic): pred <- B + c ildt_2 N(pred) in place."""
        self._chk(self.lib.dnsmi_iterate_correction_carry(self.handle, _ptr(pred), _ptr(carry), _ptr(err),
                                                          _stream()))

    def step_device_guess(self, state, nxt, rhs_a, rhs_b, hist1, hist2, mode, slots, counts=None, net=None,
                          scales=(1.0, 1.0, 1.0), esc_slots=None, hist3=None):
        """Enqueue one time step started from the corrector guess, decided on
        the device (capturable).  Up to max_corrector_iterations + 1 passes;
        mixed form: `esc_slots` FP64 slots (default 2) as in step_device."""
        p = self.params
        H, wb, bias = (32, None, None) if net is None else net
        s_1, s_2, s_out = scales
        n_esc, ratio, cap = self._esc_args(esc_slots)
        ncap = 0 if counts is None else counts.numel()
        self._chk(self.lib.dnsmi_step_device_guess(
            self.handle, _ptr(state), _ptr(nxt), _ptr(rhs_a), _ptr(rhs_b), _ptr(hist1), _ptr(hist2), _ptr(hist3),
            _ptr(wb),
            _ptr(bias), H, s_1, s_2, s_out, mode, slots, n_esc, _ptr(counts), ncap, p.corrector_tolerance,
            p.max_corrector_iterations + 1, ratio, cap, _stream()))

    def step_stats(self) -> dict:
        out = (ctypes.c_double * 7)()
        self._chk(self.lib.dnsmi_step_stats(self.handle, out))
        return {"steps": int(out[0]), "total_iterations": int(out[1]), "unconverged": int(out[2]),
                "nan": int(out[3]), "last_error": out[4], "escalated": int(out[5]),
                "rerecorded": int(out[6])}

    def split_rerecord(self) -> None:
        """Split evaluations: the next corrector evaluation records a new reference."""
        self._chk(self.lib.dnsmi_split_rerecord(self.handle))

    def step_reset(self) -> None:
        self._chk(self.lib.dnsmi_step_reset(self.handle))

    @property
    def tuning(self) -> str:
        return self.lib.dnsmi_tuning(self.handle).decode()

    def variants(self) -> tuple[int, int]:
        """(tile_cols, variants) the context runs, in Params' convention."""
        tc, v = ctypes.c_int(), ctypes.c_int()
        self._chk(self.lib.dnsmi_variants(self.handle, ctypes.byref(tc), ctypes.byref(v)))
        return tc.value, v.value

    def walls_table(self, which: int, shape: tuple):
        """Host copy of a wall-bounded set-up table (see include/dnsmi.h)."""
        import numpy as np

        out = np.zeros(shape, dtype=np.float64)
        self._chk(self.lib.dnsmi_walls_table(self.handle, which, out.ctypes.data_as(ctypes.c_void_p),
                                             out.size))
        return out

    # ---------------------------------------------------------------- timing
    def profile_begin(self) -> None:
        self._chk(self.lib.dnsmi_profile_begin(self.handle))

    def profile_end(self) -> dict:
        """{kernel name: (total ms, launches)} of the launches since begin."""
        n = _lib.KERNEL_IDS
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_longlong * n)()
        self._chk(self.lib.dnsmi_profile_end(self.handle, n, ms, cnt))
        return {self.lib.dnsmi_kernel_name(i).decode(): (ms[i], cnt[i]) for i in range(n) if cnt[i]}

    def correct_velocity(self, state, measure=False):
        self.check_field(state, "state")
        err = self.scalar() if measure else None
        self._chk(self.lib.dnsmi_correct_velocity(self.handle, _ptr(state), _ptr(err), _stream()))
        return state, err


class BatchPlan:
    """One device step of several periodic contexts of one size class, merged
    into one launch per kernel over the members (``dnsmi_batch_plan``).  The
    members' buffers are fixed at planning; ``enqueue`` issues the step of
    all of them on torch's current stream (capturable)."""

    def __init__(self, natives, states, nexts, rhs_a, rhs_b, slots, counts=None, esc_slots=None):
        n = len(natives)
        if not n or not (len(states) == len(nexts) == len(rhs_a) == len(rhs_b) == n):
            raise ValueError("need one state, next, rhs_a and rhs_b per member")
        lib = natives[0].lib
        p = natives[0].params
        tol = p.corrector_tolerance
        max_iter = p.max_corrector_iterations
        n_esc, ratio, cap = natives[0]._esc_args(esc_slots)
        for nat, *bufs in zip(natives, states, nexts, rhs_a, rhs_b):
            for t, name in zip(bufs, ("state", "next", "rhs_a", "rhs_b")):
                nat.check_field(t, name)
            q = nat.params
            if (q.corrector_tolerance, q.max_corrector_iterations) != (tol, max_iter):
                raise ValueError("members of a batch share the corrector tolerance and iteration limit")
        P = ctypes.c_void_p
        arr = lambda ts: (P * n)(*[t.data_ptr() for t in ts])  # noqa: E731
        ctxs = (P * n)(*[nat.handle.value for nat in natives])
        cnt = None if counts is None else (P * n)(*[c.data_ptr() for c in counts])
        ncap = 0 if counts is None else min(c.numel() for c in counts)
        handle = ctypes.c_void_p()
        rc = lib.dnsmi_batch_plan(ctxs, n, arr(states), arr(nexts), arr(rhs_a), arr(rhs_b), slots, n_esc, cnt, ncap,
                                  tol, max_iter, ratio, cap, ctypes.byref(handle))
        _lib.check(rc, natives[0].handle)
        self.lib, self.handle, self.members = lib, handle, n
        self._keep = (list(states), list(nexts), list(rhs_a), list(rhs_b), counts)

    @property
    def launches(self) -> int:
        return self.lib.dnsmi_batch_launches(self.handle)

    def enqueue(self) -> None:
        _lib.check(self.lib.dnsmi_batch_enqueue(self.handle, _stream()), None)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                torch.cuda.synchronize()
            except Exception:
                return
            self.lib.dnsmi_batch_destroy(h)
            self.handle = None
