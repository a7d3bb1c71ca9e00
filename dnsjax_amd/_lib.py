"""ctypes binding of libdnsmi.so (the C ABI declared in include/dnsmi.h).

The library is built in-tree by ``python -m dnsjax_amd.build`` (or
``__graft_entry__.build()``).  Loading fails loudly when it is missing:
there is no CPU fallback for any entry point.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("DNSMI_LIB") or os.path.join(HERE, "libdnsmi.so")

DNSMI_KOLMOGOROV = 0
DNSMI_WALEFFE = 1
DNSMI_DECAYING_BOX = 2
DNSMI_PLANE_COUETTE = 3
DNSMI_FP64 = 0
DNSMI_MIXED = 1

SYSTEMS = {
    "kolmogorov": DNSMI_KOLMOGOROV,
    "waleffe": DNSMI_WALEFFE,
    "decaying-box": DNSMI_DECAYING_BOX,
    "plane-couette": DNSMI_PLANE_COUETTE,
}

ERRORS = {1: "EINVAL", 2: "EHIP", 3: "ENOMEM", 4: "ESINGULAR"}


class DnsmiConfig(ctypes.Structure):
    _fields_ = [
        ("system", ctypes.c_int),
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("nz", ctypes.c_int),
        ("oversampling_factor", ctypes.c_int),
        ("oversample_y", ctypes.c_int),
        ("lx", ctypes.c_double),
        ("ly", ctypes.c_double),
        ("lz", ctypes.c_double),
        ("re", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("implicitness", ctypes.c_double),
        ("tilt_degree", ctypes.c_double),
        ("fd_order", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("correct_divergence", ctypes.c_int),
        ("tile_cols", ctypes.c_int),
        ("variants", ctypes.c_int),
        ("plain_mixed", ctypes.c_int),
        ("recenter_far", ctypes.c_double),
        ("recenter_near", ctypes.c_double),
        ("setup_on_host", ctypes.c_int),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int

# name -> (restype, argtypes); the exported surface of include/dnsmi.h
SIGNATURES = {
    "dnsmi_create": (_I, [ctypes.POINTER(DnsmiConfig), ctypes.POINTER(_P)]),
    "dnsmi_destroy": (None, [_P]),
    "dnsmi_last_error": (ctypes.c_char_p, [_P]),
    "dnsmi_spec_elems": (ctypes.c_longlong, [_P]),
    "dnsmi_get_rhs": (_I, [_P, _P, _P, _P]),
    "dnsmi_predict": (_I, [_P, _P, _P, _P, _P]),
    "dnsmi_correct": (_I, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "dnsmi_norm": (_I, [_P, _P, _P, _P]),
    "dnsmi_predict_and_correct": (_I, [_P, _P, _P, _P, _P, _P]),
    "dnsmi_iterate_correction": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "dnsmi_correct_velocity": (_I, [_P, _P, _P, _P]),
    "dnsmi_profile_begin": (_I, [_P]),
    "dnsmi_profile_end": (_I, [_P, _I, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_longlong)]),
    "dnsmi_kernel_name": (ctypes.c_char_p, [_I]),
    "dnsmi_walls_table": (_I, [_P, _I, _P, ctypes.c_longlong]),
    "dnsmi_tuning": (ctypes.c_char_p, [_P]),
    "dnsmi_get_stats": (_I, [_P, _P, _P, _P]),
    "dnsmi_step_device": (_I, [_P, _P, _P, _P, _P, _I, _I, _P, _I, ctypes.c_double, _I, ctypes.c_double,
                               ctypes.c_double, _P]),
    "dnsmi_step_stats": (_I, [_P, ctypes.POINTER(ctypes.c_double)]),
    "dnsmi_split_rerecord": (_I, [_P]),
    "dnsmi_step_reset": (_I, [_P]),
    "dnsmi_escalate": (_I, [_P, _I]),
    "dnsmi_mlp_sizes": (_I, [_I, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                             ctypes.POINTER(ctypes.c_longlong)]),
    "dnsmi_mlp_grad": (_I, [_I, _I, _P, _P, _P, _I, _P, _P, _P, _P]),
    "dnsmi_mlp_grad_static": (_I, [_I, _I, _P, _P, _P, _I, _P, _P, _P]),
    "dnsmi_mlp_adam": (_I, [_I, _P, _P, _P, _P, _I, ctypes.c_float, _P]),
    "dnsmi_pnet_params": (ctypes.c_longlong, [_I, _I, _I]),
    "dnsmi_pnet_grad": (_I, [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "dnsmi_pnet_forward": (_I, [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P]),
    "dnsmi_pnet_adam": (_I, [ctypes.c_longlong, _P, _P, _P, _P, _I, ctypes.c_float, _P]),
    "dnsmi_pnet_features": (_I, [_P, _P, _P, _P, _P, _P, _I, _I, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                 _P, _P, _P]),
    "dnsmi_guess_profile": (_I, [_P, _P, _P, _P, _P, _P, _I, _I, _I, _P, _P, _P, _P, _P, _P, ctypes.c_float,
                                 ctypes.c_float, ctypes.c_float, _P]),
    "dnsmi_mlp_pack": (_I, [_I, _P, _P, _P, _P]),
    "dnsmi_mlp_forward": (_I, [_I, _P, _I, _P, _P, _P, _P]),
    "dnsmi_mlp_features": (_I, [_P, _P, _P, _P, _P, _P, _P, _I, ctypes.c_float, ctypes.c_float,
                                ctypes.c_float, _P, _P, _P]),
    "dnsmi_guess": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I, ctypes.c_float, ctypes.c_float,
                         ctypes.c_float, _I, _I, _P]),
    "dnsmi_predict_and_correct_carry": (_I, [_P, _P, _P, _P, _P, _P]),
    "dnsmi_iterate_correction_carry": (_I, [_P, _P, _P, _P, _P]),
    "dnsmi_step_device_guess": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_float, _I, _I, _I, _P, _I, ctypes.c_double, _I, ctypes.c_double,
                                     ctypes.c_double, _P]),
    "dnsmi_variants": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "dnsmi_batch_plan": (_I, [_P, _I, _P, _P, _P, _P, _I, _I, _P, _I, ctypes.c_double, _I, ctypes.c_double,
                              ctypes.c_double, ctypes.POINTER(_P)]),
    "dnsmi_batch_enqueue": (_I, [_P, _P]),
    "dnsmi_batch_launches": (_I, [_P]),
    "dnsmi_batch_destroy": (None, [_P]),
}

KERNEL_IDS = 12

_lib = None


def load() -> ctypes.CDLL:
    """Load the in-tree library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m dnsjax_amd.build` "
            "(there is no CPU fallback for the HIP path)"
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class DnsmiError(RuntimeError):
    pass


def check(rc: int, ctx=None) -> None:
    if rc != 0:
        lib = load()
        msg = lib.dnsmi_last_error(ctx)
        raise DnsmiError(f"dnsmi error {ERRORS.get(rc, rc)}: {msg.decode() if msg else ''}")
