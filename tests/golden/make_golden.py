"""Write the committed golden fixtures from the FP64 oracle.

    python tests/golden/make_golden.py

Each .npz holds the inputs (seeded state) and the oracle's outputs of one
predict-and-correct step, one extra corrector iteration, and
correct_velocity, for a small configuration.
"""

import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from dnsjax_amd.fields import periodic_field, wall_field  # noqa: E402
from oracle import periodic as P  # noqa: E402
from oracle import stepper as S  # noqa: E402
from oracle import walls as W  # noqa: E402

PERIODIC_CASES = {
    # name: (nx, ny, nz, L, system, re, dt, rms, seed, tilt)
    "periodic_16_kolmogorov": (16, 16, 16, 2 * math.pi, "kolmogorov", 40.0, 0.01, 0.05, 0, 0.0),
    "periodic_16x8x32_waleffe_ly4": (16, 8, 32, 4.0, "waleffe", 60.0, 0.02, 0.2, 1, 0.0),
    "periodic_8_kolmogorov_tilt30": (8, 8, 8, 4.0, "kolmogorov", 30.0, 0.02, 0.3, 2, 30.0),
}

WALL_CASES = {
    # name: (nx, ny, nz, fd_order, re, dt, rms, seed)
    "walls_8x17x8_p4": (8, 17, 8, 4, 400.0, 0.01, 0.05, 2),
    "walls_16x33x8_p6": (16, 33, 8, 6, 400.0, 0.01, 0.1, 3),
}


def periodic_case(name):
    nx, ny, nz, L, system, re, dt, rms, seed, tilt = PERIODIC_CASES[name]
    p = P.PeriodicParams(nx=nx, ny=ny, nz=nz, lx=L, ly=L, lz=L, re=re, system=system, dt=dt,
                         tilt_degree=tilt)
    f = P.Flow(p)
    u = periodic_field(nx, ny, nz, L, L, L, rms, seed)
    pc, ic = S.make_stepper(P.get_rhs, P.predict, P.correct, P.norm)
    rhs0 = P.get_rhs(u, f)
    pred, rhs1, err1 = pc(u, f)
    pred2, rhs2, err2 = ic(u, pred.copy(), rhs1, f)
    cv = P.correct_velocity(pred2, f)
    return dict(state=u, rhs0=rhs0, pred1=pred, rhs1=rhs1, err1=err1, pred2=pred2, rhs2=rhs2,
                err2=err2, corrected=cv)


def wall_case(name):
    nx, ny, nz, fd, re, dt, rms, seed = WALL_CASES[name]
    p = W.WallParams(nx=nx, ny=ny, nz=nz, fd_order=fd, re=re, dt=dt)
    f = W.WallFlow(p)
    u = wall_field(nx, ny, nz, p.lx, p.lz, rms, seed)
    pc, ic = S.make_stepper(W.get_rhs, W.predict, W.correct, W.norm)
    rhs0 = W.get_rhs(u, f)
    pred, rhs1, err1 = pc(u, f)
    pred2, rhs2, err2 = ic(u, pred.copy(), rhs1, f)
    return dict(state=u, rhs0=rhs0, pred1=pred, rhs1=rhs1, err1=err1, pred2=pred2, rhs2=rhs2,
                err2=err2, **{k: f.imm[k] for k in ("Lk_ab", "Lk_piv", "Hk_ab", "Hk_piv", "M_inv", "v1", "q1")})


if __name__ == "__main__":
    for name in PERIODIC_CASES:
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **periodic_case(name))
        print("wrote", name)
    for name in WALL_CASES:
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **wall_case(name))
        print("wrote", name)
