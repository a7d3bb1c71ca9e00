"""Snapshots on the GPU: a reference-format ``.npz`` written from a golden
state loads through the device transforms onto that state, and the native
stepper started from it reproduces the golden RHS and predictor step.

Bar: the loaded state within 1e-14 of the golden state (max abs), the
saved total velocity within 1e-13 of the oracle's, and the step from the
loaded state to the same relative L2 bars as the golden-vector tests
(1e-12 periodic, 1e-11 plane Couette predictor).
"""

import math
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dnsjax_amd.params import Params  # noqa: E402
from oracle import periodic as OP  # noqa: E402
from oracle import walls as OW  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def rel(a, b):
    a = a.detach().cpu().numpy()
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_periodic_snapshot_drives_the_native_step(tmp_path):
    from dnsjax_amd.flows import TriplyPeriodicFlow

    L = 2 * math.pi
    g = np.load(os.path.join(GOLDEN, "periodic_16_kolmogorov.npz"))
    p = Params(system="kolmogorov", nx=16, ny=16, nz=16, lx=L, ly=L, lz=L, re=40.0, dt=0.01)
    op = OP.PeriodicParams(nx=16, ny=16, nz=16, lx=L, ly=L, lz=L, re=40.0, dt=0.01)
    flow = TriplyPeriodicFlow(p)
    path = str(tmp_path / "snap.npz")
    flow.save_snapshot(path, flow.to_device(g["state"]))
    with np.load(path) as f:
        total = f["velocity_phys"]
    want = OP.spec_to_phys(g["state"], op) + OP.Flow(op).base_flow
    assert np.abs(total - want).max() < 1e-13
    u = flow.init_state(snapshot=path)
    assert u.is_cuda and u.dtype == torch.complex128
    assert np.abs(u.cpu().numpy() - g["state"]).max() < 1e-14
    assert rel(flow.get_rhs(u), g["rhs0"]) < 1e-12
    pred, rhs1, err1 = flow.predict_and_correct(u)
    assert rel(pred, g["pred1"]) < 1e-12
    assert abs(float(err1) - float(g["err1"])) <= 1e-9 * float(g["err1"])


def test_walls_snapshot_drives_the_native_step(tmp_path):
    from dnsjax_amd.flows import PlaneCouetteFlow

    g = np.load(os.path.join(GOLDEN, "walls_8x17x8_p4.npz"))
    p = Params(system="plane-couette", nx=8, ny=17, nz=8, lx=2 * math.pi, lz=math.pi,
               fd_order=4, re=400.0, dt=0.01)
    op = OW.WallParams(nx=8, ny=17, nz=8, lx=2 * math.pi, lz=math.pi)
    flow = PlaneCouetteFlow(p)
    path = str(tmp_path / "walls.npz")
    flow.save_snapshot(path, flow.to_device(g["state"]))
    with np.load(path) as f:
        phys = f["velocity_phys_nonexpanded"]
    assert np.abs(phys - OW.spec_to_phys_2d(g["state"], op)).max() < 1e-13
    u = flow.init_state(snapshot=path)
    assert u.is_cuda and np.abs(u.cpu().numpy() - g["state"]).max() < 1e-14
    assert rel(flow.get_rhs(u), g["rhs0"]) < 1e-12
    pred, rhs1, err1 = flow.predict_and_correct(u)
    assert rel(pred, g["pred1"]) < 1e-11
