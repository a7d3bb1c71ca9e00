"""Every transform length the passes support, against the FP64 oracle.

The dealiased periodic RHS at cubes n = 8 .. 128 runs the padded lengths
12, 24, 48, 72, 96, 144 and 192 through every pass (y, z, x and back); the
unpadded y axis (oversample_y = False) at ny = 16 .. 128 runs 16, 32, 64 and
128.  Bar: relative L2 <= 1e-12 FP64, <= 1e-5 for the mixed form (FP32
transforms), against ``oracle/periodic.py`` get_rhs on the same seeded field.
"""

import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dnsjax_amd.params import Params  # noqa: E402
from oracle import periodic as P  # noqa: E402

L = 2 * math.pi


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def rel(a, b):
    a = a.detach().cpu().numpy()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def check(nx, ny, nz, oversample_y, precision):
    from dnsjax_amd.fields import periodic_field
    from dnsjax_amd.flows import TriplyPeriodicFlow

    kw = dict(nx=nx, ny=ny, nz=nz, lx=L, ly=L, lz=L, re=40.0, dt=0.01, system="kolmogorov",
              oversample_y=oversample_y)
    flow = TriplyPeriodicFlow(Params(precision=precision, **kw))
    of = P.Flow(P.PeriodicParams(**kw))
    u = periodic_field(nx, ny, nz, L, L, L, 0.1, 5)
    print(f"lengths: x {of.p.nx_padded}, y {of.p.ny_padded}, z {of.p.nz_padded}, {precision}", flush=True)
    err = rel(flow.get_rhs(flow.to_device(u)), P.get_rhs(u, of))
    assert err < (1e-12 if precision == "fp64" else 1e-5), err


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n", [8, 16, 32, 48, 64, 96, 128])
def test_padded_lengths(n, precision):
    check(n, n, n, True, precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("ny", [16, 32, 64, 128])
def test_unpadded_y_lengths(ny, precision):
    check(8, ny, 8, False, precision)
