"""The fixed algorithmic byte models bench.py divides by kernel time to
report roofline fractions, and its one-GPU contract.

Pinned by the model figures the measurement plan states for the workloads:
the periodic RHS pass model is 1228.0 MB at 128^3 FP64 and 18.95 MB at 32^3;
the plane Couette RHS model is 138.0 MB at 64x128x64 and 1105.7 MB at
128x256x128; one influence-matrix step is 126.0 MB and 1215.4 MB.  Each is
also restated here from its parts (field sizes, band factors, pivots), so a
dropped term or a wrong multiplier fails.
"""

from __future__ import annotations

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def mb(x):
    return round(x / 1e6, 2)


def test_periodic_rhs_model_matches_the_stated_figures():
    assert mb(bench.rhs_model_bytes(128, 16)) == pytest.approx(1228.02, abs=0.01)
    assert mb(bench.rhs_model_bytes(32, 16)) == pytest.approx(18.95, abs=0.01)
    # mixed: FP32 intermediates, the FP64 state read and RHS written
    assert mb(bench.rhs_model_bytes(128, 8)) == pytest.approx(663.56, abs=0.01)
    S = 3 * 127 * 127 * 64 * 16
    assert bench.rhs_model_bytes(128, 16) - bench.rhs_model_bytes(128, 8) == \
        (bench.rhs_model_bytes(128, 16) - 2 * S) // 2


def test_periodic_kernel_bytes_parts():
    kb = bench.kernel_bytes(128, 16)
    S = 3 * 127 * 127 * 64 * 16
    assert kb["S"] == S == 49548288
    plane = 192 * 127 * 64 * 16  # one padded-y, stored-z field
    cube = 192 * 192 * 64 * 16   # one padded-y, padded-z field
    assert kb["y_inverse"] == S + 5 * plane
    assert kb["z_inverse"] == 5 * plane + 6 * cube
    assert kb["x_nonlinear"] == 9 * cube
    assert kb["z_forward"] == 3 * cube + 3 * plane
    # the epilogues: the RHS writes N (S); the carry predictor reads u^n and
    # writes u^1 and B; a carry corrector pass reads B and the iterate and
    # writes the iterate; the incremental one also reads N_prev and writes N;
    # the split recording pass also writes B + c ildt_2 N(u_ref), which each
    # increment reads in place of B
    assert kb["y_forward:predict"] - kb["y_forward:rhs"] == 2 * S
    assert kb["y_forward:correct"] - kb["y_forward:rhs"] == 2 * S
    assert kb["y_forward:correct_incremental"] - kb["y_forward:rhs"] == 3 * S
    kb8 = bench.kernel_bytes(128, 8)
    assert kb8["y_forward:correct_record"] - kb8["y_forward:correct_split"] == S


@pytest.mark.parametrize("shape,rhs_mb,imm_mb", [((64, 128, 64, 4), 138.02, 125.99),
                                                  ((128, 256, 128, 6), 1105.72, 1215.43)])
def test_walls_models_match_the_stated_figures(shape, rhs_mb, imm_mb):
    m = bench.walls_model_bytes(*shape)
    assert mb(m["rhs"]) == pytest.approx(rhs_mb, abs=0.01)
    assert mb(m["imm"]) == pytest.approx(imm_mb, abs=0.01)


def test_walls_imm_model_parts():
    nx, ny, nz, p = 64, 128, 64, 4
    modes = (nz - 1) * (nx // 2)
    spec = 3 * modes * ny * 16
    ab = modes * (3 * p + 1) * ny * 8
    piv = modes * ny * 4
    vec = modes * ny * 8
    minv = modes * 4 * 8
    assert bench.walls_model_bytes(nx, ny, nz, p)["imm"] == 5 * spec + 2 * ab + 2 * piv + 4 * vec + minv
    assert mb(ab) == pytest.approx(26.84, abs=0.01)  # (63, 32, 13, 128) f64


def test_walls_roofline_from_event_times():
    prof = {"walls_z_inverse": (0.2, 10), "walls_x_nonlinear": (0.2, 10), "walls_z_forward": (0.1, 10),
            "walls_imm": (1.0, 10)}
    r = bench.walls_roofline(prof, 64, 128, 64, 4)
    assert r["rhs"]["t_us"] == pytest.approx(50.0)
    assert r["imm"]["t_us"] == pytest.approx(100.0)
    assert r["rhs"]["achieved"] == pytest.approx(138018816 / 50e-6 / 1e9, rel=1e-3)
    assert r["imm"]["frac"] == pytest.approx(125991936 / 100e-6 / 8e12, abs=1e-4)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stdout
