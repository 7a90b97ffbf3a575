"""VUMPS on the GPU against the oracle VUMPS (golden energy) and the Bethe-ansatz energy.

Workload: the antiferromagnetic Heisenberg chain (XXZ, Delta = -1 in the reference's convention,
J = 1, hz = 0), two-site cell, QN = [0], D = 64 in 9 U(1) sectors, seed 1001 — the two-site XXZ
VUMPS of example_config_XXZ.cfg at D = 64.  The fixture tests/golden/vumps_xxz_d64.json holds the
oracle VUMPS energy at gradient norm F <= 1e-10.  At fixed bond structure VUMPS converges to one
fixed point, so the GPU iteration (FP64, and BF16X3 operators under FP64 refinement) must reach the
same energy; the energy error of a state with gradient norm F is O(F^2), so at F <= 1e-9 the two
agree far below the 1e-12 tolerance.
"""
import json
import os

import numpy as np
import pytest

import imps_cpp_tools_amd as im

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def fixed_schedules(monkeypatch):
    """The panel-kernel autotuner chooses schedules by timing, and schedules sum in different orders;
    the early sweeps' solves stop at tolerances of ~1e-6, so such rounding differences change the
    trajectory (and the sweep count to F <= 1e-9, 100-400 sweeps).  The fixed schedule rule makes
    every run follow the same one."""
    monkeypatch.setenv("IMPS_AUTOTUNE", "0")

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "vumps_xxz_d64.json")) as f:
    GOLD = json.load(f)


def _run(prec, f_target=1e-9, max_sweeps=600):
    from imps_cpp_tools_amd.vumps import VUMPS
    ctx = im.Context(0)
    v = VUMPS(ctx, im.XXZ, GOLD["N"], GOLD["QN"], np.array(GOLD["keys0"], np.int32), np.array(GOLD["sizes0"], np.int32),
              GOLD["params"], prec=prec, seed=GOLD["seed"])
    hist = []
    for _ in range(max_sweeps):
        st = v.sweep()
        hist.append((st.energy, max(st.F)))
        if max(st.F) <= f_target:
            break
    return hist


@pytest.mark.parametrize("prec", [im.FP64, im.BF16X3], ids=["fp64", "bf16x3"])
def test_vumps_xxz_d64_energy(prec):
    hist = _run(prec)
    e, F = hist[-1]
    print("sweeps", len(hist), "energy %.15f" % e, "F %.3e" % F, flush=True)
    assert F <= 1e-9, hist[-5:]
    assert abs(e - GOLD["energy"]) <= 1e-12, (e, GOLD["energy"])
    # variational: never below the exact (Bethe ansatz) ground-state energy per site
    assert e >= GOLD["bethe_ansatz"]
    # energies decrease monotonically once the iteration is in the quadratic regime
    tail = [h[0] for h in hist[len(hist) // 2:]]
    assert max(np.diff(tail)) <= 1e-12


def test_random_state_matches_oracle():
    """randMPS_LR on the GPU (QR, TM fixed point, triangular factor, right QRs) against the oracle's
    construction from the same seeded Gaussian tensors."""
    import torch
    from imps_cpp_tools_amd.vumps import VUMPS
    from oracle_vumps import OracleVUMPS
    args = (GOLD["N"], GOLD["QN"], np.array(GOLD["keys0"], np.int32), np.array(GOLD["sizes0"], np.int32), GOLD["params"])
    v = VUMPS(im.Context(0), im.XXZ, *args, seed=GOLD["seed"])
    o = OracleVUMPS(0, *args, seed=GOLD["seed"])
    for n in range(GOLD["N"]):
        for got, ref in ((v.AL[n], o.AL[n]), (v.C[n], o.C[n]), (v.AR[n], o.AR[n])):
            g = got.cpu().numpy()
            assert g.shape == ref.data.shape
            assert np.abs(g - ref.data).max() <= 1e-12
    torch.cuda.synchronize()
