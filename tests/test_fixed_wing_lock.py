"""The oracle against the reference's locked production numbers: the fixed
AR = 2 wing diagnostic (MetalExecutionTests.swift:2204-2224,
productionMetalFixedWingDiagnosticMatchesLockedBaseline): 1,300 steps of TRT
with halfway bounce-back, far-field faces, sponge and force mode 6 on
80 x 80 x 48 cells; the reference locks CL = 0.5426947226 and
CD = 0.6728693016 to 1e-6 on its own hardware.

The FP32 oracle reproduces both to 1e-5 (measured 8.3e-6 and 5.9e-6: FP32
rounding of a different machine and compiler over 1,300 steps; the
reference sums the loads in FP32, the oracle in FP64).  A plausible slip
(a weight, a sign of the moving-wall term, the sponge profile, the
far-field face, the estimator) moves CL or CD by far more.  Side force and
roll / yaw moments stay below the reference's 1e-3 bars.
"""
import math

import numpy as np

import bfref
import cases

BAR = 2e-5


def test_fixed_wing_matches_reference_locked_coefficients():
    case = cases.FixedWing()
    p = case.params
    assert (p.nx, p.ny, p.nz) == (80, 80, 48) and case.steps == 1300
    assert int(case.mask.sum()) == 8 * 16       # chord 8 x span 16 cells, one cell thick
    f = bfref.init_equilibrium(p, case.mask, case.wall, np.float32)
    for _ in range(case.steps):
        f, load, _, _ = bfref.step(p, f, case.mask, case.mask, case.wall)
    lift, drag = case.coefficients(load)
    locked = cases.FIXED_WING_LOCKED
    print(f"CL {lift:.10f} (locked {locked['lift']:.10f}), CD {drag:.10f} (locked {locked['drag']:.10f})")
    assert abs(lift - locked["lift"]) <= BAR
    assert abs(drag - locked["drag"]) <= BAR
    side = abs(load[2]) / math.hypot(load[0], load[1])
    moment = 0.5 * 0.08 ** 2 * 2.0 * case.chord ** 3
    assert side < 1e-3
    assert math.hypot(load[3], load[4]) / moment < 1e-3
