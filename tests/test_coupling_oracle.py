"""The oracle replays the reference's indexed-dove production coupling
(MetalIndexedBirdSurfaceValidation.swift:19284-19600; archived result
ValidationArtifacts/deetjen-dove-indexed-production-coupling.json, copied as
data into tests/golden/deetjen_coupling.json): 59x53x50 periodic, TRT at the
replay's lattice scaling, signed-distance link fractions, conservative
mode-6 loads, mesh frames 0 and 1 of the Deetjen dove, 8 steps.

Integer flags and link counts must equal the reference's exactly: newly
covered and uncovered cells, persistent wall links (target fluid in both
steps, source solid in both) and occupancy transitions, each step.  The
aerodynamic impulse (force x dt, Newton seconds) agrees within a cross-
hardware bound (Apple M4 FP32 against this FP32 restatement; different
summation order of the loads).

Step 8 has one cell whose signed distance is 9.8e-7 cells (1e-8 m) on the
fluid side: below one FP32 unit in the last place of the vertex coordinates
(about 3e-6 cells here), so its sign is the Apple GPU's rounding.  The
reference has it solid; with that one cell solid the oracle reproduces the
reference's step-8 counts and impulse as well.
"""
import numpy as np

import bfref
import cases

IMPULSE_BOUND = 5e-5      # max component error / max |reference impulse|, per step


def test_coupling_counts_and_impulse_match_reference(golden_dir):
    c = cases.DeetjenCoupling(golden_dir)
    ref = c.ref["samples"]
    assert c.grid == (c.ref["gridX"], c.ref["gridY"], c.ref["gridZ"])
    rows = c.run_oracle(8)
    for r, smp in zip(rows[:7], ref[:7]):
        got = (r["covered"], r["uncovered"], r["persistent"], r["transitions"])
        expect = (smp["newlyCoveredCellCount"], smp["newlyUncoveredCellCount"],
                  smp["persistentBoundaryLinkCount"], smp["sourceLedgerTransitionCellCount"])
        assert got == expect, (smp["step"], got, expect)
        imp = r["load"][:3] * float(c.dt)
        ri = np.array(smp["aerodynamicImpulse"])
        assert np.abs(imp - ri).max() / np.abs(ri).max() < IMPULSE_BOUND, (smp["step"], imp, ri)
    # step 8: exactly one borderline cell, and it explains the difference
    m7 = rows[6]["mask"]
    m8, w8 = c.geometry(8)
    border = np.where((m8 == 0) & (np.abs(w8[:, 3]) < 2e-6))[0]
    assert len(border) == 1
    m8b = m8.copy()
    m8b[border] = 4
    f7 = rows[6]["populations"]
    fo, ld, _, _ = bfref.step(c.params, f7, m7, m8b, w8)
    _, counts = bfref.ledger_sources(c.params, f7, fo, m7, m8b, w8)
    got = (int(((m7 == 0) & (m8b != 0)).sum()), int(((m7 != 0) & (m8b == 0)).sum()), int(counts[2]),
           int(counts[3]))
    smp = ref[7]
    assert got == (smp["newlyCoveredCellCount"], smp["newlyUncoveredCellCount"],
                   smp["persistentBoundaryLinkCount"], smp["sourceLedgerTransitionCellCount"]), got
    ri = np.array(smp["aerodynamicImpulse"])
    assert np.abs(ld[:3] * float(c.dt) - ri).max() / np.abs(ri).max() < IMPULSE_BOUND
    # totals over the 8 steps, and the final per-part solid counts
    assert sum(r["covered"] for r in rows[:7]) + got[0] == c.ref["newlyCoveredCellEvents"]
    assert sum(r["uncovered"] for r in rows[:7]) + got[1] == c.ref["newlyUncoveredCellEvents"]
    assert sum(r["persistent"] for r in rows[:7]) + got[2] == c.ref["persistentBoundaryLinkEvents"]
    assert [int((m8b == i).sum()) for i in range(1, 5)] == c.ref["componentSolidCellCounts"]
