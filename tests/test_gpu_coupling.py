"""The reference's indexed-dove production coupling on the device
(MetalIndexedBirdSurfaceValidation.swift:19284-19600; fixture
tests/golden/deetjen_coupling.json): 59x53x50 periodic, TRT, signed-distance
link fractions, mode-6 loads, Deetjen mesh frames 0 and 1, 8 steps.

Each step the device's integer flags and counts must equal the oracle's
exactly (masks, per-direction link census, newly covered / uncovered cells,
persistent wall links, occupancy transitions), its populations bit for bit
(strict FP32 arithmetic) and its loads to FP64 summation order.  Steps 1-7
must also equal the reference's archived counts exactly and its impulse
within the cross-hardware bound of the oracle test; step 8 holds one cell
whose sign is the reference GPU's rounding (see test_coupling_oracle.py).
"""
import sys

import numpy as np
import pytest

import bfref
import birdflow_hip as bh
import cases

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_device_replays_reference_coupling_counts(golden_dir):
    say("case: Deetjen indexed production coupling 59x53x50, 8 steps, FP32 strict")
    c = cases.DeetjenCoupling(golden_dir)
    rows = c.run_oracle(8)
    cfg = cases.solver_config(c.params, time_step=float(c.dt), velocity_to_lattice=float(c.velocity_to_lattice))
    with bh.BirdFlowSimulation(cfg) as sim:
        sim.upload_mesh(c.points, c.times, c.triangles, c.parts, c.half_thickness, time_start=float(c.times[0]))
        sim.init_equilibrium()
        mask0, _ = c.geometry(0)
        m, _ = sim.read_geometry()
        assert np.array_equal(m, mask0)
        prev = mask0
        for s, r in enumerate(rows, start=1):
            sim.advance(1)
            census, cov, unc = sim.link_census()
            src = sim.step_momentum_sources()
            loads = sim.read_loads()
            mask, _ = sim.read_geometry()
            g = sim.read_populations()
            assert np.array_equal(mask, r["mask"]), s
            ocensus, ocov, ounc = bfref.link_census(c.params, prev, r["mask"])
            assert np.array_equal(census, ocensus) and (cov, unc) == (ocov, ounc) == (r["covered"], r["uncovered"])
            assert (src["persistent_links"], src["topology_transitions"]) == (r["persistent"], r["transitions"])
            assert np.array_equal(g.view(np.uint32), r["populations"].view(np.uint32)), s
            err = np.abs(loads[:6] - r["load"][:6]).max() / np.abs(r["load"][:3]).max()
            assert err < 1e-12, (s, err)
            if s <= 7:
                smp = c.ref["samples"][s - 1]
                assert (cov, unc, src["persistent_links"], src["topology_transitions"]) == (
                    smp["newlyCoveredCellCount"], smp["newlyUncoveredCellCount"],
                    smp["persistentBoundaryLinkCount"], smp["sourceLedgerTransitionCellCount"]), s
                ri = np.array(smp["aerodynamicImpulse"])
                assert np.abs(loads[:3] * float(c.dt) - ri).max() / np.abs(ri).max() < 5e-5, s
            say(f"step {s}: covered {cov} uncovered {unc} persistent links {src['persistent_links']} "
                f"load err vs oracle {err:.1e}")
            prev = r["mask"]
