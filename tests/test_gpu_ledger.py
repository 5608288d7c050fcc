"""GPU checks of the momentum ledger (BirdFlow.metal:4272-4526,
BirdFlowSimulation.swift:805-1085):

* the device's FP64 sums of one step equal the oracle ledger of the same
  populations, masks and wall (1e-12 relative: FP64 sums in another order)
  and its counts exactly, for FP32 and FP16 storage;
* the reference's ledger case (MetalExecutionTests.swift:968-1040): the
  compact bird in free flight under gravity (0, 0, -0.5) for 4 steps closes
  the boundary and external-system balances within 0.005, with open-boundary
  links and sponge cells in every step, persistent links and occupancy
  changes in some step and a negative gravity impulse;
* the ledger refuses a step whose inputs were overwritten.
"""
import sys

import numpy as np
import pytest

import bfref
import birdflow_hip as bh
import cases
from birdflow_hip.ledger import advance_with_momentum_ledger

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def free_bird(storage=bh.STORAGE_FP32, gravity=(0.0, 0.0, -0.5)):
    p, cfg, bird, obird, sc = cases.compact_bird_case(free_flight=True)
    cfg.storage = storage
    sim = bh.BirdFlowSimulation(cfg)
    state0 = np.array(list(p.body_position) + [0.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0])
    sim.set_body(state0, bh.BodyParameters(mass=0.025, inertia=(4e-6, 7e-6, 8e-6), gravity=gravity, substeps=1))
    sim.set_bird(bird)
    sim.init_equilibrium()
    return p, sim


@pytest.mark.parametrize("storage", [bh.STORAGE_FP32, bh.STORAGE_FP16_OFFSET])
def test_device_ledger_matches_oracle_sums(storage):
    say(f"case: ledger sums of the free-flying compact bird vs the oracle, storage {storage}, steps 1-3")
    p, sim = free_bird(storage)
    with sim:
        for s in range(3):
            f_in = sim.read_populations()
            mask_prev, _ = sim.read_geometry()
            sim.advance(1)
            src = sim.step_momentum_sources()
            fluid = sim.fluid_momentum()
            f_out = sim.read_populations()
            mask, wall = sim.read_geometry()
            sums, counts = bfref.ledger_sources(p, f_in, f_out, mask_prev, mask, wall)
            got = np.array([src["change"], src["open_boundary"], src["sponge"], src["persistent_link"]])
            scale = np.maximum(np.abs(sums).max(axis=1, keepdims=True), 1e-300)
            err = float((np.abs(got - sums) / scale).max())
            ferr = float(np.abs(fluid - bfref.ledger_fluid(p, f_out, mask)).max() / np.abs(fluid).max())
            got_counts = [src["open_boundary_links"], src["sponge_cells"], src["persistent_links"],
                          src["topology_transitions"]]
            say(f"  step {s + 1}: sums max rel err {err:.2e}, fluid totals rel err {ferr:.2e}, "
                f"counts {got_counts} vs {counts.tolist()}")
            assert err <= 1e-12 and ferr <= 1e-12
            assert got_counts == counts.tolist()
            assert counts[0] > 0 and counts[1] > 0 and counts[2] > 0


def test_free_flight_ledger_closes():
    say("case: coupled momentum ledger, compact bird, free flight, gravity (0, 0, -0.5), 4 steps")
    p, sim = free_bird()
    with sim:
        report = advance_with_momentum_ledger(sim, 4)
    for smp in report["samples"]:
        say(f"  step {smp['step']}: aero {np.array(smp['aerodynamic_impulse'])}, boundary residual "
            f"{np.array(smp['boundary_residual'])}, system residual {np.array(smp['system_residual'])}, "
            f"links {smp['open_boundary_links']}/{smp['sponge_cells']}/{smp['persistent_links']}/"
            f"{smp['topology_transitions']}")
    say(f"  relative RMS residuals: boundary {report['relative_boundary_residual']:.3e}, "
        f"system {report['relative_system_residual']:.3e}")
    assert report["finite"] and report["passed"]
    s = report["samples"]
    assert all(x["open_boundary_links"] > 0 and x["sponge_cells"] > 0 for x in s)
    assert any(x["persistent_links"] > 0 and x["topology_transitions"] > 0 for x in s)
    assert all(x["gravity_impulse"][2] < 0 for x in s)


def test_ledger_refuses_overwritten_inputs():
    say("case: ledger after a population write")
    p, sim = free_bird()
    with sim:
        sim.advance(1)
        f = sim.read_populations()
        sim.write_populations(f)
        with pytest.raises(bh.BirdFlowError):
            sim.step_momentum_sources()
        sim.advance(1)
        assert sim.step_momentum_sources()["sponge_cells"] > 0
