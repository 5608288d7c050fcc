"""The no-allocation ledger entry point, bf_ledger_step_sources_noalloc,
through the C ABI against bf_ledger_step_sources on the same step: the same
kernels and block count, so sums and counts must agree bit for bit.

Two contexts in a row share the GPU's one scratch array: the first on a
128 x 96 x 96 grid (1.18M cells: the ledger's full 4,096 block partials), the
second on a smaller 48^3 grid, so the second call runs over partials the first
left behind.  Each call is checked before and after the allocating call, and
the no-allocation call leaves the context's device allocation unchanged
(bf_allocated_bytes).  The case is the compact bird in free flight under
gravity (far-field faces, sponge, moving walls), whose step has open-boundary
links, sponge cells and persistent links.
"""
import sys

import numpy as np
import pytest

import birdflow_hip as bh
import cases

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def free_bird(shape):
    p, cfg, bird, obird, sc = cases.compact_bird_case(free_flight=True, shape=shape)
    sim = bh.BirdFlowSimulation(cfg)
    state0 = np.array(list(p.body_position) + [0.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0])
    sim.set_body(state0, bh.BodyParameters(mass=0.025, inertia=(4e-6, 7e-6, 8e-6), gravity=(0.0, 0.0, -0.5),
                                           substeps=1))
    sim.set_bird(bird)
    sim.init_equilibrium()
    return sim


def as_bits(src):
    sums = np.concatenate([src["change"], src["open_boundary"], src["sponge"], src["persistent_link"]])
    counts = [src["open_boundary_links"], src["sponge_cells"], src["persistent_links"], src["topology_transitions"]]
    return sums.view(np.uint64).tolist(), counts


def test_no_alloc_ledger_matches_the_allocating_entry_point():
    for shape in ((128, 96, 96), (48, 48, 48)):
        say(f"case: no-allocation ledger vs bf_ledger_step_sources, compact bird free flight, {shape}, 2 steps")
        with free_bird(shape) as sim:
            for step in range(2):
                sim.advance(1)
                before = sim.allocated_bytes
                first = as_bits(sim.step_momentum_sources(allocate=False))
                after = sim.allocated_bytes
                reference = as_bits(sim.step_momentum_sources(allocate=True))
                if step == 0:
                    # the allocating entry point's first call takes its buffers
                    # (so the accounting above would have seen an allocation)
                    assert sim.allocated_bytes > after
                again = as_bits(sim.step_momentum_sources(allocate=False))
                say(f"  step {step + 1}: counts {first[1]} vs {reference[1]}; sums bit-identical "
                    f"{first[0] == reference[0]}, repeat {again == reference}; allocated bytes {before} -> {after}")
                assert first == reference and again == reference
                assert after == before
                assert reference[1][0] > 0 and reference[1][1] > 0


def test_no_alloc_ledger_refuses_what_the_allocating_one_refuses():
    say("case: no-allocation ledger before any step and after a population write")
    with free_bird((48, 48, 48)) as sim:
        with pytest.raises(bh.BirdFlowError):
            sim.step_momentum_sources(allocate=False)
        sim.advance(1)
        sim.write_populations(sim.read_populations())
        with pytest.raises(bh.BirdFlowError):
            sim.step_momentum_sources(allocate=False)
        sim.advance(1)
        assert sim.step_momentum_sources(allocate=False)["sponge_cells"] > 0
