"""The two-cells-per-lane fast kernel of FP16 / BF16 storage
(fluid_fast_pair) against the one-cell-per-lane kernel (fluid_fast), which
the other reduced-storage tests hold to the FP64 oracle: each cell's
arithmetic is the same, so the stored populations, the loads (with the
limiter counts of wall-free tiles) and the link census must be bit-identical
step by step.  BIRDFLOW_PAIR_CELLS=0 selects the one-cell kernel at context
creation.

Cases cover what the pair kernel does differently: x-neighbour halves moved
between lanes and the halo element at both ends of a 128-cell row, across a
periodic x face, an inlet / outlet pair and far-field faces; tile pairs with
one tile on the general path (a moving plate and a bird crossing tile
boundaries); a sponge on every face; RR3's positivity limiter; and a grid
with an odd number of tiles along x, where the launcher keeps the one-cell
kernel."""
import os
import sys

import numpy as np
import pytest

import bfref
import birdflow_hip as bh
import cases
from birdflow_hip import workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(make, pair, steps):
    old = os.environ.get("BIRDFLOW_PAIR_CELLS")
    os.environ["BIRDFLOW_PAIR_CELLS"] = "1" if pair else "0"
    try:
        sim = make()
    finally:
        if old is None:
            del os.environ["BIRDFLOW_PAIR_CELLS"]
        else:
            os.environ["BIRDFLOW_PAIR_CELLS"] = old
    out = []
    with sim:
        for s in range(steps):
            sim.advance(1)
            if s % 4 == 3 or s == steps - 1:
                out.append((sim.read_population_storage().copy(), sim.link_census()[0].copy()))
        loads = sim.load_history(steps) if sim.config.accumulate_loads else None
    return out, loads


def plate(storage, shape=(128, 64, 64)):
    p = cases.plate_params(shape)
    steps = 24
    seq = workloads.flapping_plate(steps + 1, center=(40.0, 32.0, 32.0), span=24.0, chord=10.0, thickness=2.0,
                                   heave=2.0, period=100.0, half_thickness=0.5)

    def make():
        sim = bh.BirdFlowSimulation(cases.solver_config(p, storage=storage))
        sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
        sim.init_equilibrium()
        return sim
    return make, steps


def shear(storage, shape=(128, 32, 16)):
    p = cases.shear_wave_params(shape[1])
    p = bfref.make_params(shape, omega_plus=p.omega_plus, omega_minus=p.omega_minus,
                          collision=bfref.COLLISION_RR3)

    def make():
        sim = bh.BirdFlowSimulation(cases.solver_config(p, storage=storage))
        sim.init_shear_wave(0.05)
        return sim
    return make, 12


def bird(storage, shape=(128, 48, 48)):
    p, cfg, geometry, _, _ = cases.compact_bird_case(shape=shape)
    cfg.storage = storage
    cfg.arithmetic = bh.ARITHMETIC_DEFAULT

    def make():
        sim = bh.BirdFlowSimulation(cfg)
        sim.set_bird(geometry)
        sim.init_equilibrium()
        return sim
    return make, 12


CASES = {"plate inlet/outlet 128x64x64": plate, "shear wave periodic 128x32x16": shear,
         "bird far field + sponge 128x48x48": bird,
         "plate, odd tile count 192x64x64": lambda st: plate(st, (192, 64, 64))}


@pytest.mark.parametrize("storage", [bh.STORAGE_FP16_OFFSET, bh.STORAGE_BF16_OFFSET], ids=["fp16", "bf16"])
@pytest.mark.parametrize("case", list(CASES))
def test_pair_kernel_is_bit_identical_to_the_single_cell_kernel(case, storage):
    make, steps = CASES[case](storage)
    say(f"case: {case}, storage {storage}, {steps} steps, pair vs single-cell fast kernel")
    single, loads1 = run(make, False, steps)
    paired, loads2 = run(make, True, steps)
    for k, ((a, ca), (b, cb)) in enumerate(zip(single, paired)):
        diff = int((a != b).sum())
        say(f"  snapshot {k}: stored values differing {diff} of {a.size}; census equal {np.array_equal(ca, cb)}")
        assert diff == 0
        assert np.array_equal(ca, cb)
    if loads1 is not None:
        assert np.array_equal(loads1.view(np.uint64), loads2.view(np.uint64))
        say(f"  loads bit-identical over {len(loads1)} steps; limiter activations {loads1[:, 6].sum():.0f}")
