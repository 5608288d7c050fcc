"""Fluid-state safety (the reference's population minimum,
reducePopulationMinimum, BirdFlow.metal:6404-6459, and its non-finite
check, MetalIndexedBirdSurfaceValidation.swift:18568-18600):

* the device minimum equals numpy's over the read-back populations, ties to
  the smallest direction-major index, for FP32 and FP16 storage;
* injected values (the fault-injection hook): a negative population is the
  minimum at its index; NaN and -inf rank first, the smaller index winning;
* a NaN injected into the lattice is caught by the fluid kernels' guard in
  the next step: the advance reports a runtime-safety violation once, the
  safety record carries flag 8 and the step, and a new population state
  clears it.
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


def shear_sim(storage=bh.STORAGE_FP32, n=24):
    p = cases.shear_wave_params(n)
    sim = bh.BirdFlowSimulation(cases.solver_config(p, storage=storage))
    sim.init_shear_wave(0.05)
    return sim


@pytest.mark.parametrize("storage", [bh.STORAGE_FP32, bh.STORAGE_FP16_OFFSET], ids=["fp32", "fp16"])
def test_population_minimum_matches_numpy(storage):
    say(f"case: population minimum of a shear wave, storage {storage}")
    with shear_sim(storage) as sim:
        sim.advance(3)
        f = sim.read_populations()
        value, index, bad = sim.population_minimum()
        want = int(np.argmin(f))          # numpy: the first index of the minimum
        say(f"  device {value!r} at {index}, numpy {f[want]!r} at {want}")
        assert not bad and index == want and np.float32(value) == f[want]
        # a tie: the same minimum at a smaller index of the same direction
        # (the same stored offset for FP16 / BF16) wins
        q, cell = divmod(want, sim.n)
        tie = q * sim.n + cell // 2
        sim.inject_population(tie, float(f[want]))
        value, index, bad = sim.population_minimum()
        assert index == tie and np.float32(value) == f[want]
        # an injected negative population (FP16 stores -0.25 - w_7 rounded to
        # half precision: the minimum is what reads back)
        sim.inject_population(7 * sim.n + 5, -0.25)
        stored = sim.read_populations()[7 * sim.n + 5]
        value, index, bad = sim.population_minimum()
        assert abs(stored + 0.25) <= (0.0 if storage == bh.STORAGE_FP32 else 2e-4)
        assert (np.float32(value), index, bad) == (stored, 7 * sim.n + 5, False)


def test_non_finite_populations_rank_first():
    say("case: NaN and -inf injected, minimum ranks them first by index")
    with shear_sim() as sim:
        sim.inject_population(5 * sim.n + 17, -0.5)
        sim.inject_population(9 * sim.n + 3, float("nan"))
        sim.inject_population(11 * sim.n + 40, float("-inf"))
        value, index, bad = sim.population_minimum()
        assert bad and index == 9 * sim.n + 3 and np.isnan(value)
        sim.inject_population(9 * sim.n + 3, 0.1)
        value, index, bad = sim.population_minimum()
        assert bad and index == 11 * sim.n + 40 and value == float("-inf")


@pytest.mark.parametrize("n,storage", [(24, bh.STORAGE_FP32), (64, bh.STORAGE_FP32), (128, bh.STORAGE_FP16_OFFSET)],
                         ids=["general tiles", "fast tiles", "fast tile pairs fp16"])
def test_injected_nan_is_caught_by_the_step_guard(n, storage):
    # 24^3: every tile is partial (general kernel); 64^3: whole tiles (fast
    # kernel); 128^3 FP16: tile pairs (two cells per lane)
    say(f"case: NaN injected after 4 steps into a {n}^3 shear wave, storage {storage}; the guard of step 5")
    with shear_sim(storage, n) as sim:
        sim.advance(4)
        _, _, flags = sim.safety()
        assert flags == 0
        sim.inject_population(3 * sim.n + 100, float("nan"))
        with pytest.raises(bh.BirdFlowError) as err:
            sim.advance(2)
        assert err.value.status == 9
        _, step, flags = sim.safety()
        say(f"  error {err.value}; safety flags {flags}, step {step}")
        assert flags & 8 and step == 5
        sim.advance(1)                    # reported once
        f = sim.read_populations()
        assert np.isnan(f).any()
        sim.init_shear_wave(0.05)         # a new state clears the record
        sim.advance(2)
        assert sim.safety()[2] == 0
