"""The synthetic articulated bird (about 40k triangles: ellipsoid body 40 x
14 x 14 cells, tapered wings of half-span 100 and chord 40 -> 12, a 30 x 20
fan tail, 3-harmonic wing kinematics from seed 7) in its 512 x 256 x 256
domain: RR3, interpolated moving walls with signed-distance link fractions,
inflow u_x = 0.04 at -x, zero-gradient outlet at +x, periodic y / z,
tau+ = 0.52, strict FP32.

Against the oracle, step by step for 3 steps:
* part masks bit-identical everywhere (36,000+ triangle ids above 4,096
  included), wall fields bit-identical wherever the step reads them (the
  candidate band), the per-direction link census (fluid cells whose pull
  source is solid) and the covered / uncovered counts exactly;
* populations bit-identical and loads to 1e-12 (FP64 sums of the same FP32
  terms in another order).

And the cache of the static triangles' keys (the body and the tail do not
move: their keys are seeded each step and only the 25,600 wing triangles are
rastered) against rastering every triangle every step (BIRDFLOW_STATIC_KEYS
=0), at 256 x 128 x 128 for FP32 and FP16: masks, walls and stored
populations bit-identical for 12 steps."""
import os
import sys

import numpy as np
import pytest

import bfref
import birdflow_hip as bh
import cases
from birdflow_hip import workloads

pytestmark = pytest.mark.gpu

TAU = 0.52


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bird_setup(shape, steps, storage=bh.STORAGE_FP32):
    s = shape[0] / 512.0
    center = (160.0 * s, 128.0 * s, 128.0 * s)
    if s == 1.0:
        seq = workloads.synthetic_bird(steps + 1, seed=7, center=center)
    else:
        seq = workloads.synthetic_bird(steps + 1, seed=7, center=center, body_radii=(40.0 * s, 14.0 * s, 14.0 * s),
                                       half_span=100.0 * s, root_chord=40.0 * s, tip_chord=12.0 * s,
                                       tail=(30.0 * s, 20.0 * s), period=1200.0 * s)
    p = bfref.make_params(shape, omega_plus=1.0 / TAU, omega_minus=1.0 / (0.5 + (3.0 / 16.0) / (TAU - 0.5)),
                          bc=((bfref.BC_FARFIELD, bfref.BC_OUTLET), (bfref.BC_PERIODIC, bfref.BC_PERIODIC),
                              (bfref.BC_PERIODIC, bfref.BC_PERIODIC)),
                          collision=bfref.COLLISION_RR3, wall_rule=bfref.WALL_INTERPOLATED_SDF, force_mode=6,
                          accumulate_loads=True, far_u=(0.04, 0.0, 0.0), body_position=center)
    return p, seq, cases.solver_config(p, storage=storage)


def test_synthetic_bird_matches_oracle_at_full_size():
    steps = 3
    shape = (512, 256, 256)
    p, seq, cfg = bird_setup(shape, steps)
    say(f"case: synthetic bird {shape}, {len(seq.triangles)} triangles, RR3, strict FP32, {steps} steps vs oracle")
    assert len(seq.triangles) - 4096 > 35000      # ids above the reference's 12-bit limit
    with bh.BirdFlowSimulation(cfg) as sim:
        sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
        sim.init_equilibrium()
        mask0, wall0 = cases.mesh_geometry_oracle(p, seq, float(seq.times[0]))
        f = bfref.init_equilibrium(p, mask0, wall0, np.float32)
        prev = mask0
        for s in range(1, steps + 1):
            sim.advance(1)
            mask_g, wall_g = sim.read_geometry()
            census_g, cov_g, unc_g = sim.link_census()
            mask_o, wall_o = cases.mesh_geometry_oracle(p, seq, float(seq.times[0]) + s)
            census_o, cov_o, unc_o = bfref.link_census(p, prev, mask_o)
            f, ld, _, _ = bfref.step(p, f, prev, mask_o, wall_o)
            g = sim.read_populations()
            ld_g = sim.read_loads()
            band = wall_o[:, 3] <= BAND
            solid = int((mask_o != 0).sum())
            ids_high = int(np.unique(mask_o).size)
            diff = int((g.view(np.uint32) != f.view(np.uint32)).sum())
            rel = float(np.abs(ld_g[:6] - ld[:6]).max() / max(np.abs(ld[:6]).max(), 1e-30))
            say(f"  step {s}: solid cells {solid}, parts {ids_high - 1}, mask equal {np.array_equal(mask_g, mask_o)}, "
                f"walls in band equal {np.array_equal(wall_g[band].view(np.uint32), wall_o[band].view(np.uint32))}, "
                f"links {int(census_g.sum())} vs {int(census_o.sum())}, "
                f"covered {cov_g}/{cov_o}, uncovered {unc_g}/{unc_o}, "
                f"populations differing {diff}, loads rel err {rel:.2e}")
            assert np.array_equal(mask_g, mask_o)
            assert np.array_equal(wall_g[band].view(np.uint32), wall_o[band].view(np.uint32))
            assert np.array_equal(census_g.astype(np.int64), census_o.astype(np.int64))
            assert (cov_g, unc_g) == (cov_o, unc_o)
            assert diff == 0
            assert rel <= 1e-12
            prev = mask_o
        assert solid > 20000 and int(census_o.sum()) > 100000


# Stepping keys the cells within half thickness + 1.5 cells of the mesh:
# signed distance (distance - half thickness) / dx <= 1.5
BAND = 1.5


@pytest.mark.parametrize("storage", [bh.STORAGE_FP32, bh.STORAGE_FP16_OFFSET], ids=["fp32", "fp16"])
def test_static_key_cache_matches_the_full_raster(storage):
    steps = 12
    shape = (256, 128, 128)
    p, seq, cfg = bird_setup(shape, steps, storage)
    say(f"case: synthetic bird {shape}, static-triangle key cache vs full raster, storage {storage}, {steps} steps")
    out = {}
    for cached in (False, True):
        old = os.environ.get("BIRDFLOW_STATIC_KEYS")
        os.environ["BIRDFLOW_STATIC_KEYS"] = "1" if cached else "0"
        try:
            sim = bh.BirdFlowSimulation(cfg)
        finally:
            if old is None:
                del os.environ["BIRDFLOW_STATIC_KEYS"]
            else:
                os.environ["BIRDFLOW_STATIC_KEYS"] = old
        snaps = []
        with sim:
            sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
            sim.init_equilibrium()
            for s in range(steps):
                sim.advance(1)
                if s % 4 == 3:
                    mask, wall = sim.read_geometry()
                    snaps.append((mask, wall, sim.read_population_storage().copy()))
            loads = sim.load_history(steps)
        out[cached] = (snaps, loads)
    for k, (a, b) in enumerate(zip(out[False][0], out[True][0])):
        mask_same = np.array_equal(a[0], b[0])
        wall_same = np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
        pop_diff = int((a[2] != b[2]).sum())
        say(f"  snapshot {k}: masks equal {mask_same}, walls equal {wall_same}, stored values differing {pop_diff}, "
            f"solid cells {int((a[0] != 0).sum())}")
        assert mask_same and wall_same and pop_diff == 0
    assert np.array_equal(out[False][1].view(np.uint64), out[True][1].view(np.uint64))
