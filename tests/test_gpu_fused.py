"""GPU parity of the fused arithmetic (the production default) and of the
limiter accounting of wall-free tiles.

The fused collision regroups the reference's operators over opposite
direction pairs with fused multiply-adds, so it differs from the oracle by
rounding only.  Bars: populations within 1e-5 relative L2 of the FP32
oracle after 1,000 steps and velocities as close to FP64 as the FP32 oracle
itself; force history within 1e-4 relative; masks and link counts bit-exact
(the geometry kernels do not depend on the arithmetic mode).
"""
import json
import math
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


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


FUSED = bh.ARITHMETIC_FUSED
STRICT = bh.ARITHMETIC_STRICT


def test_fused_shear_wave_1000_steps_64(golden_dir):
    # Velocity is a small difference of O(0.05) populations: after 1000
    # steps any change of FP32 rounding order moves it by ~1e-4 relative.
    # The fused path is held to the populations bar against the FP32
    # oracle and to the FP32 oracle's own velocity error against FP64.
    say("case: fused shear wave 64^3, 1000 steps, TRT tau 0.8 magic 1/4")
    p = cases.shear_wave_params(64)
    with bh.BirdFlowSimulation(cases.solver_config(p, arithmetic=FUSED)) as sim:
        sim.init_shear_wave(cases.SHEAR_AMPLITUDE)
        sim.advance(1000)
        g = sim.read_populations()
        _, u_g = sim.copy_macroscopic_fields()
    f = bfref.init_shear_wave(p, cases.SHEAR_AMPLITUDE, np.float32)
    f64 = bfref.init_shear_wave(p, cases.SHEAR_AMPLITUDE, np.float64)
    for _ in range(1000):
        f, *_ = bfref.step(p, f)
        f64, *_ = bfref.step(p, f64)
    _, u_o = bfref.moments(p, f)
    _, u_64 = bfref.moments(p, f64)
    e_fused, e_oracle = rel_l2(u_g, u_64), rel_l2(u_o, u_64)
    say(f"  populations rel L2 vs FP32 oracle {rel_l2(g, f):.3e}; velocity rel L2 vs FP64: fused {e_fused:.3e}, "
        f"FP32 oracle {e_oracle:.3e}; fused vs FP32 oracle {rel_l2(u_g, u_o):.3e}")
    assert rel_l2(g, f) <= 1e-5
    assert e_fused <= 2.0 * e_oracle + 1e-6
    ref = json.load(open(os.path.join(golden_dir, "shear_wave_decay.json")))["n64_tau0.8_magic0.25_1000steps"]
    amp, amp32 = cases.mode_amplitude(p, g), cases.mode_amplitude(p, f)
    ref_amp = ref["measured_amplitude"]
    say(f"  amplitude {amp:.10e} vs reference FP64 {ref_amp:.10e} (FP32 oracle {amp32:.10e})")
    # fused FP32 rounding shifts the 1000-step amplitude by about 1.3e-4
    assert abs(amp - ref_amp) / ref_amp <= 2e-4
    k = 2 * math.pi / 64
    exact = 0.01 * math.exp(-0.1 * k * k * 1000)
    assert abs(amp - exact) / exact < 1e-3


@pytest.mark.parametrize("collision", [bh.COLLISION_TRT_SYMMETRIC_LIMITER, bh.COLLISION_REGULARIZED,
                                       bh.COLLISION_RR3])
def test_fused_shear_wave_200_steps_48(collision):
    say(f"case: fused shear wave 48^3, 200 steps, collision {collision}")
    p = cases.shear_wave_params(48)
    p.collision = collision
    with bh.BirdFlowSimulation(cases.solver_config(p, arithmetic=FUSED)) as sim:
        sim.init_shear_wave(0.02)
        sim.advance(200)
        g = sim.read_populations()
    f = bfref.init_shear_wave(p, 0.02, np.float32)
    for _ in range(200):
        f, *_ = bfref.step(p, f)
    err = rel_l2(g - bfref.init_equilibrium(p, dtype=np.float32), f - bfref.init_equilibrium(p, dtype=np.float32))
    say(f"  populations rel L2 {rel_l2(g, f):.3e}; non-equilibrium part rel L2 {err:.3e}")
    assert rel_l2(g, f) <= 1e-5 and err <= 1e-4


@pytest.mark.parametrize("collision", [bh.COLLISION_TRT, bh.COLLISION_TRT_SYMMETRIC_LIMITER,
                                       bh.COLLISION_REGULARIZED, bh.COLLISION_RR3])
def test_fused_static_walls_farfield_sponge_one_step(collision):
    say(f"case: fused, static random walls, far field + sponge, collision {collision}")
    rng = np.random.default_rng(3 + collision)
    shape = (70, 30, 24)
    n = shape[0] * shape[1] * shape[2]
    p = bfref.make_params(shape, omega_plus=1.0 / 0.53, omega_minus=1.0 / 1.2,
                          bc=((1, 1), (1, 1), (1, 1)), collision=collision,
                          wall_rule=bfref.WALL_INTERPOLATED_SDF, force_mode=6, accumulate_loads=True,
                          sponge_strength=0.3, sponge_width=4.0, far_u=(0.06, 0.01, 0.0),
                          body_position=(35.0, 15.0, 12.0))
    solid = np.zeros((shape[2], shape[1], shape[0]), np.uint8)
    solid[8:16, 10:20, 20:40] = 2
    prev = solid.copy()
    prev[8:16, 10:20, 21:41] = 2
    wall = np.zeros((n, 4), np.float32)
    wall[:, :3] = rng.uniform(-0.05, 0.05, (n, 3))
    wall[:, 3] = rng.uniform(-1.0, 2.0, n)
    f = (bfref.init_equilibrium(p, dtype=np.float32).reshape(19, n)
         * rng.uniform(0.97, 1.03, (19, n)).astype(np.float32)).ravel()
    fo, lo, *_ = bfref.step(p, f, prev.ravel(), solid.ravel(), wall)
    with bh.BirdFlowSimulation(cases.solver_config(p, arithmetic=FUSED)) as sim:
        sim.set_geometry(solid.ravel(), wall, mask_prev=prev.ravel())
        sim.write_populations(f)
        sim.advance(1)
        g = sim.read_populations()
        lg = sim.read_loads()
    err = np.abs(g - fo).max()
    say(f"  max |gpu - oracle| {err:.3e}, rel L2 {rel_l2(g, fo):.3e}; loads {lg[:3]} vs {lo[:3]}")
    assert rel_l2(g, fo) <= 1e-6 and err <= 1e-6
    # the loads of one step come from the pulled populations and the wall
    # rule, which do not depend on the collision arithmetic
    assert np.allclose(lg[:6], lo[:6], rtol=1e-9, atol=1e-9 * np.abs(lo[:6]).max())


def test_fused_flapping_plate_history():
    steps = 60
    p = cases.plate_params()
    seq = workloads.flapping_plate(steps, center=(24.0, 24.0, 24.0), span=20.0, chord=8.0, thickness=2.0,
                                   heave=2.0, period=200.0, half_thickness=0.5)
    say(f"case: fused flapping plate {p.nx}x{p.ny}x{p.nz}, RR3, interpolated walls, {steps} steps")
    f_o, loads_o, geo_o = cases.run_mesh_oracle(p, seq, steps, record_masks=True)
    with bh.BirdFlowSimulation(cases.solver_config(p, arithmetic=FUSED)) as sim:
        sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
        sim.init_equilibrium()
        sim.advance(steps)
        g = sim.read_populations()
        hist = sim.load_history()
        mask, _ = sim.read_geometry()
    fscale = np.abs(loads_o[:, :3]).max()
    err = np.abs(hist[:, :3] - loads_o[:, :3]).max() / fscale
    say(f"  populations rel L2 {rel_l2(g, f_o):.3e}; force history max rel err {err:.3e}")
    assert np.array_equal(mask, geo_o[-1][0])
    assert rel_l2(g, f_o) <= 1e-5
    assert err <= 1e-4


def test_fused_translating_sphere_history(golden_dir):
    ref = json.load(open(os.path.join(golden_dir, "translating_body.json")))
    say("case: fused translating sphere 24^3, 40 steps, force mode 6")
    p = cases.translating_sphere_params(6)
    fx_o, *_ = cases.run_translating_sphere_oracle(force_mode=6)
    with bh.BirdFlowSimulation(cases.solver_config(p, arithmetic=FUSED)) as sim:
        sim.set_translating_sphere(cases.SPHERE_CENTER_RADIUS, cases.SPHERE_GEOMETRY_VELOCITY,
                                   cases.SPHERE_WALL_VELOCITY)
        sim.init_equilibrium()
        sim.advance(40)
        hist = sim.load_history()
    err = np.abs(hist[:, 0] - fx_o).max() / np.abs(fx_o).max()
    say(f"  Fx history max rel err {err:.3e}; mean {hist[:, 0].mean():.10f} vs reference "
        f"{ref['conservativeMeanForceX']:.10f}")
    assert err <= 1e-4
    assert abs(hist[:, 0].mean() - ref["conservativeMeanForceX"]) / abs(ref["conservativeMeanForceX"]) < 1e-5


def _limiter_case(collision):
    shape = (128, 8, 8)
    n = shape[0] * shape[1] * shape[2]
    p = bfref.make_params(shape, omega_plus=1.0 / 0.55, omega_minus=1.0 / 1.2, collision=collision,
                          accumulate_loads=True)
    rng = np.random.default_rng(11 + collision)
    f = (bfref.init_equilibrium(p, dtype=np.float32).reshape(19, n)
         * rng.uniform(0.1, 1.9, (19, n)).astype(np.float32)).ravel()
    return p, f


@pytest.mark.parametrize("collision", [bh.COLLISION_TRT_SYMMETRIC_LIMITER, bh.COLLISION_REGULARIZED,
                                       bh.COLLISION_RR3])
def test_limiter_activations_of_wall_free_tiles(collision):
    # Periodic box without walls or sponge: every tile takes the fast path,
    # and strong perturbations make the positivity limiters act in many cells.
    say(f"case: limiter accounting on wall-free tiles, 128x8x8, collision {collision}")
    p, f = _limiter_case(collision)
    fo, lo, *_ = bfref.step(p, f)
    assert lo[6] > 50
    for arith in (STRICT, FUSED):
        with bh.BirdFlowSimulation(cases.solver_config(p, arithmetic=arith)) as sim:
            sim.write_populations(f)
            sim.advance(1)
            g = sim.read_populations()
            lg = sim.read_loads()
            sim.advance(1)
            lg2 = sim.read_loads()
        say(f"  arithmetic {arith}: activations {lg[6]:.0f} vs oracle {lo[6]:.0f}, max restriction "
            f"{lg[7]:.6f} vs {lo[7]:.6f}; rel L2 {rel_l2(g, fo):.3e}")
        if arith == STRICT:
            assert np.array_equal(g, fo)
            assert lg[6] == lo[6] and lg[7] == lo[7]
        else:
            assert rel_l2(g, fo) <= 1e-6
            assert abs(lg[6] - lo[6]) <= max(2.0, 0.002 * lo[6]) and abs(lg[7] - lo[7]) <= 1e-4
        # the counters start from zero again on the next step
        fo2, lo2, *_ = bfref.step(p, fo)
        assert abs(lg2[6] - lo2[6]) <= (0 if arith == STRICT else max(2.0, 0.002 * lo2[6]))


@pytest.mark.parametrize("storage,bar", [(bh.STORAGE_FP16_OFFSET, 5e-3), (bh.STORAGE_BF16_OFFSET, 5e-2)])
def test_reduced_storage_plate_force_history_vs_fp64(storage, bar):
    # FP16 offsets: force history within 0.5% of an FP64 run on the same
    # inputs; BF16 offsets carry ~3 significant digits and are held to 5%.
    steps = 60
    p = cases.plate_params()
    seq = workloads.flapping_plate(steps, center=(24.0, 24.0, 24.0), span=20.0, chord=8.0, thickness=2.0,
                                   heave=2.0, period=200.0, half_thickness=0.5)
    say(f"case: flapping plate, storage {storage}, {steps} steps, force history vs FP64 oracle")
    _, loads64, _ = cases.run_mesh_oracle(p, seq, steps, dtype=np.float64)
    with bh.BirdFlowSimulation(cases.solver_config(p, storage=storage)) as sim:
        sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
        sim.init_equilibrium()
        sim.advance(steps)
        hist = sim.load_history()
    scale = np.abs(loads64[:, :3]).max()
    err = np.abs(hist[:, :3] - loads64[:, :3]).max() / scale
    mean_err = np.abs(hist[:, 2].mean() - loads64[:, 2].mean()) / abs(loads64[:, 2].mean())
    say(f"  force history max |diff| / max |F| = {err:.3e}; mean lift rel diff {mean_err:.3e}")
    assert err <= bar


BIRD_SHAPE = (256, 128, 128)
BIRD_STEPS = 80


def half_bird(steps):
    """The 512x256x256 synthetic bird scaled by 1/2 (body, wings, tail and
    stroke period) into 256x128x128."""
    return workloads.synthetic_bird(steps, seed=7, center=(80.0, 64.0, 64.0), body_radii=(20.0, 7.0, 7.0),
                                    half_span=50.0, root_chord=20.0, tip_chord=6.0, tail=(15.0, 10.0),
                                    period=600.0)


def bird_params():
    return bfref.make_params(
        BIRD_SHAPE, omega_plus=1.0 / 0.52, omega_minus=1.0 / (0.5 + (3.0 / 16.0) / 0.02),
        bc=((bfref.BC_FARFIELD, bfref.BC_OUTLET), (bfref.BC_PERIODIC, bfref.BC_PERIODIC),
            (bfref.BC_PERIODIC, bfref.BC_PERIODIC)),
        collision=bfref.COLLISION_RR3, wall_rule=bfref.WALL_INTERPOLATED_SDF, force_mode=6, accumulate_loads=True,
        far_u=(0.04, 0.0, 0.0), body_position=(80.0, 64.0, 64.0))


@pytest.fixture(scope="module")
def bird_fp64_loads():
    """FP64 oracle force history of the half-scale bird on the same inputs:
    each step's masks and wall field are the GPU's (bit-identical to the
    oracle's FP32 raster, checked elsewhere), read back step by step."""
    p = bird_params()
    seq = half_bird(BIRD_STEPS)
    loads = np.zeros((BIRD_STEPS, 8))
    with bh.BirdFlowSimulation(cases.solver_config(p)) as sim:
        sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
        # the device populations are stepped too (the fluid kernels' guard
        # rejects a state never initialized)
        sim.init_equilibrium()
        prev, wall = sim.read_geometry()
        f = bfref.init_equilibrium(p, prev, wall, np.float64)
        for s in range(BIRD_STEPS):
            sim.advance(1)
            mask, wall = sim.read_geometry()
            f, l, _, _ = bfref.step(p, f, prev, mask, wall)
            loads[s] = l
            prev = mask
    return loads


@pytest.mark.parametrize("storage,bar", [(bh.STORAGE_FP16_OFFSET, 5e-3), (bh.STORAGE_BF16_OFFSET, 5e-2)])
def test_reduced_storage_bird_force_history_vs_fp64(bird_fp64_loads, storage, bar):
    # The precision case on the articulated bird (40k triangles, RR3, moving
    # walls with cover / uncover) at half scale: FP16 offsets within 0.5 % of
    # an FP64 run on the same inputs, BF16 within 5 %.
    p = bird_params()
    seq = half_bird(BIRD_STEPS)
    say(f"case: half-scale bird {BIRD_SHAPE}, storage {storage}, {BIRD_STEPS} steps, force history vs FP64 oracle")
    with bh.BirdFlowSimulation(cases.solver_config(p, storage=storage)) as sim:
        sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
        sim.init_equilibrium()
        sim.advance(BIRD_STEPS)
        hist = sim.load_history()
    ref = bird_fp64_loads
    scale = np.abs(ref[:, :3]).max()
    err = np.abs(hist[:, :3] - ref[:, :3]).max() / scale
    tail = slice(BIRD_STEPS // 2, None)
    mean_err = np.abs(hist[tail, :3].mean(0) - ref[tail, :3].mean(0)).max() / np.abs(ref[tail, :3].mean(0)).max()
    say(f"  force history max |diff| / max |F| = {err:.3e}; mean force (second half) rel diff {mean_err:.3e}")
    assert err <= bar
