"""The exact-root wall rule (BF_WALL_INTERPOLATED_RAY) on the GPU against the
oracle's bfo_ray_links / step with that table, bit for bit (strict FP32):

* the indexed Deetjen dove at 26.5 ms on the reference audit's D12 and D16
  grids (112 x 103 x 99 and 149 x 136 x 131 cells, 7.5 mm half thickness):
  the whole table equals the oracle's, and the 15 links of the reference's
  link-ray-root audit (deetjen-dove-moving-wall-link-ray-root.json) carry
  its exact fractions to 1e-4 of the link;
* a static box (the 12-triangle plate mesh) across the periodic y face and
  against the zero-gradient outlet at +x: links through the periodic wrap
  and the outlet's clamped pulls; table and populations over 6 steps;
* the flapping plate (96 x 48 x 48, RR3, inflow / outlet, periodic y / z):
  table every 5 steps and populations and loads after 20 steps;
* the synthetic bird scaled to 256 x 128 x 128 (40,000 triangles, moving
  wings): table, populations and loads over 3 steps;
* the articulated synthetic bird (kinematics on the device; its static
  body and tail entries cached once) at 128 x 64 x 64, 3 steps, bit for bit;
* FP16 offset storage under the rule: identical tables, force history
  within 0.5 % of FP32.
"""
import dataclasses
import json
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


def ray_config(p, storage=bh.STORAGE_FP32):
    return dataclasses.replace(cases.solver_config(p, storage=storage), wall_rule=bh.WALL_INTERPOLATED_RAY)


def links_params(p):
    """The oracle steps the exact roots through its link-table rule."""
    q = type(p).from_buffer_copy(p)
    q.wall_rule = bfref.WALL_INTERPOLATED_LINKS
    return q


def oracle_geometry(p, seq, t):
    first, _, _ = bfref.frame_interval(seq.times, t)
    pos, vel = bfref.prepare_vertices(seq.points, seq.times, first, t)
    mask, wall, _ = bfref.raster_resolve(p, pos, vel, seq.triangles, seq.parts, seq.half_thickness, key_bits=64)
    links = bfref.ray_links(p, pos, seq.triangles, seq.half_thickness, mask, wall)
    return mask, wall, links


def compare_tables(tag, got, want, mask):
    got = got.reshape(19, -1)
    want = want.reshape(19, -1)
    links = want != 0
    differing = int((got.view(np.uint32) != want.view(np.uint32)).sum())
    say(f"  {tag}: links {int(links.sum())} (GPU {int((got != 0).sum())}), entries differing {differing}, "
        f"fractions below 1/2 {int((want[links] < 0.5).sum())}")
    assert int(links.sum()) > 0
    assert differing == 0


def run_pair(p, seq, steps, check_every, storage=bh.STORAGE_FP32):
    """GPU and oracle side by side; tables compared every check_every steps,
    populations and loads at the end."""
    pl = links_params(p)
    mask0, wall0, _ = oracle_geometry(p, seq, float(seq.times[0]))
    f = bfref.init_equilibrium(p, mask0, wall0, np.float32)
    prev = mask0
    loads_o = []
    with bh.BirdFlowSimulation(ray_config(p, storage)) as sim:
        sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
        sim.init_equilibrium()
        for s in range(1, steps + 1):
            sim.advance(1)
            mask, wall, links = oracle_geometry(p, seq, float(seq.times[0]) + s)
            f, ld, _, _ = bfref.step_with_links(pl, f, prev, mask, wall, links)
            loads_o.append(ld)
            prev = mask
            if s % check_every == 0 or s == steps:
                mask_g, _ = sim.read_geometry()
                assert np.array_equal(mask_g, mask), s
                compare_tables(f"step {s}", sim.ray_links(), links, mask)
        g = sim.read_populations()
        loads = sim.load_history(steps)
    loads_o = np.array(loads_o)
    differing = int((g.view(np.uint32) != f.view(np.uint32)).sum())
    rel = float(np.abs(loads[:, :6] - loads_o[:, :6]).max() / max(np.abs(loads_o[:, :6]).max(), 1e-30))
    say(f"  after {steps} steps: populations differing {differing}, loads max rel err {rel:.2e}, "
        f"last force {loads[-1, :3]}")
    assert differing == 0
    assert rel <= 1e-12
    return loads


@pytest.mark.parametrize("case", ["d12", "d16"])
def test_deetjen_dove_table_matches_oracle_and_reference_roots(golden_dir, case):
    d = np.load(os.path.join(golden_dir, "deetjen_subset.npz"))
    ref = json.load(open(os.path.join(golden_dir, "deetjen_ray_roots.json")))
    cd = ref["cases"][case]
    grid = tuple(int(v) for v in cd["grid"])
    origin = tuple(float(np.float32(v)) for v in cd["origin"])
    dx = float(np.float32(cd["cellSize"]))
    t = float(np.float32(ref["timeSeconds"]))
    h = float(np.float32(ref["halfThicknessMeters"]))
    frames = {int(f): d["points"][i] for i, f in enumerate(d["frames"])}
    pts = np.stack([frames[26], frames[27]]).astype(np.float32)
    times = d["times"][26:28].astype(np.float32)
    say(f"case: Deetjen dove at {t * 1e3:.1f} ms on the {case} grid {grid}, exact-root table vs oracle and archive")
    p = bfref.make_params(grid, omega_plus=1.0, omega_minus=1.0, origin=origin, dx=dx,
                          bc=((bfref.BC_FARFIELD, bfref.BC_FARFIELD),) * 3,
                          wall_rule=bfref.WALL_INTERPOLATED_SDF)
    with bh.BirdFlowSimulation(ray_config(p)) as sim:
        sim.upload_mesh(pts, times, d["triangles"], d["triangle_parts"], h, time_start=t)
        mask_g, _ = sim.read_geometry()
        got = sim.ray_links()
    first, _, _ = bfref.frame_interval(times, t)
    pos, vel = bfref.prepare_vertices(pts, times, first, t)
    mask, wall, _ = bfref.raster_resolve(p, pos, vel, d["triangles"], d["triangle_parts"], h, 64)
    want = bfref.ray_links(p, pos, d["triangles"], h, mask, wall)
    assert np.array_equal(mask_g, mask)
    compare_tables(case, got, want, mask)
    nx, ny = grid[0], grid[1]
    worst = 0.0
    for link in cd["links"]:
        fx, fy, fz = link["fluidCell"]
        q = link["directionIndex"]
        gid = fx + nx * (fy + ny * fz)
        v = float(got[q, gid])
        exact = link["exactGlobalFluidToIntersectionFraction"]
        worst = max(worst, abs(v - exact))
        say(f"    link {link['solidCell']} -> {link['fluidCell']} q {q}: GPU {v:.6f}, reference exact {exact:.6f}, "
            f"reference production {link['productionFluidToIntersectionFraction']:.6f}")
        assert abs(v - exact) <= 1e-4
    say(f"  worst |GPU - reference exact root| = {worst:.2e}")


def test_static_box_across_periodic_and_outlet_faces():
    shape = (40, 32, 24)
    steps = 6
    p = cases.plate_params(shape, collision=bfref.COLLISION_TRT, omega_plus=1.0 / 0.7, inflow=0.03)
    # 6 x 8 x 3 box centred 3.3 cells from the outlet face and 1.3 cells
    # above the periodic y face: it crosses y = 0, and cells at the outlet
    # pull from clamped sources inside it
    seq = workloads.flapping_plate(steps + 1, center=(shape[0] - 3.3, 1.3, 11.6), span=8.0, chord=6.0, thickness=3.0,
                                   heave=0.0, pitch_deg=0.0, half_thickness=0.5)
    say(f"case: static box across the periodic y face at the +x outlet, {shape}, TRT, {steps} steps, exact roots")
    run_pair(p, seq, steps, check_every=3)


def test_flapping_plate_matches_oracle():
    steps = 20
    p = cases.plate_params()
    seq = workloads.flapping_plate(steps + 1, center=(24.0, 24.0, 24.0), span=20.0, chord=8.0, thickness=2.0,
                                   heave=2.0, pitch_deg=20.0, period=200.0, half_thickness=0.5)
    say(f"case: flapping plate {p.nx}x{p.ny}x{p.nz}, RR3, {steps} steps, exact roots vs oracle")
    run_pair(p, seq, steps, check_every=5)


def test_synthetic_bird_matches_oracle():
    steps = 3
    shape = (256, 128, 128)
    s = shape[0] / 512.0
    center = (160.0 * s, 128.0 * s, 128.0 * s)
    seq = workloads.synthetic_bird(steps + 1, seed=7, center=center, body_radii=(40.0 * s, 14.0 * s, 14.0 * s),
                                   half_span=100.0 * s, root_chord=40.0 * s, tip_chord=12.0 * s,
                                   tail=(30.0 * s, 20.0 * s), period=1200.0 * s)
    tau = 0.52
    p = bfref.make_params(shape, omega_plus=1.0 / tau, omega_minus=1.0 / (0.5 + (3.0 / 16.0) / (tau - 0.5)),
                          bc=((bfref.BC_FARFIELD, bfref.BC_OUTLET), (bfref.BC_PERIODIC, bfref.BC_PERIODIC),
                              (bfref.BC_PERIODIC, bfref.BC_PERIODIC)),
                          collision=bfref.COLLISION_RR3, wall_rule=bfref.WALL_INTERPOLATED_SDF, force_mode=6,
                          accumulate_loads=True, far_u=(0.04, 0.0, 0.0), body_position=center)
    say(f"case: synthetic bird {shape}, {len(seq.triangles)} triangles, RR3, {steps} steps, exact roots vs oracle")
    run_pair(p, seq, steps, check_every=1)


def test_articulated_bird_matches_oracle():
    # the exact-root rule on an articulated mesh (kinematics on the device):
    # its rigid joint's triangles are static, so the pass merges their cached
    # entries and walks only the wings; bit for bit against the oracle
    steps = 3
    shape = (128, 64, 64)
    s = shape[0] / 512.0
    center = (160.0 * s, 128.0 * s, 128.0 * s)
    mesh = workloads.synthetic_bird_articulated(seed=7, center=center, body_radii=(40.0 * s, 14.0 * s, 14.0 * s),
                                                half_span=100.0 * s, root_chord=40.0 * s, tip_chord=12.0 * s,
                                                tail=(30.0 * s, 20.0 * s), period=1200.0 * s)
    tau = 0.52
    p = bfref.make_params(shape, omega_plus=1.0 / tau, omega_minus=1.0 / (0.5 + (3.0 / 16.0) / (tau - 0.5)),
                          bc=((bfref.BC_FARFIELD, bfref.BC_OUTLET), (bfref.BC_PERIODIC, bfref.BC_PERIODIC),
                              (bfref.BC_PERIODIC, bfref.BC_PERIODIC)),
                          collision=bfref.COLLISION_RR3, wall_rule=bfref.WALL_INTERPOLATED_SDF, force_mode=6,
                          accumulate_loads=True, far_u=(0.04, 0.0, 0.0), body_position=center)
    k = cases.oracle_kinematics(mesh)
    say(f"case: articulated bird {shape}, {len(mesh.triangles)} triangles "
        f"({int((mesh.vertex_joint[mesh.triangles] == 0).all(axis=1).sum())} static), RR3, {steps} steps, exact roots")

    def geometry(t):
        pos, vel = bfref.articulated_vertices(mesh.rest, mesh.vertex_joint, mesh.span_fraction, k, t)
        mask, wall, _ = bfref.raster_resolve(p, pos, vel, mesh.triangles, mesh.parts, mesh.half_thickness, 64)
        return mask, wall, bfref.ray_links(p, pos, mesh.triangles, mesh.half_thickness, mask, wall)

    pl = links_params(p)
    mask0, wall0, _ = geometry(0.0)
    f = bfref.init_equilibrium(p, mask0, wall0, np.float32)
    prev = mask0
    loads_o = []
    with bh.BirdFlowSimulation(ray_config(p)) as sim:
        sim.upload_articulated(mesh, 0.0)
        sim.init_equilibrium()
        for st in range(1, steps + 1):
            sim.advance(1)
            mask, wall, links = geometry(float(st))
            f, ld, _, _ = bfref.step_with_links(pl, f, prev, mask, wall, links)
            loads_o.append(ld)
            prev = mask
            assert np.array_equal(sim.read_geometry()[0], mask), st
            compare_tables(f"step {st}", sim.ray_links(), links, mask)
        g = sim.read_populations()
        loads = sim.load_history(steps)
    loads_o = np.array(loads_o)
    differing = int((g.view(np.uint32) != f.view(np.uint32)).sum())
    rel = float(np.abs(loads[:, :6] - loads_o[:, :6]).max() / np.abs(loads_o[:, :6]).max())
    say(f"  after {steps} steps: populations differing {differing}, loads max rel err {rel:.2e}")
    assert differing == 0 and rel <= 1e-12


def test_offset_storage_with_exact_roots_tracks_fp32():
    # FP16 offsets under the exact-root rule: the same tables (they depend
    # on the geometry only) and a force history within 0.5 % of the FP32
    # run's largest force over 20 steps of the flapping plate
    steps = 20
    p = cases.plate_params()
    seq = workloads.flapping_plate(steps + 1, center=(24.0, 24.0, 24.0), span=20.0, chord=8.0, thickness=2.0,
                                   heave=2.0, pitch_deg=20.0, period=200.0, half_thickness=0.5)
    say(f"case: flapping plate, exact roots, FP16 offsets against FP32, {steps} steps")
    out = {}
    for storage in (bh.STORAGE_FP32, bh.STORAGE_FP16_OFFSET):
        with bh.BirdFlowSimulation(ray_config(p, storage)) as sim:
            sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
            sim.init_equilibrium()
            sim.advance(steps)
            out[storage] = (sim.ray_links(), sim.load_history(steps))
    t32, l32 = out[bh.STORAGE_FP32]
    t16, l16 = out[bh.STORAGE_FP16_OFFSET]
    force = float(np.abs(l16[:, :3] - l32[:, :3]).max() / np.abs(l32[:, :3]).max())
    say(f"  tables identical: {np.array_equal(t16.view(np.uint32), t32.view(np.uint32))}; force history max diff / "
        f"max force {force:.2e}")
    assert np.array_equal(t16.view(np.uint32), t32.view(np.uint32))
    assert force <= 5e-3


def test_ray_rule_refusals():
    p = cases.plate_params(shape=(32, 16, 16))
    with bh.BirdFlowSimulation(ray_config(p)) as sim:
        with pytest.raises(bh.BirdFlowError):
            sim.advance(1)                   # no mesh
        with pytest.raises(bh.BirdFlowError):
            sim.ray_links()
