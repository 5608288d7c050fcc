"""On-device kinematics (bf_mesh_upload_articulated) against the oracle
(bfo_articulated_vertices, pinned by tests/test_kinematics_oracle.py):

* the prepared vertices (bf_mesh_vertices_at) bit-identical at times across
  more than a period;
* the synthetic bird of the 512 x 256 x 256 workload uploaded articulated
  (no frames): masks bit-identical, walls bit-identical where the step reads
  them, link census exact and populations bit-identical for 3 steps, strict
  FP32, RR3, interpolated moving walls; its device memory is the rest pose,
  not 2,400 frames;
* (tests/test_gpu_checkpoint.py: an archive of an articulated mesh
  restores and continues bit for bit);
* the adjoint's forward pass over an articulated mesh is the ordinary step
  (its gradients against FP64 differences: tests/test_gpu_adjoint.py).
"""
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


def oracle_kinematics(mesh):
    return bfref.make_kinematics(mesh.period, mesh.amplitude_scale, mesh.center,
                                 [dict(side=j.side, root=j.root, stroke=j.stroke, deviation=j.deviation,
                                       pitch=j.pitch, twist=j.twist) for j in mesh.joints],
                                 harmonics=mesh.harmonics)


def bird_params(shape):
    s = shape[0] / 512.0
    center = (160.0 * s, 128.0 * s, 128.0 * s)
    return bfref.make_params(shape, omega_plus=1.0 / TAU, omega_minus=1.0 / (0.5 + (3.0 / 16.0) / (TAU - 0.5)),
                             bc=((bfref.BC_FARFIELD, bfref.BC_OUTLET), (bfref.BC_PERIODIC, bfref.BC_PERIODIC),
                                 (bfref.BC_PERIODIC, bfref.BC_PERIODIC)),
                             collision=bfref.COLLISION_RR3, wall_rule=bfref.WALL_INTERPOLATED_SDF, force_mode=6,
                             accumulate_loads=True, far_u=(0.04, 0.0, 0.0), body_position=center), center


def oracle_geometry(p, mesh, k, t):
    pos, vel = bfref.articulated_vertices(mesh.rest, mesh.vertex_joint, mesh.span_fraction, k, t)
    mask, wall, _ = bfref.raster_resolve(p, pos, vel, mesh.triangles, mesh.parts, mesh.half_thickness, key_bits=64)
    return mask, wall


def test_prepared_vertices_match_oracle():
    shape = (512, 256, 256)
    p, center = bird_params(shape)
    mesh = workloads.synthetic_bird_articulated(seed=7, center=center)
    k = oracle_kinematics(mesh)
    V = len(mesh.rest)
    say(f"case: articulated synthetic bird, {V} vertices, {len(mesh.triangles)} triangles, prepared vertices vs oracle")
    with bh.BirdFlowSimulation(cases.solver_config(p)) as sim:
        sim.upload_articulated(mesh)
        for t in (0.0, 1.0, 333.0, 1199.0, 1200.0, 2399.5, 5001.0):
            pos, vel = sim.mesh_vertices_at(t, V)
            po, vo = bfref.articulated_vertices(mesh.rest, mesh.vertex_joint, mesh.span_fraction, k, t)
            dp = int((pos.view(np.uint32) != po.view(np.uint32)).sum())
            dv = int((vel.view(np.uint32) != vo.view(np.uint32)).sum())
            say(f"  t {t}: positions differing {dp}, velocities differing {dv}, max wing speed "
                f"{float(np.linalg.norm(vo, axis=1).max()):.4f}")
            assert dp == 0 and dv == 0


def test_articulated_bird_steps_match_oracle():
    steps = 3
    shape = (512, 256, 256)
    p, center = bird_params(shape)
    mesh = workloads.synthetic_bird_articulated(seed=7, center=center)
    k = oracle_kinematics(mesh)
    say(f"case: articulated synthetic bird {shape}, RR3, strict FP32, {steps} steps vs oracle")
    with bh.BirdFlowSimulation(cases.solver_config(p)) as sim:
        before = sim.allocated_bytes
        sim.upload_articulated(mesh)
        mesh_bytes = sim.allocated_bytes - before
        sim.init_equilibrium()
        mask0, wall0 = oracle_geometry(p, mesh, k, 0.0)
        f = bfref.init_equilibrium(p, mask0, wall0, np.float32)
        prev = mask0
        for s in range(1, steps + 1):
            sim.advance(1)
            mask_g, wall_g = sim.read_geometry()
            census_g, cov_g, unc_g = sim.link_census()
            mask_o, wall_o = oracle_geometry(p, mesh, k, float(s))
            census_o, cov_o, unc_o = bfref.link_census(p, prev, mask_o)
            f, ld, _, _ = bfref.step(p, f, prev, mask_o, wall_o)
            g = sim.read_populations()
            band = wall_o[:, 3] <= 1.5
            diff = int((g.view(np.uint32) != f.view(np.uint32)).sum())
            say(f"  step {s}: mask equal {np.array_equal(mask_g, mask_o)}, walls in band equal "
                f"{np.array_equal(wall_g[band].view(np.uint32), wall_o[band].view(np.uint32))}, "
                f"links {int(census_g.sum())}, covered {cov_g}/{cov_o}, uncovered {unc_g}/{unc_o}, "
                f"populations differing {diff}")
            assert np.array_equal(mask_g, mask_o)
            assert np.array_equal(wall_g[band].view(np.uint32), wall_o[band].view(np.uint32))
            assert np.array_equal(census_g.astype(np.int64), census_o.astype(np.int64))
            assert (cov_g, unc_g) == (cov_o, unc_o)
            assert diff == 0
            prev = mask_o
    # the mesh's device memory: the key grid (8 B per cell) and per-vertex
    # arrays, no frames (2,400 steps of frames would add 590 MB)
    V = len(mesh.rest)
    frames_bytes = 2401 * V * 12
    say(f"  device memory of the articulated mesh {mesh_bytes / 1e6:.1f} MB (keys {8 * p.nx * p.ny * p.nz / 1e6:.1f} "
        f"MB); 2,400 frames would add {frames_bytes / 1e6:.0f} MB")
    assert mesh_bytes <= 8 * p.nx * p.ny * p.nz + 64 * V + 20 * len(mesh.triangles) + (1 << 20)


def test_adjoint_forward_pass_over_articulated_mesh_is_the_ordinary_step():
    # the adjoint's forward pass steps the articulated mesh as bf_step does:
    # its objective equals the weighted loads of a plain run, bit for bit
    # (its gradients: tests/test_gpu_adjoint.py)
    shape = (64, 32, 32)
    steps = 6
    p, center = bird_params(shape)
    mesh = workloads.synthetic_bird_articulated(seed=7, center=center, body_radii=(5.0, 2.0, 2.0), half_span=12.0,
                                                root_chord=5.0, tip_chord=2.0, tail=(4.0, 3.0), period=100.0,
                                                body_segments=(12, 8), wing_segments=(12, 4), tail_segments=(6, 3))
    weights = np.random.default_rng(3).uniform(-1.0, 1.0, (steps, 6))
    with bh.BirdFlowSimulation(cases.solver_config(p)) as sim:
        sim.upload_articulated(mesh)
        sim.init_equilibrium()
        sim.advance(steps)
        loads = sim.load_history()
    with bh.BirdFlowSimulation(cases.solver_config(p)) as sim:
        sim.upload_articulated(mesh)
        sim.adjoint_set_kinematic_parameters([bh.KinematicParameter((0, 1), "stroke", 0, "sin")])
        sim.init_equilibrium()
        J, grad, _ = sim.adjoint_gradient(steps, weights, checkpoint_interval=4)
    want = 0.0                      # the library's summation order
    for i in range(steps):
        for k in range(6):
            want += float(weights[i, k]) * float(loads[i, k])
    say(f"case: adjoint forward pass over an articulated bird {shape}: J {J!r}, plain run {want!r}, grad {grad}")
    assert J == want and grad.size == 1 and np.isfinite(grad).all()
