"""The oracle's articulated-mesh kinematics (bfo_articulated_vertices, the
on-device kinematics of bf_mesh_upload_articulated), pinned independently of
its own formula:

* against the synthetic bird's frames, which numpy builds in FP64 from
  rotation matrices (workloads.SyntheticBird.frame): every vertex within
  2e-4 cells at times across more than one period;
* velocities against FP64 central differences of those frames;
* closed forms: a pure stroke of amplitude A turns a span-axis vertex on the
  circle (0, s y cos(s a), s y sin(s a)); a pure pitch turns a chord-axis
  vertex in the x-z plane; twist grows linearly with the span fraction;
* the rigid vertices sit at centre + rest exactly, and t and t + period give
  the same pose;
* the FP64 form (articulated_vertices_f64, the finite-difference reference
  of the articulated adjoint) against the same FP64 frames to 5e-6 cells
  (the FP32 rounding of the mesh's rest positions), its
  velocities against their central differences, and its response to a
  perturbed Fourier coefficient (ArticulatedMesh.perturbed) against the
  frames rebuilt with the same coefficient perturbed.
"""
import math

import numpy as np
import pytest

import bfref
from birdflow_hip import workloads
from birdflow_hip.simulation import ArticulatedMesh, Joint, KinematicParameter

import cases


def oracle_kinematics(mesh: ArticulatedMesh):
    return bfref.make_kinematics(mesh.period, mesh.amplitude_scale, mesh.center,
                                 [dict(side=j.side, root=j.root, stroke=j.stroke, deviation=j.deviation,
                                       pitch=j.pitch, twist=j.twist) for j in mesh.joints],
                                 harmonics=mesh.harmonics)


def small_bird():
    bird = workloads.SyntheticBird(seed=7, center=(40.0, 32.0, 32.0), body_radii=(10.0, 3.5, 3.5), half_span=25.0,
                                   root_chord=10.0, tip_chord=3.0, tail=(7.5, 5.0), period=300.0,
                                   body_segments=(24, 16), wing_segments=(40, 10), tail_segments=(12, 6))
    bird.limit_mach(0.04, 0.15)
    return bird


def test_articulated_bird_matches_fp64_frames():
    bird = small_bird()
    mesh = bird.articulated()
    k = oracle_kinematics(mesh)
    worst = 0.0
    for t in (0.0, 17.0, 123.0, 299.0, 300.0, 451.5, 1003.0):
        pos, _ = bfref.articulated_vertices(mesh.rest, mesh.vertex_joint, mesh.span_fraction, k, t)
        want = bird.frame(t)
        worst = max(worst, float(np.abs(pos - want).max()))
    print(f"max |oracle - FP64 frame| = {worst:.2e} cells")
    assert worst <= 2e-4


def test_articulated_velocities_are_the_time_derivative():
    bird = small_bird()
    mesh = bird.articulated()
    k = oracle_kinematics(mesh)
    eps = 1e-3
    for t in (5.0, 77.0, 260.0):
        _, vel = bfref.articulated_vertices(mesh.rest, mesh.vertex_joint, mesh.span_fraction, k, t, velocity_scale=1.0)
        fd = (bird.frame(t + eps) - bird.frame(t - eps)) / (2 * eps)
        scale = float(np.abs(fd).max())
        err = float(np.abs(vel - fd).max())
        print(f"t {t}: max |v - central difference| = {err:.2e} (max speed {scale:.3e})")
        assert scale > 0.01
        assert err <= 2e-3 * scale
    # the velocity scale multiplies
    _, v1 = bfref.articulated_vertices(mesh.rest, mesh.vertex_joint, mesh.span_fraction, k, 5.0, velocity_scale=1.0)
    _, v2 = bfref.articulated_vertices(mesh.rest, mesh.vertex_joint, mesh.span_fraction, k, 5.0, velocity_scale=0.5)
    assert np.array_equal(v2, (v1 * np.float32(0.5)).astype(np.float32))


def single_joint(**series):
    zero = np.zeros((1, 2), np.float32)
    j = Joint(series.pop("side", 1.0), (0.0, 0.0, 0.0), series.get("stroke", zero), series.get("deviation", zero),
              series.get("pitch", zero), series.get("twist", zero))
    return j


def run_one(joint, rest, span, t, period=100.0):
    rest = np.asarray(rest, np.float32).reshape(-1, 3)
    mesh = ArticulatedMesh(rest, np.ones(len(rest), np.uint8), np.asarray(span, np.float32),
                           np.zeros((0, 3), np.uint32), np.zeros(0, np.uint8), 1.0, period, 1.0, (0.0, 0.0, 0.0),
                           [joint], harmonics=1)
    return bfref.articulated_vertices(rest, mesh.vertex_joint, mesh.span_fraction, oracle_kinematics(mesh), t)


@pytest.mark.parametrize("side", [1.0, -1.0])
def test_closed_forms(side):
    A = 0.4
    t = 13.0
    a = A * math.sin(2 * math.pi * t / 100.0)
    adot = A * math.cos(2 * math.pi * t / 100.0) * 2 * math.pi / 100.0
    # pure stroke (about x) of a span-axis vertex
    pos, vel = run_one(single_joint(side=side, stroke=np.array([[A, 0.0]])), [(0.0, 10.0, 0.0)], [0.0], t)
    want = np.array([0.0, side * 10.0 * math.cos(side * a), side * 10.0 * math.sin(side * a)])
    assert np.abs(pos[0] - want).max() <= 2e-5
    dwant = np.array([0.0, -10.0 * math.sin(side * a) * adot, 10.0 * math.cos(side * a) * adot])
    assert np.abs(vel[0] - dwant).max() <= 2e-6
    # pure pitch (about y) of a chord-axis vertex: (x cos a, 0, -x sin a)
    pos, _ = run_one(single_joint(side=side, pitch=np.array([[A, 0.0]])), [(4.0, 0.0, 0.0)], [0.0], t)
    assert np.abs(pos[0] - [4.0 * math.cos(a), 0.0, -4.0 * math.sin(a)]).max() <= 1e-5
    # twist scales with the span fraction: at fraction 0.5 it is half a pitch
    p1, _ = run_one(single_joint(side=side, twist=np.array([[A, 0.0]])), [(4.0, 0.0, 0.0)], [0.5], t)
    p2, _ = run_one(single_joint(side=side, pitch=np.array([[0.5 * A, 0.0]])), [(4.0, 0.0, 0.0)], [0.0], t)
    assert np.abs(p1 - p2).max() <= 1e-6
    # deviation (about z) of a span-axis vertex: (-s y sin(s b), s y cos(s b), 0)
    pos, _ = run_one(single_joint(side=side, deviation=np.array([[0.0, A]])), [(0.0, 10.0, 0.0)], [0.0], t)
    b = A * math.cos(2 * math.pi * t / 100.0)
    assert np.abs(pos[0] - [-side * 10.0 * math.sin(side * b), side * 10.0 * math.cos(side * b), 0.0]).max() <= 2e-5


def test_rigid_vertices_and_periodicity():
    bird = small_bird()
    mesh = bird.articulated()
    k = oracle_kinematics(mesh)
    pos, vel = bfref.articulated_vertices(mesh.rest, mesh.vertex_joint, mesh.span_fraction, k, 42.0)
    rigid = mesh.vertex_joint == 0
    center = np.asarray(mesh.center, np.float32)
    assert np.array_equal(pos[rigid], (center + mesh.rest[rigid]).astype(np.float32))
    assert not vel[rigid].any()
    p2, _ = bfref.articulated_vertices(mesh.rest, mesh.vertex_joint, mesh.span_fraction, k, 42.0 + mesh.period)
    assert np.abs(p2 - pos).max() <= 1e-4
    # mirror symmetry of the wings: the kinematics of the two sides differ
    # (drawn separately), so mirror one side's series onto the other
    j = mesh.joints
    mirrored = ArticulatedMesh(mesh.rest, mesh.vertex_joint, mesh.span_fraction, mesh.triangles, mesh.parts,
                               mesh.half_thickness, mesh.period, mesh.amplitude_scale, mesh.center,
                               [j[0], Joint(-1.0, j[0].root, j[0].stroke, j[0].deviation, j[0].pitch, j[0].twist)])
    pm, _ = bfref.articulated_vertices(mesh.rest, mesh.vertex_joint, mesh.span_fraction, oracle_kinematics(mirrored),
                                       42.0)
    left, right = pm[mesh.vertex_joint == 1], pm[mesh.vertex_joint == 2]
    assert np.abs(left[:, [0, 2]] - right[:, [0, 2]]).max() <= 1e-4
    assert np.abs((left[:, 1] - center[1]) + (right[:, 1] - center[1])).max() <= 1e-4


def test_fp64_kinematics_match_the_fp64_frames():
    bird = small_bird()
    mesh = bird.articulated()
    k64 = cases.kinematics_f64(mesh)
    worst = 0.0
    for t in (0.0, 17.0, 123.0, 299.0, 451.5):
        pos, vel = bfref.articulated_vertices_f64(mesh.rest, mesh.vertex_joint, mesh.span_fraction, k64, t)
        worst = max(worst, float(np.abs(pos - bird.frame(t)).max()))
        eps = 1e-3
        fd = (bird.frame(t + eps) - bird.frame(t - eps)) / (2 * eps)
        assert np.abs(vel - fd).max() <= 1e-6 * np.abs(fd).max()
        # and the FP32 oracle within its rounding
        p32, v32 = bfref.articulated_vertices(mesh.rest, mesh.vertex_joint, mesh.span_fraction,
                                              oracle_kinematics(mesh), t)
        assert np.abs(p32 - pos).max() <= 2e-4
        assert np.abs(v32 - vel).max() <= 2e-3 * np.abs(vel).max()
    # the mesh holds its rest positions and span fractions in FP32 (up to
    # 40 cells: 2.4e-6 per coordinate), the frames in FP64
    print(f"max |FP64 kinematics - FP64 frames| = {worst:.2e} cells")
    assert worst <= 5e-6


@pytest.mark.parametrize("channel,harmonic,term,joints", [("stroke", 0, "sin", (0, 1)), ("pitch", 1, "cos", (1,)),
                                                          ("twist", 2, "sin", (0,)), ("deviation", 0, "cos", (0, 1))])
def test_perturbed_coefficient_moves_the_mesh_like_the_frames(channel, harmonic, term, joints):
    # d(vertex)/d(coefficient) by central differences of the FP64 kinematics
    # (the adjoint's reference) and of the FP64 frames with the bird's own
    # coefficient perturbed in every named wing
    bird = small_bird()
    mesh = bird.articulated()
    q = KinematicParameter(joints, channel, harmonic, term)
    h = 1e-6
    t = 77.0
    series = getattr(bird.kin, channel)
    comp = 0 if term == "sin" else 1
    d_mesh, d_frame = [], []
    for sign in (1.0, -1.0):
        m = mesh.perturbed([q], [sign * h])
        pos, _ = bfref.articulated_vertices_f64(m.rest, m.vertex_joint, m.span_fraction, cases.kinematics_f64(m), t)
        d_mesh.append(pos)
        saved = series.copy()
        for w in joints:
            series[w, harmonic, comp] += sign * h
        d_frame.append(bird.frame(t))
        series[...] = saved
    dm = (d_mesh[0] - d_mesh[1]) / (2 * h)
    df = (d_frame[0] - d_frame[1]) / (2 * h)
    moved = np.abs(df).max()
    print(f"{channel} h{harmonic} {term} joints {joints}: max |d vertex / d c| = {moved:.3e}, "
          f"difference {np.abs(dm - df).max():.2e}")
    assert moved > 0.1
    assert np.abs(dm - df).max() <= 1e-6 * moved
    # the joints not named stay put
    other = [j for j in (0, 1) if j not in joints]
    for j in other:
        assert not dm[mesh.vertex_joint == j + 1].any()
