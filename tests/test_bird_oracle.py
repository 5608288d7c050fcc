"""Pins the oracle's analytic signed-distance bird (BirdFlow.metal:486-533,
611-938) with closed forms and invariants, none of which retypes the
oracle's expressions:

  * sin / cos against FP64 numpy (<= 2 ulp where |value| > 1e-3, 1e-7
    absolute everywhere) and exact at 0;
  * wing frames at rest are the body axes (span +-y, chord x, normal +-z),
    the roots sit at the body position plus the mirrored root offset;
  * the stroke rotates the left span to (0, cos theta, sin theta) and the
    right one to (0, -cos theta, sin theta), theta = bias + A sin(2 pi f t);
  * chord, span, normal are orthonormal and right-handed for any pose;
  * the wing angular velocity reproduces central differences of the frames;
  * the grid radius is the closed-form bound plus two cells;
  * occupancy of an isolated ellipsoid body equals the implicit inside
    test away from the surface;
  * body-cell wall velocities are the rigid-body field v + w x r (scaled),
    wing cells add the wing's relative rotation about its root;
  * a symmetric pose gives a mask mirrored across the body x-z plane with
    the wing part ids swapped;
  * a cell solid in the previous mask is evaluated even beyond the radius;
  * measured keyframes (BirdFlow.metal:550-609): keyframes sampled from the
    sinusoid reproduce the analytic frame, keyframe values are hit at their
    phases, a deviation rotates span and chord about the stroke normal in
    closed form, and the angular velocities (deviation and tip-twist rates
    included) match frame differences.
"""
import math

import numpy as np
import pytest

import bfref

DEMO = dict(body_radii=(0.18, 0.065, 0.075), wing_span=0.34, wing_root_chord=0.14, wing_tip_chord=0.055,
            wing_thickness=0.018, wing_sweep=0.055, wing_root_offset=(0.02, 0.055, 0.02), tail_length=0.17,
            tail_half_width=0.11, tail_thickness=0.018, frequency=4.5, stroke_amplitude=math.radians(42),
            stroke_bias=math.radians(2), pitch_mean=math.radians(8), pitch_amplitude=math.radians(24),
            pitch_phase=math.radians(78))
COMPACT = dict(body_radii=(0.015, 0.011, 0.012), wing_span=0.025, wing_root_chord=0.020, wing_tip_chord=0.012,
               wing_thickness=0.008, wing_sweep=0.004, wing_root_offset=(0.0, 0.010, 0.003), tail_length=0.020,
               tail_half_width=0.014, tail_thickness=0.008, frequency=3.0, stroke_amplitude=0.25, stroke_bias=0.0,
               pitch_mean=0.0, pitch_amplitude=0.15, pitch_phase=0.4)
FIELDS = {name: i for i, name in enumerate(bfref.BIRD_FRAME_FIELDS)}


def row(frame, name):
    return frame[FIELDS[name]].astype(np.float64)


def test_sincos_against_fp64():
    rng = np.random.default_rng(1)
    x = np.concatenate([np.linspace(-400, 400, 400001), rng.uniform(-1e3, 1e3, 100000),
                        np.arange(-64, 65) * (math.pi / 4)]).astype(np.float32)
    s, c = bfref.sincos(x)
    xd = x.astype(np.float64)
    for got, want in ((s, np.sin(xd)), (c, np.cos(xd))):
        err = np.abs(got.astype(np.float64) - want)
        assert err.max() <= 1e-7
        big = np.abs(want) > 1e-3
        ulp = np.spacing(np.abs(want[big]).astype(np.float32)).astype(np.float64)
        assert (err[big] / ulp).max() <= 2.0
    s0, c0 = bfref.sincos([0.0])
    assert s0[0] == 0.0 and c0[0] == 1.0


def test_wing_frames_at_rest_are_the_body_axes():
    b = bfref.make_bird(**dict(COMPACT, stroke_amplitude=0.0, pitch_amplitude=0.0))
    pos = np.array([0.3, 0.2, 0.1])
    f = bfref.bird_prepare(b, pos, time=0.123, dx=0.002)
    off = np.array(COMPACT["wing_root_offset"])
    for side, name in ((1, "left"), (-1, "right")):
        assert np.allclose(row(f, f"{name}_span")[:3], (0, side, 0), atol=1e-7)
        assert np.allclose(row(f, f"{name}_chord")[:3], (1, 0, 0), atol=1e-7)
        assert np.allclose(row(f, f"{name}_normal")[:3], (0, 0, side), atol=1e-7)
        assert np.allclose(row(f, f"{name}_root")[:3], pos + off * (1, side, 1), atol=1e-7)
        assert np.allclose(row(f, f"{name}_angular"), 0.0, atol=0)
    assert np.allclose(row(f, "body_position")[:3], pos, atol=1e-7)
    assert np.array_equal(row(f, "orientation"), [0, 0, 0, 1])


@pytest.mark.parametrize("t", [0.0, 0.01, 0.037, 0.1, 0.29])
def test_stroke_rotates_the_spans_about_the_body_x_axis(t):
    b = bfref.make_bird(**dict(DEMO, pitch_mean=0.0, pitch_amplitude=0.0))
    f = bfref.bird_prepare(b, (0, 0, 0), time=t, dx=0.005)
    theta = DEMO["stroke_bias"] + DEMO["stroke_amplitude"] * math.sin(2 * math.pi * DEMO["frequency"] * t)
    rate = 2 * math.pi * DEMO["frequency"] * DEMO["stroke_amplitude"] * math.cos(2 * math.pi * DEMO["frequency"] * t)
    assert np.allclose(row(f, "left_span")[:3], (0, math.cos(theta), math.sin(theta)), atol=2e-6)
    assert np.allclose(row(f, "right_span")[:3], (0, -math.cos(theta), math.sin(theta)), atol=2e-6)
    assert np.allclose(row(f, "left_normal")[:3], (0, -math.sin(theta), math.cos(theta)), atol=2e-6)
    assert np.allclose(row(f, "left_angular")[:3], (rate, 0, 0), rtol=2e-6, atol=2e-6)
    assert np.allclose(row(f, "right_angular")[:3], (-rate, 0, 0), rtol=2e-6, atol=2e-6)


def random_pose(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    return rng.uniform(-1, 1, 3), q, rng.normal(size=3), rng.normal(size=3)


def test_frames_are_orthonormal_and_right_handed():
    rng = np.random.default_rng(7)
    b = bfref.make_bird(**DEMO)
    for _ in range(50):
        pos, q, v, w = random_pose(rng)
        f = bfref.bird_prepare(b, pos, q * rng.uniform(0.5, 2.0), v, w, time=rng.uniform(0, 2), dx=0.004)
        assert abs(np.linalg.norm(row(f, "orientation")) - 1) < 1e-6
        for name in ("left", "right"):
            c, s, n = (row(f, f"{name}_{k}")[:3] for k in ("chord", "span", "normal"))
            for a in (c, s, n):
                assert abs(np.linalg.norm(a) - 1) < 2e-6
            assert abs(c @ s) < 2e-6 and abs(c @ n) < 2e-6 and abs(s @ n) < 2e-6
            assert np.allclose(np.cross(c, s), n, atol=4e-6)


def test_wing_angular_velocity_matches_frame_differences():
    rng = np.random.default_rng(3)
    b = bfref.make_bird(**DEMO)
    pos, q, _, _ = random_pose(rng)
    h = 2e-4
    for t in (0.013, 0.071, 0.18):
        fp = bfref.bird_prepare(b, pos, q, time=t + h, dx=0.004)
        fm = bfref.bird_prepare(b, pos, q, time=t - h, dx=0.004)
        f = bfref.bird_prepare(b, pos, q, time=t, dx=0.004)
        for name in ("left", "right"):
            w = row(f, f"{name}_angular")[:3]
            for k in ("chord", "span", "normal"):
                d = (row(fp, f"{name}_{k}")[:3] - row(fm, f"{name}_{k}")[:3]) / (2 * h)
                want = np.cross(w, row(f, f"{name}_{k}")[:3])
                assert np.linalg.norm(d - want) <= 2e-3 * np.linalg.norm(w) + 1e-3


def test_geometry_radius_is_the_closed_form_bound():
    b = bfref.make_bird(**DEMO)
    dx = 0.0037
    f = bfref.bird_prepare(b, (0, 0, 0), time=0.0, dx=dx)
    r = np.array(DEMO["body_radii"])
    body = np.linalg.norm(r)
    tail = np.linalg.norm([r[0] + DEMO["tail_length"], max(r[1], DEMO["tail_half_width"]),
                           r[2] + 0.5 * DEMO["tail_thickness"]])
    wing = (np.linalg.norm(DEMO["wing_root_offset"]) + DEMO["wing_span"] + abs(DEMO["wing_sweep"])
            + 0.5 * max(DEMO["wing_root_chord"], DEMO["wing_tip_chord"]) + 0.5 * DEMO["wing_thickness"])
    assert abs(row(f, "omega_world_radius")[3] - (max(body, tail, wing) + 2 * dx)) < 1e-6


def grid(shape, dx, origin=(0.0, 0.0, 0.0)):
    return bfref.make_params(shape, omega_plus=1.0, omega_minus=1.0, origin=origin, dx=dx)


def cell_centers(p):
    z, y, x = np.meshgrid(np.arange(p.nz), np.arange(p.ny), np.arange(p.nx), indexing="ij")
    o = np.array(p.origin, np.float64)
    return np.stack([o[0] + (x.ravel() + 0.5) * p.dx, o[1] + (y.ravel() + 0.5) * p.dx,
                     o[2] + (z.ravel() + 0.5) * p.dx], axis=1)


def test_isolated_ellipsoid_occupancy_is_the_implicit_inside_test():
    tiny = dict(COMPACT, wing_span=1e-4, wing_root_chord=1e-4, wing_tip_chord=1e-4, wing_thickness=1e-4,
                wing_sweep=0.0, wing_root_offset=(0.0, 0.0, 0.0), tail_length=1e-4, tail_half_width=1e-4,
                tail_thickness=1e-4)
    b = bfref.make_bird(**tiny)
    p = grid((40, 32, 32), 0.001)
    pos = np.array([0.0201, 0.0163, 0.0158])
    f = bfref.bird_prepare(b, pos, time=0.0, dx=p.dx)
    mask, _ = bfref.bird_build(p, b, f)
    rel = cell_centers(p) - pos
    F = ((rel / np.array(tiny["body_radii"])) ** 2).sum(axis=1)
    assert (mask[F < 0.999] == 1).all()
    assert (mask[F > 1.001] == 0).all()
    assert (mask != 0).sum() > 1000


def test_wall_velocity_is_the_rigid_field_plus_wing_rotation():
    rng = np.random.default_rng(11)
    b = bfref.make_bird(**COMPACT)
    p = grid((48, 48, 48), 0.002)
    pos = np.array([0.048, 0.048, 0.048])
    q = np.array([0.05, -0.08, 0.1, 1.0])
    v, w = np.array([1.0, -0.3, 0.2]), np.array([2.0, -1.0, 3.0])
    scale = 0.7
    f = bfref.bird_prepare(b, pos, q, v, w, time=0.05, dx=p.dx)
    mask, wall = bfref.bird_build(p, b, f, velocity_to_lattice=scale)
    centers = cell_centers(p)
    rel = centers - row(f, "body_position")[:3]
    omega = row(f, "omega_world_radius")[:3]
    rigid = v + np.cross(omega, rel)
    body = (mask == 1) | (mask == 4)
    assert body.sum() > 500
    assert np.allclose(wall[body, :3], rigid[body] * scale, rtol=1e-5, atol=1e-6)
    for part, name in ((2, "left"), (3, "right")):
        cells = mask == part
        assert cells.sum() > 100
        rel_root = centers[cells] - row(f, f"{name}_root")[:3]
        want = rigid[cells] + np.cross(row(f, f"{name}_angular")[:3], rel_root)
        assert np.allclose(wall[cells, :3], want * scale, rtol=1e-5, atol=1e-6)
    assert np.array_equal(wall[:, 3], np.zeros(len(wall), np.float32))
    # beyond the radius: fluid and zero velocity
    far = (np.linalg.norm(rel, axis=1) > row(f, "omega_world_radius")[3] * 1.001)
    assert (mask[far] == 0).all() and (wall[far] == 0).all()


def test_symmetric_pose_gives_a_mirrored_mask():
    b = bfref.make_bird(**COMPACT)
    p = grid((48, 48, 48), 0.002)
    pos = np.array([0.048, 0.048, 0.048])       # the y mid-plane between cells 23 and 24
    for t in (0.0, 0.04, 0.1):
        f = bfref.bird_prepare(b, pos, time=t, dx=p.dx)
        mask, wall = bfref.bird_build(p, b, f)
        m = mask.reshape(48, 48, 48)
        mirrored = m[:, ::-1, :].copy()
        swap = mirrored.copy()
        swap[mirrored == 2] = 3
        swap[mirrored == 3] = 2
        differ = int((swap != m).sum())
        solid = int((m != 0).sum())
        assert (m == 2).sum() > 100 and (m == 3).sum() > 100
        assert differ <= max(2, solid // 1000)
        wy = wall[:, 1].reshape(48, 48, 48)
        same = swap == m
        assert np.allclose(wy[same & (m != 0)], -wy[:, ::-1, :][same & (m != 0)], atol=1e-6)


def test_previously_solid_cell_beyond_the_radius_is_evaluated():
    b = bfref.make_bird(**COMPACT)
    p = grid((48, 48, 48), 0.002)
    pos = np.array([0.048, 0.048, 0.048])
    v = np.array([1.5, 0.0, 0.0])
    f = bfref.bird_prepare(b, pos, linear_velocity=v, time=0.0, dx=p.dx)
    prev = np.zeros(48 ** 3, np.uint8)
    gid = 1 + 48 * (1 + 48 * 1)           # a corner cell, far beyond the radius
    prev[gid] = 1
    m0, w0 = bfref.bird_build(p, b, f)
    m1, w1 = bfref.bird_build(p, b, f, solid_prev=prev)
    assert m1[gid] == 0 and w0[gid, 0] == 0.0 and w1[gid, 0] == np.float32(1.5)
    other = np.ones(48 ** 3, bool)
    other[gid] = False
    assert np.array_equal(m0[other], m1[other]) and np.array_equal(w0[other], w1[other])


# ---------------------------------------------------------------------------
# measured-keyframe kinematics (BirdFlow.metal:550-609)
# ---------------------------------------------------------------------------
def sinusoid_keyframes(cfg, count, deviation=(0.0, 0.0), twist=(0.0, 0.0)):
    """Keyframes sampling the analytic stroke / pitch sinusoids (exact rates),
    plus optional deviation and tip-twist sinusoids (amplitude, phase)."""
    f = cfg["frequency"]
    w = 2 * math.pi * f
    kf = np.zeros((count, 17))
    for k in range(count):
        ph = k / count
        t = ph / f
        stroke = cfg["stroke_bias"] + cfg["stroke_amplitude"] * math.sin(w * t)
        srate = w * cfg["stroke_amplitude"] * math.cos(w * t)
        pitch = cfg["pitch_mean"] + cfg["pitch_amplitude"] * math.sin(w * t + cfg["pitch_phase"])
        prate = w * cfg["pitch_amplitude"] * math.cos(w * t + cfg["pitch_phase"])
        dev = deviation[0] * math.sin(w * t + deviation[1])
        drate = w * deviation[0] * math.cos(w * t + deviation[1])
        tw = twist[0] * math.sin(w * t + twist[1])
        trate = w * twist[0] * math.cos(w * t + twist[1])
        side = [stroke, dev, pitch, tw, srate, drate, prate, trate]
        kf[k] = [ph] + side + side
    return kf.astype(np.float32)


def test_measured_keyframes_of_the_sinusoid_reproduce_the_analytic_frame():
    b = bfref.make_bird(**DEMO)
    kf = sinusoid_keyframes(DEMO, 96)
    rng = np.random.default_rng(5)
    pos, q, _, _ = random_pose(rng)
    worst = 0.0
    for t in rng.uniform(0.0, 2.0, 40):
        fa = bfref.bird_prepare(b, pos, q, time=t, dx=0.004)
        fm = bfref.bird_prepare(b, pos, q, time=t, dx=0.004, keyframes=kf)
        for name in ("left", "right"):
            for k in ("root", "chord", "span", "normal"):
                worst = max(worst, np.abs(row(fa, f"{name}_{k}")[:3] - row(fm, f"{name}_{k}")[:3]).max())
            wa, wm = row(fa, f"{name}_angular")[:3], row(fm, f"{name}_angular")[:3]
            assert np.linalg.norm(wa - wm) <= 2e-4 * max(1.0, np.linalg.norm(wa))
    assert worst < 2e-5


def test_measured_keyframe_values_are_hit_exactly_at_their_phases():
    b = bfref.make_bird(**dict(DEMO, frequency=4.0))
    kf = sinusoid_keyframes(dict(DEMO, frequency=4.0), 8, deviation=(0.2, 0.3), twist=(0.4, 1.1))
    for k in range(8):
        t = float(np.float32(kf[k, 0])) / 4.0
        f = bfref.bird_prepare(b, (0, 0, 0), time=t, dx=0.004, keyframes=kf)
        assert row(f, "left_chord")[3] == pytest.approx(float(kf[k, 4]), abs=2e-6)
        assert row(f, "right_chord")[3] == pytest.approx(-float(kf[k, 12]), abs=2e-6)
        assert row(f, "left_angular")[3] == pytest.approx(float(kf[k, 8]), rel=1e-4, abs=1e-4)


def test_deviation_rotates_span_and_chord_about_the_stroke_normal():
    b = bfref.make_bird(**DEMO)
    for dev in (-0.4, 0.15, 0.7):
        kf = np.zeros((4, 17), np.float32)
        kf[:, 0] = [0.0, 0.25, 0.5, 0.75]
        kf[:, 2] = dev
        kf[:, 10] = dev
        f = bfref.bird_prepare(b, (0, 0, 0), time=0.3, dx=0.004, keyframes=kf)
        c, s = math.cos(dev), math.sin(dev)
        assert np.allclose(row(f, "left_span")[:3], (-s, c, 0), atol=2e-6)
        assert np.allclose(row(f, "left_chord")[:3], (c, s, 0), atol=2e-6)
        # right side: rotation about its own stroke normal (-z)
        assert np.allclose(row(f, "right_span")[:3], (-s, -c, 0), atol=2e-6)
        assert np.allclose(row(f, "right_chord")[:3], (c, -s, 0), atol=2e-6)


def test_measured_wing_angular_velocity_matches_frame_differences():
    b = bfref.make_bird(**DEMO)
    kf = sinusoid_keyframes(DEMO, 64, deviation=(0.25, 0.7), twist=(0.3, 0.2))
    rng = np.random.default_rng(9)
    pos, q, _, _ = random_pose(rng)
    h = 2e-4
    for t in (0.021, 0.09, 0.16):
        fp = bfref.bird_prepare(b, pos, q, time=t + h, dx=0.004, keyframes=kf)
        fm = bfref.bird_prepare(b, pos, q, time=t - h, dx=0.004, keyframes=kf)
        f = bfref.bird_prepare(b, pos, q, time=t, dx=0.004, keyframes=kf)
        for name in ("left", "right"):
            w = row(f, f"{name}_angular")[:3]
            for k in ("chord", "span", "normal"):
                d = (row(fp, f"{name}_{k}")[:3] - row(fm, f"{name}_{k}")[:3]) / (2 * h)
                assert np.linalg.norm(d - np.cross(w, row(f, f"{name}_{k}")[:3])) <= 2e-3 * np.linalg.norm(w) + 1e-3
            # tip twist rate against differences of the tip twist
            dtw = (row(fp, f"{name}_chord")[3] - row(fm, f"{name}_chord")[3]) / (2 * h)
            assert abs(dtw - row(f, f"{name}_angular")[3]) <= 2e-3 * abs(dtw) + 1e-3


# ---------------------------------------------------------------------------
# Prescribed-wing hinge reaction (updateWingInertialReaction,
# BirdFlow.metal:6493-6572): closed forms on hand-built frames.
# ---------------------------------------------------------------------------
def wing_frame(theta, omega, root=(0.0, 0.0, 0.0), body=(0.0, 0.0, 0.0)):
    """A frame whose two wings turn about the z axis: span (cos t, sin t, 0),
    chord (-sin t, cos t, 0), normal z, angular velocity (0, 0, omega); the
    right wing mirrors the left across the x-z plane (span (cos t, -sin t, 0),
    angular velocity (0, 0, -omega))."""
    f = np.zeros((14, 4), np.float32)
    f[0, :3] = body
    c, s = math.cos(theta), math.sin(theta)
    f[4, :3], f[5, :3], f[6, :3], f[7, :3], f[8, :3] = root, (-s, c, 0.0), (c, s, 0.0), (0, 0, 1), (0, 0, omega)
    f[9, :3], f[10, :3], f[11, :3], f[12, :3], f[13, :3] = ((root[0], -root[1], root[2]), (-s, -c, 0.0), (c, -s, 0.0),
                                                            (0, 0, -1), (0, 0, -omega))
    return f


def test_wing_reaction_of_a_centred_isotropic_wing_is_the_spin_change():
    # centre of mass on the hinge: no linear momentum; isotropic inertia I:
    # angular momentum I w whatever the axes, so the reaction torque is
    # -I (w2 - w1) / dt and the force zero
    d = bfref.make_wing_dynamics(0.03, (0.0, 0.0, 0.0), (2e-4, 2e-4, 2e-4))
    dt = 0.002
    f1, f2 = wing_frame(0.3, 5.0), wing_frame(0.31, 7.0)
    _, prev = bfref.wing_reaction(f1, d, d, dt, np.zeros(16), initialize=True)
    r, _ = bfref.wing_reaction(f2, d, d, dt, prev)
    assert np.abs(r[0, :3]).max() == 0.0 and np.abs(r[2, :3]).max() == 0.0
    assert r[1, 2] == pytest.approx(-2e-4 * (7.0 - 5.0) / dt, rel=1e-5)
    assert r[3, 2] == pytest.approx(+2e-4 * (7.0 - 5.0) / dt, rel=1e-5)
    assert abs(r[1, 0]) + abs(r[1, 1]) <= 1e-9


@pytest.mark.parametrize("omega,dt", [(20.0, 0.001), (3.0, 0.004)])
def test_wing_reaction_of_a_steady_stroke_is_the_centrifugal_pull(omega, dt):
    # constant rate w about z, centre of mass r along the span, hinge at the
    # body position: P = m w r chord turns with the wing, so the reaction
    # -dP/dt = m w r (2 sin(w dt / 2) / dt) along the mean span (outward: the
    # wing pulls the hinge), and L = m w r^2 z + I_n w z is constant (zero
    # torque); the mirrored right wing gives the mirrored force
    m, r = 0.04, 0.12
    d = bfref.make_wing_dynamics(m, (0.0, r, 0.0), (1e-4, 2e-5, 3e-4))
    t1 = 0.4
    t2 = t1 + omega * dt
    f1, f2 = wing_frame(t1, omega), wing_frame(t2, omega)
    _, prev = bfref.wing_reaction(f1, d, d, dt, np.zeros(16), initialize=True)
    rxn, _ = bfref.wing_reaction(f2, d, d, dt, prev)
    mid = 0.5 * (t1 + t2)
    span_mid = np.array([math.cos(mid), math.sin(mid), 0.0])
    magnitude = m * omega * r * 2.0 * math.sin(omega * dt / 2.0) / dt
    force = rxn[0, :3].astype(np.float64)
    assert np.linalg.norm(force) == pytest.approx(magnitude, rel=2e-4)
    assert float(np.dot(force, span_mid)) == pytest.approx(magnitude, rel=2e-4)
    assert np.abs(rxn[1, :3]).max() <= 1e-4 * magnitude * r
    right = rxn[2, :3].astype(np.float64)
    assert right == pytest.approx(force * np.array([1.0, -1.0, 1.0]), rel=1e-4, abs=1e-6 * magnitude)


def test_wing_reaction_torque_of_an_offset_hinge():
    # hinge at distance a along x from the body position, centre of mass on
    # the hinge, no inertia, the wing translating with the body frame... the
    # momentum is zero; then the centre of mass r along the span with the
    # hinge offset: L = (a x + r span) x P adds a x P to the centred case,
    # so the torque reaction gains a x (force reaction)
    m, r, a = 0.02, 0.1, 0.05
    d = bfref.make_wing_dynamics(m, (0.0, r, 0.0), (0.0, 0.0, 0.0))
    omega, dt = 10.0, 0.002
    f1, f2 = wing_frame(0.2, omega, root=(a, 0.0, 0.0)), wing_frame(0.2 + omega * dt, omega, root=(a, 0.0, 0.0))
    g1, g2 = wing_frame(0.2, omega), wing_frame(0.2 + omega * dt, omega)
    _, p1 = bfref.wing_reaction(f1, d, d, dt, np.zeros(16), initialize=True)
    off, _ = bfref.wing_reaction(f2, d, d, dt, p1)
    _, q1 = bfref.wing_reaction(g1, d, d, dt, np.zeros(16), initialize=True)
    cen, _ = bfref.wing_reaction(g2, d, d, dt, q1)
    assert off[0, :3] == pytest.approx(cen[0, :3], rel=1e-5)
    extra = np.cross(np.array([a, 0.0, 0.0]), off[0, :3].astype(np.float64))
    assert (off[1, :3] - cen[1, :3]).astype(np.float64) == pytest.approx(extra, rel=1e-3, abs=1e-7)


def test_wing_reaction_first_frame_is_zero_and_of_the_demonstration_bird_is_mirrored():
    bird = bfref.make_bird(**DEMO)
    d = bfref.make_wing_dynamics(0.02, (0.03, 0.09, 0.0), (4e-5, 1e-5, 5e-5))
    dt = 0.001
    f0 = bfref.bird_prepare(bird, (0.5, 0.4, 0.3), time=0.010, dx=0.004)
    f1 = bfref.bird_prepare(bird, (0.5, 0.4, 0.3), time=0.011, dx=0.004)
    r0, prev = bfref.wing_reaction(f0, d, d, dt, np.full(16, 7.0), initialize=True)
    assert np.abs(r0).max() == 0.0
    r1, _ = bfref.wing_reaction(f1, d, d, dt, prev)
    assert np.abs(r1[0, :3]).max() > 0.0
    # the symmetric bird: the right wing's reaction mirrors the left one's
    # across the body x-z plane (force y flips; torque x and z flip)
    assert r1[2, :3] == pytest.approx(r1[0, :3] * np.array([1, -1, 1]), rel=1e-5, abs=1e-9)
    assert r1[3, :3] == pytest.approx(r1[1, :3] * np.array([-1, 1, -1]), rel=1e-5, abs=1e-9)


def test_wing_reaction_uses_the_inertia_about_each_axis():
    # anisotropic inertia (I_chord, I_span, I_normal): a rate change about
    # the normal alone is resisted by I_normal, one about the span (a wing
    # pitching about its own span) by I_span
    d = bfref.make_wing_dynamics(0.0, (0.0, 0.0, 0.0), (1e-4, 2e-4, 3e-4))
    dt = 0.001
    f1, f2 = wing_frame(0.0, 4.0), wing_frame(0.0, 9.0)
    _, prev = bfref.wing_reaction(f1, d, d, dt, np.zeros(16), initialize=True)
    r, _ = bfref.wing_reaction(f2, d, d, dt, prev)
    assert r[1, 2] == pytest.approx(-3e-4 * 5.0 / dt, rel=1e-5)
    g1, g2 = wing_frame(0.0, 0.0), wing_frame(0.0, 0.0)
    g1[8, :3] = (2.0, 0.0, 0.0)          # theta 0: the span is x
    g2[8, :3] = (6.0, 0.0, 0.0)
    _, prev = bfref.wing_reaction(g1, d, d, dt, np.zeros(16), initialize=True)
    r, _ = bfref.wing_reaction(g2, d, d, dt, prev)
    assert r[1, 0] == pytest.approx(-2e-4 * 4.0 / dt, rel=1e-5)
