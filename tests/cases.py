"""Canonical cases shared by the oracle tests and the GPU parity tests."""
from __future__ import annotations

import math

import numpy as np

import bfref

# Shear wave (MetalShearWaveValidation.swift:666-700; Reference/shear_wave_reference.py:144-178)
SHEAR_AMPLITUDE = 0.01


def shear_wave_omegas(viscosity=0.1, magic=0.25):
    tau_p = 0.5 + 3.0 * viscosity
    tau_m = 0.5 + magic / (tau_p - 0.5)
    return 1.0 / tau_p, 1.0 / tau_m


def shear_wave_params(n, omega_plus=None, omega_minus=None, **kw):
    if omega_plus is None:
        omega_plus, omega_minus = shear_wave_omegas()
    return bfref.make_params((n, n, n), omega_plus=omega_plus, omega_minus=omega_minus, **kw)


def mode_amplitude(p, f):
    """Sine projection of u_x over y (MetalShearWaveValidation.swift:357-369)."""
    _, u = bfref.moments(p, f)
    ux = u[:, 0].astype(np.float64).reshape(p.nz, p.ny, p.nx)
    y = np.arange(p.ny)
    s = np.sin(2 * np.pi * y / p.ny)
    return float(2.0 * (ux * s[None, :, None]).sum() / ux.size)


# Translating sphere (MetalTranslatingBodyTopologyValidation.swift:894-902, 4829-4860)
SPHERE_CENTER_RADIUS = (8.0, 12.0, 12.0, 3.25)
SPHERE_GEOMETRY_VELOCITY = (0.05, 0.0, 0.0)
SPHERE_WALL_VELOCITY = (0.05, 0.0, 0.0, 0.0)


def translating_sphere_omegas():
    """LatticeScaling for characteristic length 8 cells, U = 0.05, nu = 0.1, magic 3/16 (FP32)."""
    f32 = np.float32
    re = f32(0.05) * f32(8) / f32(0.1)
    nu = f32(0.05) * f32(8) / re
    tau_p = f32(0.5) + f32(3) * nu
    tau_m = f32(0.5) + f32(3.0 / 16.0) / (tau_p - f32(0.5))
    return float(f32(1) / tau_p), float(f32(1) / tau_m)


def translating_sphere_params(force_mode=6):
    op, om = translating_sphere_omegas()
    return bfref.make_params((24, 24, 24), omega_plus=op, omega_minus=om,
                             wall_rule=bfref.WALL_INTERPOLATED_HALF, force_mode=force_mode,
                             load_component=0, selected_part=1, accumulate_loads=True,
                             body_position=SPHERE_CENTER_RADIUS[:3])


def run_translating_sphere_oracle(force_mode=6, steps=40):
    p = translating_sphere_params(force_mode)
    m0, w0 = bfref.translating_sphere(p, SPHERE_CENTER_RADIUS, SPHERE_GEOMETRY_VELOCITY, SPHERE_WALL_VELOCITY, 0.0)
    f = bfref.init_equilibrium(p, m0, w0)
    prev = m0
    fx, cov, unc, transitions = [], 0, 0, 0
    for s in range(1, steps + 1):
        m, w = bfref.translating_sphere(p, SPHERE_CENTER_RADIUS, SPHERE_GEOMETRY_VELOCITY, SPHERE_WALL_VELOCITY, float(s))
        c = int(((m != 0) & (prev == 0)).sum())
        u = int(((m == 0) & (prev != 0)).sum())
        cov, unc = cov + c, unc + u
        transitions += int(c > 0 or u > 0)
        f, loads, _, _ = bfref.step(p, f, prev, m, w)
        fx.append(loads[0])
        prev = m
    return np.array(fx), cov, unc, transitions


# Deetjen dove indexed mesh (ValidationInputs/deetjen-ob-f03-surface-v1)
def deetjen_points(d):
    return {int(fr): d["points"][i] for i, fr in enumerate(d["frames"])}


def deetjen_frame_points(d, frame, points_by_frame):
    """Two-frame point array, first frame index and query time reproducing the
    reference's prepare at the exact time of `frame`.  Interior and first
    frames interpolate with blend 0 (position = that frame); the final frame
    uses the preceding segment with blend 1."""
    times = d["times"]
    last = len(times) - 1
    t = float(times[frame])
    if frame == last:
        first = last - 1
        pts = np.stack([points_by_frame[first], points_by_frame[last]])
    else:
        first = frame
        pts = np.stack([points_by_frame[frame], points_by_frame[frame]])
    return pts, first, t


# ---------------------------------------------------------------------------
# GPU-side helpers: a solver configuration equal to an oracle parameter block,
# and the oracle's indexed-mesh replay loop.
# ---------------------------------------------------------------------------
def solver_config(p, storage=0, collision=None, time_step=1.0, velocity_to_lattice=1.0, arithmetic=None):
    """SolverConfig of oracle parameters p (default arithmetic: strict for
    FP32 storage, bit-comparable with the oracle; fused for reduced storage)."""
    import birdflow_hip as bh
    if arithmetic is None:
        arithmetic = bh.ARITHMETIC_DEFAULT
    return bh.SolverConfig(arithmetic=arithmetic,
        shape=(p.nx, p.ny, p.nz), omega_plus=float(p.omega_plus), omega_minus=float(p.omega_minus),
        storage=storage, collision=p.collision if collision is None else collision, wall_rule=p.wall_rule,
        force_mode=p.force_mode, load_component=p.load_component, selected_part=p.selected_part,
        accumulate_loads=bool(p.accumulate_loads), boundary=[[p.bc[a][0], p.bc[a][1]] for a in range(3)],
        origin=tuple(p.origin), cell_size=float(p.dx), time_step=time_step,
        velocity_to_lattice=velocity_to_lattice, force_to_physical=float(p.force_scale),
        sponge_strength=float(p.sponge_strength), sponge_width=float(p.sponge_width),
        far_velocity=tuple(p.far_u), far_density=float(p.far_rho), body_position=tuple(p.body_position))


def mesh_geometry_oracle(p, seq, t, velocity_scale=1.0):
    first, _, _ = bfref.frame_interval(seq.times, t)
    pos, vel = bfref.prepare_vertices(seq.points, seq.times, first, t, velocity_scale=velocity_scale)
    mask, wall, _ = bfref.raster_resolve(p, pos, vel, seq.triangles, seq.parts, seq.half_thickness, key_bits=64)
    return mask, wall


def run_mesh_oracle(p, seq, steps, dtype=np.float32, record_masks=False):
    """Oracle replay: geometry of step s at t = s (frame s), fluid step with the
    previous / current masks (MetalIndexedBirdSurfaceValidation.swift:18347-18602)."""
    mask0, wall0 = mesh_geometry_oracle(p, seq, float(seq.times[0]))
    f = bfref.init_equilibrium(p, mask0, wall0, dtype)
    prev = mask0
    loads, masks = [], []
    for s in range(1, steps + 1):
        mask, wall = mesh_geometry_oracle(p, seq, float(seq.times[0]) + s)
        f, ld, _, _ = bfref.step(p, f, prev, mask, wall)
        loads.append(ld)
        if record_masks:
            masks.append((mask, wall))
        prev = mask
    return f, np.array(loads), masks


def plate_params(shape=(96, 48, 48), collision=bfref.COLLISION_RR3, omega_plus=1.0 / 0.6, inflow=0.05):
    """Scaled flapping-plate case: inlet (far-field equilibrium) at -x,
    zero-gradient outlet at +x, periodic y/z, interpolated moving walls with
    signed-distance link fractions, conservative loads."""
    return bfref.make_params(
        shape, omega_plus=omega_plus, omega_minus=1.0 / (0.5 + (3.0 / 16.0) / (1.0 / omega_plus - 0.5)),
        bc=((bfref.BC_FARFIELD, bfref.BC_OUTLET), (bfref.BC_PERIODIC, bfref.BC_PERIODIC),
            (bfref.BC_PERIODIC, bfref.BC_PERIODIC)),
        collision=collision, wall_rule=bfref.WALL_INTERPOLATED_SDF, force_mode=6, accumulate_loads=True,
        far_u=(inflow, 0.0, 0.0), body_position=(shape[0] / 4, shape[1] / 2, shape[2] / 2))


def integrate_body_oracle(body, loads, params, dt):
    """One free-flight update of an oracle Body from a step's loads (FP64)."""
    force = np.asarray(loads[:3], np.float64) + np.asarray(params["external_force"], np.float64)
    torque = np.asarray(loads[3:6], np.float64) + np.asarray(params["external_torque"], np.float64)
    return bfref.body_integrate(body, force, torque, params["mass"], params["inertia"], params["gravity"], dt,
                                params["substeps"])


def run_coupled_body_oracle(p, seq, steps, body, params, dtype=np.float32):
    """Free flight with the mesh carried by the body (frames in the body frame):
    geometry of step s from the pose after step s-1 at time s, fluid step with
    the torque reference at that pose, loads, then the body update.  Returns
    the final populations, the load and body histories and the final mask."""
    def geometry(t):
        first, _, _ = bfref.frame_interval(seq.times, t)
        lp, lv = bfref.prepare_vertices(seq.points, seq.times, first, t, velocity_scale=1.0)
        pos, vel = bfref.place_vertices(body, lp, lv, velocity_scale=1.0)
        mask, wall, _ = bfref.raster_resolve(p, pos, vel, seq.triangles, seq.parts, seq.half_thickness, 64)
        return mask, wall

    for i in range(3):
        p.body_position[i] = body.position[i]
    mask0, wall0 = geometry(float(seq.times[0]))
    f = bfref.init_equilibrium(p, mask0, wall0, dtype)
    prev = mask0
    loads, states = [], []
    for s in range(1, steps + 1):
        mask, wall = geometry(float(seq.times[0]) + s)
        for i in range(3):
            p.body_position[i] = body.position[i]
        f, ld, _, _ = bfref.step(p, f, prev, mask, wall)
        integrate_body_oracle(body, ld, params, 1.0)
        loads.append(ld)
        states.append(bfref.body_array(body))
        prev = mask
    return f, np.array(loads), np.array(states), prev


# ---------------------------------------------------------------------------
# Analytic signed-distance bird: the reference's compact execution case
# (Tests/BirdFlowMetalTests/MetalExecutionTests.swift:411-470).
# ---------------------------------------------------------------------------
COMPACT_BIRD = dict(body_radii=(0.015, 0.011, 0.012), wing_span=0.025, wing_root_chord=0.020, wing_tip_chord=0.012,
                    wing_thickness=0.008, wing_sweep=0.004, wing_root_offset=(0.0, 0.010, 0.003), tail_length=0.020,
                    tail_half_width=0.014, tail_thickness=0.008, frequency=3.0, stroke_amplitude=0.25,
                    stroke_bias=0.0, pitch_mean=0.0, pitch_amplitude=0.15, pitch_phase=0.4)


def compact_bird_case(free_flight=False, shape=(48, 48, 48)):
    """(oracle params, SolverConfig, BirdGeometry, oracle Bird, scaling): the
    compact bird at the domain centre, LatticeScaling(chord 0.02 m over 8
    cells, 1 m/s, Re 100, air 1.225, lattice speed 0.03), far field -1 m/s
    in x (0 in free flight), sponge 0.04 over 4 cells, far-field faces,
    TRT, halfway walls, force mode 6."""
    import birdflow_hip as bh
    sc = bh.LatticeScaling.make(0.020, 8, 1.0, 100.0, 1.225, lattice_reference_speed=0.03)
    dx = np.float32(sc.cell_size)
    center = tuple(float(np.float32(0.5 * n * dx)) for n in shape)
    far = (0.0, 0.0, 0.0) if free_flight else (float(np.float32(-1.0) * np.float32(sc.velocity_to_lattice)), 0.0, 0.0)
    p = bfref.make_params(shape, omega_plus=sc.omega_plus, omega_minus=sc.omega_minus,
                          bc=((bfref.BC_FARFIELD, bfref.BC_FARFIELD),) * 3, collision=bfref.COLLISION_TRT,
                          wall_rule=bfref.WALL_HALFWAY, force_mode=6, accumulate_loads=True, sponge_strength=0.04,
                          sponge_width=4.0, far_u=far, origin=(0.0, 0.0, 0.0), dx=float(dx),
                          force_scale=sc.force_to_physical, body_position=center)
    cfg = solver_config(p, time_step=sc.time_step, velocity_to_lattice=sc.velocity_to_lattice)
    return p, cfg, bh.BirdGeometry(**COMPACT_BIRD), bfref.make_bird(**COMPACT_BIRD), sc


def bird_frame_oracle(bird, body, time, dx, keyframes=None):
    """Oracle frame from an FP64 body 13-vector, cast to FP32 as the device does."""
    b = np.asarray(body, np.float64)
    return bfref.bird_prepare(bird, b[0:3].astype(np.float32), b[3:7].astype(np.float32),
                              b[7:10].astype(np.float32), b[10:13].astype(np.float32), time=time, dx=dx,
                              keyframes=keyframes)


def compact_bird_keyframes(count=24):
    """Measured-style keyframes for the compact bird: stroke and pitch of its
    sinusoids, deviation 0.2 sin(.+0.5) and tip twist 0.35 sin(.+1.0) on the
    left, the right side lagging by a tenth of a cycle (asymmetric)."""
    c = COMPACT_BIRD
    f = c["frequency"]
    w = 2 * math.pi * f
    kf = np.zeros((count, 17))
    for k in range(count):
        ph = k / count
        row = [ph]
        for lag in (0.0, 0.1):
            t = (ph - lag) / f
            a = [c["stroke_bias"] + c["stroke_amplitude"] * math.sin(w * t), 0.2 * math.sin(w * t + 0.5),
                 c["pitch_mean"] + c["pitch_amplitude"] * math.sin(w * t + c["pitch_phase"]),
                 0.35 * math.sin(w * t + 1.0)]
            r = [w * c["stroke_amplitude"] * math.cos(w * t), w * 0.2 * math.cos(w * t + 0.5),
                 w * c["pitch_amplitude"] * math.cos(w * t + c["pitch_phase"]), w * 0.35 * math.cos(w * t + 1.0)]
            row += a + r
        kf[k] = row
    return kf.astype(np.float32)


WING_SHAPE = (64, 64, 32)
WING_ROOT = (32.0, 32.0, 16.0)
WING_CYCLE = 200.0
WING_TIME_STEP = 2.0
WING_VELOCITY_TO_LATTICE = 0.05


def wing_case(root=WING_ROOT, periodic=False):
    """(oracle params, SolverConfig, FlappingWing, oracle Wing): the Li &
    Nabawy planform with an 8-cell mean chord rooted at `root` (default the
    centre) of a 64 x 64 x 32 box in cells (dx 1), cycle 200 time units at 2
    per step, far-field faces with a weak stream and a sponge 0.04 over 4
    cells (or, periodic, every face periodic and no sponge), TRT, the
    link-table wall rule, force mode 6."""
    import birdflow_hip as bh
    bc = ((bfref.BC_PERIODIC, bfref.BC_PERIODIC) if periodic else (bfref.BC_FARFIELD, bfref.BC_FARFIELD),) * 3
    p = bfref.make_params(WING_SHAPE, omega_plus=1.6, omega_minus=1.0 / (0.1875 / (1.0 / 1.6 - 0.5) + 0.5),
                          bc=bc, collision=bfref.COLLISION_TRT,
                          wall_rule=bfref.WALL_INTERPOLATED_LINKS, force_mode=6, accumulate_loads=True,
                          sponge_strength=0.0 if periodic else 0.04, sponge_width=4.0, far_u=(0.01, 0.0, 0.0),
                          origin=(0.0, 0.0, 0.0), dx=1.0, force_scale=1.0, body_position=root)
    cfg = solver_config(p, time_step=WING_TIME_STEP, velocity_to_lattice=WING_VELOCITY_TO_LATTICE)
    return (p, cfg, bh.FlappingWing.li_nabawy(8.0, root, WING_CYCLE),
            bfref.li_nabawy_wing(8.0, root, WING_CYCLE))


# ---------------------------------------------------------------------------
# The reference's indexed-dove production coupling replay
# (MetalIndexedBirdSurfaceValidation.swift:19284-19600, archived as
# ValidationArtifacts/deetjen-dove-indexed-production-coupling.json):
# 59x53x50 periodic, TRT, signed-distance link fractions, force mode 6,
# geometry of step s at time s * dt from mesh frames 0 and 1.
# ---------------------------------------------------------------------------
class DeetjenCoupling:
    def __init__(self, golden_dir):
        import json
        import os
        d = np.load(os.path.join(golden_dir, "deetjen_subset.npz"))
        self.ref = json.load(open(os.path.join(golden_dir, "deetjen_coupling.json")))
        f32 = np.float32
        self.dx = f32(d["dx"])
        self.dt = f32(self.ref["timeStepSeconds"])
        # LatticeScaling (SimulationConfiguration.swift:137-183) in FP32 with the
        # replay's reference length 8 cells, Re 1000, lattice speed 0.04, magic 3/16
        nu = f32(f32(0.04) * f32(8)) / f32(1000)
        tau_p = f32(0.5) + f32(3) * nu
        tau_m = f32(0.5) + f32(f32(3.0) / f32(16.0)) / f32(tau_p - f32(0.5))
        self.omega_plus, self.omega_minus = f32(1) / tau_p, f32(1) / tau_m
        self.velocity_to_lattice = f32(self.dt / self.dx)
        self.force_to_physical = f32(f32(self.dx ** f32(4)) / f32(self.dt ** f32(2)))
        pts = deetjen_points(d)
        self.points = np.stack([pts[0], pts[1]]).astype(np.float32)
        self.times = d["times"][:2].astype(np.float32)
        self.triangles = d["triangles"]
        self.parts = d["triangle_parts"]
        self.half_thickness = f32(d["half_thickness_cells"]) * self.dx
        self.origin = tuple(float(v) for v in d["origin"])
        self.grid = tuple(int(v) for v in d["grid"])
        self.params = bfref.make_params(
            self.grid, omega_plus=float(self.omega_plus), omega_minus=float(self.omega_minus),
            collision=bfref.COLLISION_TRT, wall_rule=bfref.WALL_INTERPOLATED_SDF, force_mode=6,
            accumulate_loads=True, origin=self.origin, dx=float(self.dx),
            force_scale=float(self.force_to_physical))

    def time(self, step):
        """Mesh time of step s: initial time + Float(s) * dt (FP32)."""
        return float(np.float32(self.times[0]) + np.float32(step) * self.dt)

    def geometry(self, step):
        t = self.time(step)
        pos, vel = bfref.prepare_vertices(self.points, self.times, 0, t,
                                          velocity_scale=float(self.velocity_to_lattice))
        mask, wall, _ = bfref.raster_resolve(self.params, pos, vel, self.triangles, self.parts,
                                             self.half_thickness, key_bits=64)
        return mask, wall

    def run_oracle(self, steps=8):
        """Per step: covered, uncovered, persistent links, load, masks."""
        p = self.params
        mask0, wall0 = self.geometry(0)
        f = bfref.init_equilibrium(p, mask0, wall0, np.float32)
        prev = mask0
        rows = []
        for s in range(1, steps + 1):
            mask, wall = self.geometry(s)
            fo, ld, _, _ = bfref.step(p, f, prev, mask, wall)
            _, counts = bfref.ledger_sources(p, f, fo, prev, mask, wall)
            cov = int(((prev == 0) & (mask != 0)).sum())
            unc = int(((prev != 0) & (mask == 0)).sum())
            rows.append(dict(covered=cov, uncovered=unc, persistent=int(counts[2]), transitions=int(counts[3]),
                             load=ld, mask=mask, populations=fo))
            f, prev = fo, mask
        return rows


# ---------------------------------------------------------------------------
# The reference's production fixed-wing diagnostic
# (MetalWingValidator.runSingleCase(resolution: 80),
# MetalWingValidation.swift:102-147, 295-436; locked in
# MetalExecutionTests.swift:2204-2224): the Taira & Colonius AR = 2 flat
# plate at Re 100 and 30 degrees, a one-cell voxel plate
# (initializeFixedWingCase, BirdFlow.metal:2357-2396) in uniform flow at
# 0.08 inclined by the angle of attack, 80 x 80 x 48 cells (8 per chord),
# far-field faces, a 4-cell sponge of strength 0.04, TRT with magic 3/16,
# halfway bounce-back, force mode 6; the force of step 1,300 (U t / c = 13).
# ---------------------------------------------------------------------------
FIXED_WING_LOCKED = {"lift": 0.5426947226244379, "drag": 0.672869301552491}


class FixedWing:
    def __init__(self, resolution=80):
        f32 = np.float32
        self.chord = resolution // 10
        span = 2 * self.chord
        shape = (resolution, self.chord * 10, self.chord * 6)
        self.shape = shape
        speed = f32(0.08)
        # LatticeScaling (SimulationConfiguration.swift:137-183) in FP32:
        # length = chord cells, reference speed = lattice speed, Re 100
        nu = f32(speed * f32(self.chord)) / f32(100)
        tau_p = f32(f32(0.5) + f32(3) * nu)
        tau_m = f32(f32(0.5) + f32(f32(3.0) / f32(16.0)) / f32(tau_p - f32(0.5)))
        self.alpha = 30.0 * math.pi / 180.0
        a32 = f32(self.alpha)
        far = (float(f32(speed * f32(math.cos(a32)))), float(f32(speed * f32(math.sin(a32)))), 0.0)
        center = (f32(0.3) * f32(resolution), f32(0.5) * f32(shape[1]) + f32(0.5), f32(0.5) * f32(shape[2]))
        z, y, x = np.meshgrid(np.arange(shape[2]), np.arange(shape[1]), np.arange(shape[0]), indexing="ij")
        rx = (x.astype(f32) + f32(0.5)) - center[0]
        ry = (y.astype(f32) + f32(0.5)) - center[1]
        rz = (z.astype(f32) + f32(0.5)) - center[2]
        plate = (np.abs(rx) <= f32(0.5 * self.chord)) & (np.abs(rz) <= f32(0.5 * span)) & (np.abs(ry) <= f32(0.5))
        self.mask = plate.astype(np.uint8).ravel()
        self.wall = np.zeros((self.mask.size, 4), np.float32)
        self.params = bfref.make_params(
            shape, omega_plus=float(f32(1) / tau_p), omega_minus=float(f32(1) / tau_m),
            bc=((bfref.BC_FARFIELD, bfref.BC_FARFIELD),) * 3, collision=bfref.COLLISION_TRT,
            wall_rule=bfref.WALL_HALFWAY, force_mode=6, accumulate_loads=True, sponge_strength=0.04,
            sponge_width=float(max(4, self.chord // 2)), far_u=far, body_position=tuple(float(c) for c in center))
        self.steps = int(round(13.0 * self.chord / 0.08))

    def coefficients(self, load):
        """(lift, drag) of a load (forceCoefficients, MetalWingValidation.swift:421-436)."""
        q = 0.5 * 0.08 * 0.08 * (2.0 * self.chord * self.chord)
        s, c = math.sin(self.alpha), math.cos(self.alpha)
        return (-load[0] * s + load[1] * c) / q, (load[0] * c + load[1] * s) / q


def kinematics_f64(mesh):
    """An ArticulatedMesh's kinematics as the dict bfref.articulated_vertices_f64
    takes, coefficients kept in FP64."""
    return {"period": float(mesh.period), "amplitude_scale": float(mesh.amplitude_scale),
            "harmonics": int(mesh.harmonics), "center": [float(v) for v in mesh.center],
            "joints": [{"side": float(j.side), "root": [float(v) for v in j.root],
                        **{n: np.asarray(getattr(j, n), np.float64) for n in ("stroke", "deviation", "pitch", "twist")}}
                       for j in mesh.joints]}


def oracle_kinematics(mesh):
    """An ArticulatedMesh's kinematics for the FP32 oracle (bfref.Kinematics)."""
    return bfref.make_kinematics(mesh.period, mesh.amplitude_scale, mesh.center,
                                 [dict(side=j.side, root=j.root, stroke=j.stroke, deviation=j.deviation,
                                       pitch=j.pitch, twist=j.twist) for j in mesh.joints],
                                 harmonics=mesh.harmonics)
