"""ctypes binding of the CPU oracle (oracle/libbfref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py``, always as the checker or the
timed CPU baseline, never as the product path.

The functions restate the reference solver's arithmetic; bfref.h lists the
reference file:line each one follows.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libbfref.so")

BC_PERIODIC, BC_FARFIELD, BC_OUTLET = 0, 1, 2
COLLISION_TRT, COLLISION_TRT_SYMMETRIC_LIMITER, COLLISION_REGULARIZED, COLLISION_RR3 = 0, 1, 2, 3
WALL_HALFWAY, WALL_INTERPOLATED_SDF, WALL_INTERPOLATED_HALF, WALL_INTERPOLATED_LINKS = 0, 1, 2, 3

C = np.array(
    [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
     (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1),
     (1, 0, -1), (-1, 0, 1), (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1)], dtype=np.int64)
OPP = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


class Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("bc", (ctypes.c_int32 * 2) * 3),
        ("collision", ctypes.c_int32),
        ("wall_rule", ctypes.c_int32),
        ("force_mode", ctypes.c_int32),
        ("load_component", ctypes.c_int32),
        ("selected_part", ctypes.c_int32),
        ("accumulate_loads", ctypes.c_int32),
        ("capture_fields", ctypes.c_int32),
        ("pad0", ctypes.c_int32),
        ("omega_plus", ctypes.c_float), ("omega_minus", ctypes.c_float),
        ("sponge_strength", ctypes.c_float), ("sponge_width", ctypes.c_float),
        ("far_u", ctypes.c_float * 3), ("far_rho", ctypes.c_float),
        ("origin", ctypes.c_float * 3), ("dx", ctypes.c_float),
        ("force_scale", ctypes.c_float),
        ("body_position", ctypes.c_float * 3),
    ]


class Body(ctypes.Structure):
    _fields_ = [
        ("position", ctypes.c_double * 3),
        ("orientation", ctypes.c_double * 4),
        ("linear_velocity", ctypes.c_double * 3),
        ("angular_velocity_body", ctypes.c_double * 3),
    ]


class WingDynamics(ctypes.Structure):
    """bfo_wing_dynamics: mass, centre of mass from the hinge along chord /
    span / normal, principal inertia about those axes."""
    _fields_ = [("mass", ctypes.c_float), ("center_of_mass", ctypes.c_float * 3), ("inertia", ctypes.c_float * 3)]


def make_wing_dynamics(mass, center_of_mass, inertia) -> WingDynamics:
    d = WingDynamics()
    d.mass = mass
    d.center_of_mass[:] = list(center_of_mass)
    d.inertia[:] = list(inertia)
    return d


class Bird(ctypes.Structure):
    """bfo_bird: GPUBirdParameters geometry and analytic kinematics."""
    _fields_ = [
        ("body_radii", ctypes.c_float * 3),
        ("wing_span", ctypes.c_float), ("wing_root_chord", ctypes.c_float), ("wing_tip_chord", ctypes.c_float),
        ("wing_thickness", ctypes.c_float), ("wing_sweep", ctypes.c_float),
        ("wing_root_offset", ctypes.c_float * 3),
        ("tail_length", ctypes.c_float), ("tail_half_width", ctypes.c_float), ("tail_thickness", ctypes.c_float),
        ("frequency", ctypes.c_float), ("stroke_amplitude", ctypes.c_float), ("stroke_bias", ctypes.c_float),
        ("pitch_mean", ctypes.c_float), ("pitch_amplitude", ctypes.c_float), ("pitch_phase", ctypes.c_float),
    ]


class Wing(ctypes.Structure):
    """bfo_wing: GPUFlappingWingParameters."""
    _fields_ = [(name, ctypes.c_float) for name in (
        "root_x", "root_y", "root_z", "chord_mean", "radius", "thickness", "beta_exponent", "beta_normalization",
        "cycle", "stroke_amplitude", "stroke_duration", "pitch_duration", "pitch_low", "pitch_high",
        "stroke_rate_max", "reference_speed")]


class Joint(ctypes.Structure):
    """bfo_joint: one articulated joint (side, hinge, Fourier series)."""
    _fields_ = [("side", ctypes.c_float), ("root", ctypes.c_float * 3)] + [
        (name, (ctypes.c_float * 2) * 4) for name in ("stroke", "deviation", "pitch", "twist")]


class Kinematics(ctypes.Structure):
    """bfo_kinematics: period, amplitude scale, harmonics, joints."""
    _fields_ = [("period", ctypes.c_float), ("amplitude_scale", ctypes.c_float), ("harmonics", ctypes.c_uint32),
                ("joint_count", ctypes.c_uint32), ("center", ctypes.c_float * 3), ("reserved", ctypes.c_float),
                ("joints", Joint * 8)]


def make_kinematics(period, amplitude_scale, center, joints, harmonics=None) -> Kinematics:
    """joints: list of dicts with side, root[3] and stroke / deviation /
    pitch / twist arrays [harmonics][2] (sin, cos coefficients, radians)."""
    k = Kinematics()
    k.period, k.amplitude_scale = period, amplitude_scale
    k.center[:] = [float(v) for v in center]
    k.joint_count = len(joints)
    h_max = 1
    for i, j in enumerate(joints):
        J = k.joints[i]
        J.side = float(j["side"])
        J.root[:] = [float(v) for v in j["root"]]
        for name in ("stroke", "deviation", "pitch", "twist"):
            c = np.asarray(j[name], np.float32).reshape(-1, 2)
            h_max = max(h_max, c.shape[0])
            arr = getattr(J, name)
            for h in range(c.shape[0]):
                arr[h][0], arr[h][1] = float(c[h, 0]), float(c[h, 1])
    k.harmonics = harmonics if harmonics is not None else h_max
    return k


def articulated_vertices(rest, vertex_joint, span_fraction, k: Kinematics, t, velocity_scale=1.0):
    """Positions and velocities [V][3] of an articulated mesh at time t."""
    rest = np.ascontiguousarray(rest, np.float32).reshape(-1, 3)
    vj = np.ascontiguousarray(vertex_joint, np.uint8)
    sf = np.ascontiguousarray(span_fraction, np.float32)
    pos = np.empty_like(rest)
    vel = np.empty_like(rest)
    lib().bfo_articulated_vertices(_ptr(rest), _ptr(vj), _ptr(sf), rest.shape[0], ctypes.byref(k), np.float32(t),
                                   np.float32(velocity_scale), _ptr(pos), _ptr(vel))
    return pos, vel


def articulated_vertices_f64(rest, vertex_joint, span_fraction, kin: dict, t, velocity_scale=1.0):
    """bfo_articulated_vertices (bfref_kin.inc) in FP64 with FP64
    coefficients: positions and velocities [V][3] at time t, the reference
    for finite differences with respect to the Fourier coefficients.  kin:
    period, amplitude_scale, harmonics, center and joints (dicts with side,
    root and stroke / deviation / pitch / twist [harmonic][sin, cos])."""
    rest = np.asarray(rest, np.float64).reshape(-1, 3)
    vj = np.asarray(vertex_joint).astype(np.int64)
    sf = np.asarray(span_fraction, np.float64)
    period = float(kin["period"])
    scale = float(kin["amplitude_scale"])
    H = int(kin["harmonics"])
    omega = 2.0 * np.pi / period
    phi = omega * (t - np.floor(t / period) * period)
    h = np.arange(1, H + 1, dtype=np.float64)
    sh, ch = np.sin(h * phi), np.cos(h * phi)
    center = np.asarray(kin["center"], np.float64)
    pos = center + rest
    vel = np.zeros_like(rest)
    for j, J in enumerate(kin["joints"]):
        sel = vj == j + 1
        if not sel.any():
            continue
        ang, rate = [], []
        for name in ("stroke", "deviation", "pitch", "twist"):
            c = np.asarray(J[name], np.float64).reshape(-1, 2)[:H]
            ang.append(scale * float(c[:, 0] @ sh + c[:, 1] @ ch))
            rate.append(scale * omega * float((h * (c[:, 0] * ch - c[:, 1] * sh)).sum()))
        s = float(J["side"])
        sa, ca = np.sin(s * ang[0]), np.cos(s * ang[0])
        sb, cb = np.sin(s * ang[1]), np.cos(s * ang[1])
        l = rest[sel] * np.array([1.0, s, 1.0])
        a = ang[2] + ang[3] * sf[sel]
        adot = rate[2] + rate[3] * sf[sel]
        sp, cp = np.sin(a), np.cos(a)
        u3 = np.stack([cp * l[:, 0] + sp * l[:, 2], l[:, 1], -sp * l[:, 0] + cp * l[:, 2]], 1)
        u2 = np.stack([cb * u3[:, 0] - sb * u3[:, 1], sb * u3[:, 0] + cb * u3[:, 1], u3[:, 2]], 1)
        u1 = np.stack([u2[:, 0], ca * u2[:, 1] - sa * u2[:, 2], sa * u2[:, 1] + ca * u2[:, 2]], 1)
        root = np.asarray(J["root"], np.float64) * np.array([1.0, s, 1.0])
        pos[sel] = center + root + u1
        ad, bd = s * rate[0], s * rate[1]
        t1 = np.stack([np.zeros(len(u1)), -ad * u1[:, 2], ad * u1[:, 1]], 1)
        w2 = np.stack([-bd * u2[:, 1], bd * u2[:, 0], np.zeros(len(u2))], 1)
        t2 = np.stack([w2[:, 0], ca * w2[:, 1] - sa * w2[:, 2], sa * w2[:, 1] + ca * w2[:, 2]], 1)
        w3 = np.stack([adot * u3[:, 2], np.zeros(len(u3)), -adot * u3[:, 0]], 1)
        z3 = np.stack([cb * w3[:, 0] - sb * w3[:, 1], sb * w3[:, 0] + cb * w3[:, 1], w3[:, 2]], 1)
        t3 = np.stack([z3[:, 0], ca * z3[:, 1] - sa * z3[:, 2], sa * z3[:, 1] + ca * z3[:, 2]], 1)
        vel[sel] = t1 + t2 + t3
    return pos, vel * velocity_scale


WING_FRAME_FIELDS = ("root", "chord", "span", "normal", "angular", "state")


BIRD_FRAME_FIELDS = ("body_position", "orientation", "linear_velocity", "omega_world_radius",
                     "left_root", "left_chord", "left_span", "left_normal", "left_angular",
                     "right_root", "right_chord", "right_span", "right_normal", "right_angular")


def build() -> str:
    """Compile the oracle with its Makefile (used where a toolchain exists)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _declare(l):
    vp, P = ctypes.c_void_p, ctypes.POINTER(Params)
    for sfx in ("f32", "f64"):
        getattr(l, f"bfo_step_{sfx}").argtypes = [P, vp, vp, vp, vp, vp, vp, vp, vp]
        getattr(l, f"bfo_init_equilibrium_{sfx}").argtypes = [P, vp, vp, vp]
        getattr(l, f"bfo_moments_{sfx}").argtypes = [P, vp, vp, vp]
    l.bfo_step_f64w.argtypes = [P, vp, vp, vp, vp, vp, vp, vp, vp]
    l.bfo_init_shear_wave_f32.argtypes = [P, ctypes.c_float, vp]
    l.bfo_init_shear_wave_f64.argtypes = [P, ctypes.c_double, vp]
    l.bfo_equilibrium_f32.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float]
    l.bfo_equilibrium_f32.restype = ctypes.c_float
    l.bfo_equilibrium_f64.argtypes = [ctypes.c_int] + [ctypes.c_double] * 4
    l.bfo_equilibrium_f64.restype = ctypes.c_double
    l.bfo_link_fraction_f64.argtypes = [ctypes.c_double, ctypes.c_double]
    l.bfo_link_fraction_f64.restype = ctypes.c_double
    l.bfo_link_fraction_f32.argtypes = [ctypes.c_float, ctypes.c_float]
    l.bfo_link_fraction_f32.restype = ctypes.c_float
    l.bfo_bouzidi_population_f64.argtypes = [ctypes.c_double] * 5
    l.bfo_bouzidi_population_f64.restype = ctypes.c_double
    l.bfo_frame_interval.argtypes = [vp, ctypes.c_int, ctypes.c_float, vp, vp]
    l.bfo_frame_interval.restype = ctypes.c_int
    l.bfo_prepare_vertices.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                       vp, ctypes.c_float, vp, vp]
    l.bfo_raster.argtypes = [P, vp, vp, ctypes.c_int, ctypes.c_float, ctypes.c_int, vp]
    l.bfo_resolve.argtypes = [P, vp, vp, vp, vp, ctypes.c_int, ctypes.c_float, ctypes.c_int, vp, vp, vp]
    l.bfo_clear_geometry.argtypes = [P, vp, vp]
    l.bfo_translating_sphere.argtypes = [P, vp, vp, vp, ctypes.c_float, vp, vp]
    l.bfo_link_census.argtypes = [P, vp, vp, vp, vp, vp]
    l.bfo_body_integrate.argtypes = [ctypes.POINTER(Body), vp, vp, ctypes.c_double, vp, vp,
                                     ctypes.c_double, ctypes.c_int]
    l.bfo_body_safety.argtypes = [ctypes.POINTER(Body), vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  vp, vp, vp, ctypes.c_double, ctypes.c_uint64, vp, vp]
    l.bfo_body_safety.restype = ctypes.c_uint32
    l.bfo_place_vertices.argtypes = [ctypes.POINTER(Body), ctypes.c_int, ctypes.c_float, vp, vp, vp, vp]
    l.bfo_articulated_vertices.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.POINTER(Kinematics), ctypes.c_float,
                                           ctypes.c_float, vp, vp]
    l.bfo_sincos_n.argtypes = [vp, ctypes.c_int, vp, vp]
    l.bfo_set_link_table.argtypes = [vp]
    l.bfo_wing_pow.argtypes = [ctypes.c_float, ctypes.c_float]
    l.bfo_wing_pow.restype = ctypes.c_float
    l.bfo_wing_prepare.argtypes = [ctypes.POINTER(Wing), ctypes.c_float, vp]
    l.bfo_wing_link_fraction.argtypes = [ctypes.POINTER(Wing), vp, ctypes.c_float, vp, vp]
    l.bfo_wing_link_fraction.restype = ctypes.c_float
    l.bfo_wing_build.argtypes = [P, ctypes.POINTER(Wing), vp, vp, ctypes.c_float, vp, vp]
    l.bfo_wing_links.argtypes = [P, ctypes.POINTER(Wing), vp, vp, vp]
    l.bfo_bird_prepare.argtypes = [ctypes.POINTER(Bird), vp, vp, vp, vp, ctypes.c_float, ctypes.c_float, vp,
                                   ctypes.c_int, vp]
    l.bfo_bird_build.argtypes = [P, ctypes.POINTER(Bird), vp, vp, ctypes.c_float, vp, vp]
    l.bfo_ray_links.argtypes = [P, vp, vp, ctypes.c_int, ctypes.c_float, vp, vp, vp]
    l.bfo_ray_shell.argtypes = [vp, vp, vp, vp, vp, ctypes.c_float]
    l.bfo_ray_shell.restype = ctypes.c_float
    l.bfo_ray_links_winners.argtypes = [P, vp, vp, ctypes.c_int, ctypes.c_float, vp, vp, vp, vp]
    l.bfo_ray_shell_f64_n.argtypes = [ctypes.c_int, vp, vp, vp, vp, vp, ctypes.c_double, vp]
    l.bfo_set_link_table_f64.argtypes = [vp]
    l.bfo_wing_reaction.argtypes = [vp, ctypes.POINTER(WingDynamics), ctypes.POINTER(WingDynamics), ctypes.c_float,
                                    ctypes.c_int, vp, vp]


def make_params(shape, *, omega_plus, omega_minus, bc=None, collision=COLLISION_TRT,
                wall_rule=WALL_HALFWAY, force_mode=6, load_component=0, selected_part=0,
                accumulate_loads=False, capture_fields=False, sponge_strength=0.0,
                sponge_width=4.0, far_u=(0.0, 0.0, 0.0), far_rho=1.0, origin=(0.0, 0.0, 0.0),
                dx=1.0, force_scale=1.0, body_position=(0.0, 0.0, 0.0)) -> Params:
    p = Params()
    p.nx, p.ny, p.nz = (int(v) for v in shape)
    if bc is None:
        bc = ((BC_PERIODIC, BC_PERIODIC),) * 3
    for a in range(3):
        for s in range(2):
            p.bc[a][s] = int(bc[a][s])
    p.collision, p.wall_rule, p.force_mode = collision, wall_rule, force_mode
    p.load_component, p.selected_part = load_component, selected_part
    p.accumulate_loads, p.capture_fields = int(bool(accumulate_loads)), int(bool(capture_fields))
    p.omega_plus, p.omega_minus = omega_plus, omega_minus
    p.sponge_strength, p.sponge_width = sponge_strength, sponge_width
    p.far_u[:] = list(far_u)
    p.far_rho = far_rho
    p.origin[:] = list(origin)
    p.dx = dx
    p.force_scale = force_scale
    p.body_position[:] = list(body_position)
    return p


def cells(p: Params) -> int:
    return p.nx * p.ny * p.nz


def _dtype_sfx(dtype):
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def step(p: Params, fin, solid_prev=None, solid_cur=None, wall4=None, capture=False):
    """One fused step; returns (fout, loads[8], density, velocity4)."""
    n = cells(p)
    sfx = _dtype_sfx(fin.dtype)
    fin = np.ascontiguousarray(fin)
    fout = np.empty_like(fin)
    zeros = np.zeros(n, np.uint8)
    solid_prev = zeros if solid_prev is None else np.ascontiguousarray(solid_prev, np.uint8)
    solid_cur = zeros if solid_cur is None else np.ascontiguousarray(solid_cur, np.uint8)
    if wall4 is None:
        wall4 = np.zeros((n, 4), np.float32)
        wall4[:, 3] = 16.0
    if sfx == "f64" and np.asarray(wall4).dtype == np.float64:
        sfx = "f64w"            # FP64 wall field (reference gradients)
        wall4 = np.ascontiguousarray(wall4, np.float64)
    else:
        wall4 = np.ascontiguousarray(wall4, np.float32)
    loads = np.zeros(8, np.float64)
    density = velocity4 = None
    p.capture_fields = int(bool(capture))
    if capture:
        density = np.empty(n, fin.dtype)
        velocity4 = np.empty((n, 4), fin.dtype)
    getattr(lib(), f"bfo_step_{sfx}")(ctypes.byref(p), _ptr(fin), _ptr(fout), _ptr(solid_prev),
                                      _ptr(solid_cur), _ptr(wall4), _ptr(density), _ptr(velocity4),
                                      _ptr(loads))
    return fout, loads, density, velocity4


def init_shear_wave(p: Params, amplitude: float, dtype=np.float32):
    f = np.empty(19 * cells(p), dtype)
    getattr(lib(), f"bfo_init_shear_wave_{_dtype_sfx(dtype)}")(ctypes.byref(p), amplitude, _ptr(f))
    return f


def init_equilibrium(p: Params, solid=None, wall4=None, dtype=np.float32):
    f = np.empty(19 * cells(p), dtype)
    solid = None if solid is None else np.ascontiguousarray(solid, np.uint8)
    wall4 = None if wall4 is None else np.ascontiguousarray(wall4, np.float32)
    getattr(lib(), f"bfo_init_equilibrium_{_dtype_sfx(dtype)}")(ctypes.byref(p), _ptr(f), _ptr(solid), _ptr(wall4))
    return f


def moments(p: Params, f):
    n = cells(p)
    rho = np.empty(n, f.dtype)
    u = np.empty((n, 3), f.dtype)
    getattr(lib(), f"bfo_moments_{_dtype_sfx(f.dtype)}")(ctypes.byref(p), _ptr(np.ascontiguousarray(f)), _ptr(rho), _ptr(u))
    return rho, u


def equilibrium_f32(q, rho, u):
    return lib().bfo_equilibrium_f32(q, rho, *u)


def equilibrium_f64(q, rho, u):
    return lib().bfo_equilibrium_f64(q, rho, *u)


def link_fraction_f64(fluid, solid):
    return lib().bfo_link_fraction_f64(fluid, solid)


def bouzidi_population_f64(q, reflected, farther, previous, correction):
    return lib().bfo_bouzidi_population_f64(q, reflected, farther, previous, correction)


def frame_interval(times, t):
    times = np.ascontiguousarray(times, np.float32)
    blend = np.zeros(1, np.float32)
    dur = np.zeros(1, np.float32)
    first = lib().bfo_frame_interval(_ptr(times), len(times), np.float32(t), _ptr(blend), _ptr(dur))
    return first, float(blend[0]), float(dur[0])


def prepare_vertices(points, times, first, query_time, translation=(0, 0, 0), velocity_scale=1.0):
    points = np.ascontiguousarray(points, np.float32)
    times = np.ascontiguousarray(times, np.float32)
    frames, verts = points.shape[0], points.shape[1]
    pos = np.empty((verts, 3), np.float32)
    vel = np.empty((verts, 3), np.float32)
    tr = np.ascontiguousarray(translation, np.float32)
    lib().bfo_prepare_vertices(_ptr(points), _ptr(times), frames, verts, int(first), np.float32(query_time),
                               _ptr(tr), np.float32(velocity_scale), _ptr(pos), _ptr(vel))
    return pos, vel


def raster_resolve(p: Params, pos, vel, tri, tri_part, half_thickness, key_bits=64):
    """Clear + raster + resolve; returns (mask, wall4, keys)."""
    n = cells(p)
    tri = np.ascontiguousarray(tri, np.uint32)
    tri_part = np.ascontiguousarray(tri_part, np.uint8)
    pos = np.ascontiguousarray(pos, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    keys = np.full(n, np.iinfo(np.uint32 if key_bits == 32 else np.uint64).max,
                   np.uint32 if key_bits == 32 else np.uint64)
    mask = np.zeros(n, np.uint8)
    wall4 = np.zeros((n, 4), np.float32)
    lib().bfo_clear_geometry(ctypes.byref(p), _ptr(mask), _ptr(wall4))
    lib().bfo_raster(ctypes.byref(p), _ptr(pos), _ptr(tri), len(tri), np.float32(half_thickness), key_bits, _ptr(keys))
    lib().bfo_resolve(ctypes.byref(p), _ptr(pos), _ptr(vel), _ptr(tri), _ptr(tri_part), len(tri),
                      np.float32(half_thickness), key_bits, _ptr(keys), _ptr(mask), _ptr(wall4))
    return mask, wall4, keys


def translating_sphere(p: Params, center_radius, geometry_velocity, wall_velocity, t):
    n = cells(p)
    mask = np.zeros(n, np.uint8)
    wall4 = np.zeros((n, 4), np.float32)
    cr = np.ascontiguousarray(center_radius, np.float32)
    gv = np.ascontiguousarray(geometry_velocity, np.float32)
    wv = np.ascontiguousarray(wall_velocity, np.float32)
    lib().bfo_translating_sphere(ctypes.byref(p), _ptr(cr), _ptr(gv), _ptr(wv), np.float32(t), _ptr(mask), _ptr(wall4))
    return mask, wall4


def link_census(p: Params, solid_prev, solid_cur):
    counts = np.zeros(19, np.uint64)
    cov = np.zeros(1, np.uint64)
    unc = np.zeros(1, np.uint64)
    lib().bfo_link_census(ctypes.byref(p), _ptr(np.ascontiguousarray(solid_prev, np.uint8)),
                          _ptr(np.ascontiguousarray(solid_cur, np.uint8)), _ptr(counts), _ptr(cov), _ptr(unc))
    return counts, int(cov[0]), int(unc[0])


def make_body(position=(0, 0, 0), orientation=(0, 0, 0, 1), linear_velocity=(0, 0, 0),
              angular_velocity_body=(0, 0, 0)) -> Body:
    b = Body()
    b.position[:] = [float(v) for v in position]
    b.orientation[:] = [float(v) for v in orientation]
    b.linear_velocity[:] = [float(v) for v in linear_velocity]
    b.angular_velocity_body[:] = [float(v) for v in angular_velocity_body]
    return b


def body_array(body: Body) -> np.ndarray:
    """position, orientation (x y z w), linear velocity, body angular velocity."""
    return np.array(list(body.position) + list(body.orientation) + list(body.linear_velocity)
                    + list(body.angular_velocity_body), np.float64)


class Safety:
    """Running safety record: metrics [max Mach, min clearance, Mach and
    clearance at the first event], event [step lo, step hi, flags]."""

    def __init__(self):
        self.metrics = np.array([0.0, np.finfo(np.float64).max, 0.0, 0.0], np.float64)
        self.event = np.zeros(3, np.uint32)


def body_safety(body: Body, record: Safety, step, extent, rotation_radius, wing_speed, mach_limit, domain_min,
                domain_max, far_velocity_world, velocity_to_lattice):
    e = np.ascontiguousarray(extent, np.float64)
    lo = np.ascontiguousarray(domain_min, np.float64)
    hi = np.ascontiguousarray(domain_max, np.float64)
    fv = np.ascontiguousarray(far_velocity_world, np.float64)
    return int(lib().bfo_body_safety(ctypes.byref(body), _ptr(e), float(rotation_radius), float(wing_speed),
                                     float(mach_limit), _ptr(lo), _ptr(hi), _ptr(fv), float(velocity_to_lattice),
                                     int(step), _ptr(record.metrics), _ptr(record.event)))


def place_vertices(body: Body, local_pos, local_vel, velocity_scale=1.0):
    lp = np.ascontiguousarray(local_pos, np.float32)
    lv = np.ascontiguousarray(local_vel, np.float32)
    pos = np.empty_like(lp)
    vel = np.empty_like(lv)
    lib().bfo_place_vertices(ctypes.byref(body), lp.shape[0], np.float32(velocity_scale), _ptr(lp), _ptr(lv),
                             _ptr(pos), _ptr(vel))
    return pos, vel


def body_integrate(body: Body, force, torque, mass, inertia, gravity, dt, substeps=1):
    f = np.ascontiguousarray(force, np.float64)
    t = np.ascontiguousarray(torque, np.float64)
    i = np.ascontiguousarray(inertia, np.float64)
    g = np.ascontiguousarray(gravity, np.float64)
    lib().bfo_body_integrate(ctypes.byref(body), _ptr(f), _ptr(t), float(mass), _ptr(i), _ptr(g), float(dt), int(substeps))
    return body


def wall_field_f64(p: Params, pos, vel, tri, tri_of_cell, half_thickness):
    """FP64 wall field (velocity xyz, signed distance / dx) of the cells whose
    nearest triangle is given (tri_of_cell >= 0), by the closest-point
    regions of BirdFlow.metal:1529-1583 and the resolve of :1697-1738
    evaluated in FP64; other cells get (0, 0, 0, 16).  With the triangle
    assignment frozen this is the smooth wall field the adjoint
    differentiates."""
    n = cells(p)
    wall = np.zeros((n, 4), np.float64)
    wall[:, 3] = 16.0
    cellsel = np.nonzero(tri_of_cell >= 0)[0]
    if cellsel.size == 0:
        return wall
    t = tri_of_cell[cellsel]
    x = cellsel % p.nx
    y = (cellsel // p.nx) % p.ny
    z = cellsel // (p.nx * p.ny)
    P = np.stack([p.origin[0] + (x + 0.5) * p.dx, p.origin[1] + (y + 0.5) * p.dx,
                  p.origin[2] + (z + 0.5) * p.dx], 1).astype(np.float64)
    tri = np.asarray(tri, np.int64)
    pos = np.asarray(pos, np.float64)
    vel = np.asarray(vel, np.float64)
    A, B, C = pos[tri[t, 0]], pos[tri[t, 1]], pos[tri[t, 2]]
    dot = lambda u, v: np.einsum("ij,ij->i", u, v)
    ab, ac, ap = B - A, C - A, P - A
    d1, d2 = dot(ab, ap), dot(ac, ap)
    bp = P - B
    d3, d4 = dot(ab, bp), dot(ac, bp)
    cp_ = P - C
    d5, d6 = dot(ab, cp_), dot(ac, cp_)
    vc = d1 * d4 - d3 * d2
    vb = d5 * d2 - d1 * d6
    va = d3 * d6 - d5 * d4
    m = len(t)
    bar = np.zeros((m, 3))
    done = np.zeros(m, bool)

    def take(cond, b0, b1, b2):
        sel = cond & ~done
        bar[sel, 0], bar[sel, 1], bar[sel, 2] = b0[sel], b1[sel], b2[sel]
        done[sel] = True

    one, zero = np.ones(m), np.zeros(m)
    with np.errstate(divide="ignore", invalid="ignore"):
        take((d1 <= 0) & (d2 <= 0), one, zero, zero)
        take((d3 >= 0) & (d4 <= d3), zero, one, zero)
        v = d1 / (d1 - d3)
        take((vc <= 0) & (d1 >= 0) & (d3 <= 0), 1 - v, v, zero)
        take((d6 >= 0) & (d5 <= d6), zero, zero, one)
        w = d2 / (d2 - d6)
        take((vb <= 0) & (d2 >= 0) & (d6 <= 0), 1 - w, zero, w)
        w = (d4 - d3) / ((d4 - d3) + (d5 - d6))
        take((va <= 0) & ((d4 - d3) >= 0) & ((d5 - d6) >= 0), zero, 1 - w, w)
        inv = 1.0 / np.maximum(va + vb + vc, 1e-20)
        take(np.ones(m, bool), 1 - vb * inv - vc * inv, vb * inv, vc * inv)
    cp = bar[:, :1] * A + bar[:, 1:2] * B + bar[:, 2:] * C
    u = bar[:, :1] * vel[tri[t, 0]] + bar[:, 1:2] * vel[tri[t, 1]] + bar[:, 2:] * vel[tri[t, 2]]
    sd = np.linalg.norm(P - cp, axis=1) - half_thickness
    wall[cellsel, :3] = u
    wall[cellsel, 3] = sd / p.dx
    return wall


def interpolate_frames_f64(points, times, t, velocity_scale=1.0):
    """FP64 frame interpolation at time t (the prepare step,
    BirdFlow.metal:1589-1622): positions and lattice velocities."""
    first, _, _ = frame_interval(np.asarray(times, np.float32), t)
    times = np.asarray(times, np.float64)
    a, b = np.asarray(points[first], np.float64), np.asarray(points[first + 1], np.float64)
    duration = times[first + 1] - times[first]
    blend = min(max((t - times[first]) / duration, 0.0), 1.0)
    return a + blend * (b - a), (b - a) / duration * velocity_scale


# ---------------------------------------------------------------------------
# analytic signed-distance bird
# ---------------------------------------------------------------------------
def make_bird(**kw) -> Bird:
    """bfo_bird from keyword fields (sequences for the 3-vectors)."""
    b = Bird()
    for k, v in kw.items():
        if k in ("body_radii", "wing_root_offset"):
            getattr(b, k)[:] = [float(x) for x in v]
        else:
            setattr(b, k, float(v))
    return b


def sincos(x):
    x = np.ascontiguousarray(x, np.float32).ravel()
    s = np.empty_like(x)
    c = np.empty_like(x)
    lib().bfo_sincos_n(_ptr(x), x.size, _ptr(s), _ptr(c))
    return s, c


def bird_prepare(bird: Bird, position, orientation=(0, 0, 0, 1), linear_velocity=(0, 0, 0),
                 angular_velocity_body=(0, 0, 0), time=0.0, dx=1.0, keyframes=None) -> np.ndarray:
    """Prepared frame [14, 4] (rows in BIRD_FRAME_FIELDS order); keyframes:
    [count, 17] measured wing keyframes or None (analytic sinusoid)."""
    out = np.zeros((14, 4), np.float32)
    arrs = [np.ascontiguousarray(v, np.float32) for v in (position, orientation, linear_velocity,
                                                          angular_velocity_body)]
    kf = None if keyframes is None else np.ascontiguousarray(keyframes, np.float32).reshape(-1, 17)
    lib().bfo_bird_prepare(ctypes.byref(bird), *[_ptr(a) for a in arrs], np.float32(time), np.float32(dx),
                           _ptr(kf), 0 if kf is None else kf.shape[0], _ptr(out))
    return out


def wing_reaction(frame, left: WingDynamics, right: WingDynamics, dt, previous, initialize=False):
    """updateWingInertialReaction: (reaction [4, 4], new previous [4, 4]);
    rows left force, left torque, right force, right torque / the momenta."""
    fr = np.ascontiguousarray(frame, np.float32)
    prev = np.array(previous, np.float32).reshape(16).copy()
    out = np.zeros(16, np.float32)
    lib().bfo_wing_reaction(_ptr(fr), ctypes.byref(left), ctypes.byref(right), np.float32(dt), int(initialize),
                            _ptr(prev), _ptr(out))
    return out.reshape(4, 4), prev.reshape(4, 4)


def bird_build(p: Params, bird: Bird, frame, solid_prev=None, velocity_to_lattice=1.0):
    n = cells(p)
    mask = np.zeros(n, np.uint8)
    wall4 = np.zeros((n, 4), np.float32)
    fr = np.ascontiguousarray(frame, np.float32)
    prev = None if solid_prev is None else np.ascontiguousarray(solid_prev, np.uint8)
    lib().bfo_bird_build(ctypes.byref(p), ctypes.byref(bird), _ptr(fr), _ptr(prev), np.float32(velocity_to_lattice),
                         _ptr(mask), _ptr(wall4))
    return mask, wall4


# ---------------------------------------------------------------------------
# momentum ledger (FP64 restatement, numpy)
#   measureFluidMassMomentum / measureFluidMassMomentumChange /
#   captureExternalFluidSources, BirdFlow.metal:4272-4526, with per-face
#   boundaries (far-field and outlet faces are open boundaries) and every
#   term formed in FP64 from the FP32 populations.
# ---------------------------------------------------------------------------
def _grid_coords(p: Params):
    z, y, x = np.meshgrid(np.arange(p.nz), np.arange(p.ny), np.arange(p.nx), indexing="ij")
    return x.ravel(), y.ravel(), z.ravel()


def ledger_fluid(p: Params, f, solid):
    """(mass, px, py, pz) of the fluid cells (mask 0)."""
    n = cells(p)
    f = np.asarray(f, np.float64).reshape(19, n)
    fluid = np.asarray(solid).ravel() == 0
    m = f[:, fluid]
    return np.array([m.sum()] + [(m * C[:, a][:, None]).sum() for a in range(3)])


def _sponge_factor(p: Params, x, y, z):
    """Sponge factor of each cell, FP32 as the device (BirdFlow.metal:2427-2446)."""
    f32 = np.float32
    if not p.sponge_strength > 0:
        return np.zeros(x.shape, f32)
    d = np.minimum.reduce([x, y, z, p.nx - 1 - x, p.ny - 1 - y, p.nz - 1 - z]).astype(f32)
    width = f32(p.sponge_width)
    nrm = np.clip((width - d) / np.maximum(width, f32(1.0)), f32(0), f32(1)).astype(f32)
    return (f32(p.sponge_strength) * nrm * nrm).astype(f32)


def ledger_sources(p: Params, f_in, f_out, solid_prev, solid_cur, wall4):
    """Sums (change, open boundary, sponge, persistent link) as 4 x (mass,
    px, py, pz) and counts (open-boundary links, sponge cells, persistent
    links, occupancy changes) of one step."""
    n = cells(p)
    fi = np.asarray(f_in, np.float32).reshape(19, n).astype(np.float64)
    fo = np.asarray(f_out, np.float32).reshape(19, n).astype(np.float64)
    was = np.asarray(solid_prev).ravel() != 0
    is_ = np.asarray(solid_cur).ravel() != 0
    wall = np.asarray(wall4, np.float32).reshape(n, 4)
    x, y, z = _grid_coords(p)
    sums = np.zeros((4, 4))
    counts = np.zeros(4, np.int64)
    d = np.where(~is_, fo, 0.0) - np.where(~was, fi, 0.0)
    sums[0] = [d.sum()] + [(d * C[:, a][:, None]).sum() for a in range(3)]
    counts[3] = int((was != is_).sum())
    fluid = ~is_
    far_eq = np.array([bfref_far_eq(p, q) for q in range(19)], np.float64)
    w32 = W.astype(np.float32).astype(np.float64)
    dims = (p.nx, p.ny, p.nz)
    coords = (x, y, z)
    for q in range(1, 19):
        far = np.zeros(n, bool)
        opn = np.zeros(n, bool)
        src = []
        for a in range(3):
            u = coords[a] - C[q, a]
            lo, hi = p.bc[a][0], p.bc[a][1]
            s = u.copy()
            below, above = u < 0, u >= dims[a]
            far |= (below & (lo == BC_FARFIELD)) | (above & (hi == BC_FARFIELD))
            opn |= (below & (lo == BC_OUTLET)) | (above & (hi == BC_OUTLET))
            s[below] = dims[a] - 1 if lo == BC_PERIODIC else 0
            s[above] = 0 if hi == BC_PERIODIC else dims[a] - 1
            src.append(s)
        sid = src[0] + p.nx * (src[1] + p.ny * src[2])
        out = fi[OPP[q]]
        cq = C[q].astype(np.float64)
        sel = fluid & far
        inc = far_eq[q]
        sums[1, 0] += (inc - out[sel]).sum()
        sums[1, 1:] += (inc + out[sel]).sum() * cq
        sel2 = fluid & ~far & opn
        inc2 = fi[q][sid[sel2]]
        sums[1, 0] += (inc2 - out[sel2]).sum()
        sums[1, 1:] += (inc2 + out[sel2]).sum() * cq
        counts[0] += int(sel.sum() + sel2.sum())
        sel3 = fluid & ~far & ~opn & ~was
        sel3 &= is_[np.where(sel3, sid, 0)] & was[np.where(sel3, sid, 0)]
        u = wall[sid[sel3], :3].astype(np.float64)
        corr = 2.0 * w32[q] * float(np.float32(p.far_rho)) * (u @ cq) * 3.0
        sums[3, 0] += corr.sum()
        sums[3, 1:] += (2.0 * out[sel3] + corr).sum() * cq
        counts[2] += int(sel3.sum())
    s = _sponge_factor(p, x, y, z).astype(np.float64)
    sp = fluid & (s > 0)
    if sp.any():
        inv = 1.0 / np.maximum(1.0 - s[sp], 1e-7)
        dlt = s[sp][None, :] * (far_eq[:, None] - fo[:, sp]) * inv[None, :]
        sums[2] = [dlt.sum()] + [(dlt * C[:, a][:, None]).sum() for a in range(3)]
        counts[1] = int(sp.sum())
    return sums, counts


def bfref_far_eq(p: Params, q):
    """FP32 far-field equilibrium of direction q (the value the step uses)."""
    return float(equilibrium_f32(q, float(np.float32(p.far_rho)), tuple(float(v) for v in p.far_u)))


# ---------------------------------------------------------------------------
# prescribed beta-planform flapping wing
# ---------------------------------------------------------------------------
def make_wing(**kw) -> Wing:
    w = Wing()
    for k, v in kw.items():
        if k == "root":
            w.root_x, w.root_y, w.root_z = (float(c) for c in v)
        else:
            setattr(w, k, float(v))
    return w


def li_nabawy_wing(chord_cells, root, cycle_steps, reference_speed=0.0):
    """The Li & Nabawy (2022) Re = 100 beta planform of the reference's
    prescribed-wing case (MetalFlappingWingValidation.swift:528-569,
    3449-3470): AR 3, thickness 0.05 c, beta shape 1.48915, stroke half
    amplitude 80 deg, pitch 45 / 135 deg, acceleration and pitch durations
    0.25 of a cycle, in cells and steps."""
    import math as _m
    half = 80.0 * _m.pi / 180.0
    rmax = 2 * half / (0.25 + 2 * 0.25 / _m.pi)
    return make_wing(root=root, chord_mean=chord_cells, radius=3.0 * chord_cells, thickness=0.05 * chord_cells,
                     beta_exponent=1.4891506583557845 - 1.0, beta_normalization=2.497931737168364,
                     cycle=cycle_steps, stroke_amplitude=half, stroke_duration=0.25, pitch_duration=0.25,
                     pitch_low=45.0 * _m.pi / 180.0, pitch_high=135.0 * _m.pi / 180.0, stroke_rate_max=rmax,
                     reference_speed=reference_speed)


def wing_prepare(wing: Wing, time) -> np.ndarray:
    out = np.zeros((6, 4), np.float32)
    lib().bfo_wing_prepare(ctypes.byref(wing), np.float32(time), _ptr(out))
    return out


def wing_link_fraction(wing: Wing, frame, dx, fluid, solid):
    fr = np.ascontiguousarray(frame, np.float32)
    a = np.ascontiguousarray(fluid, np.float32)
    b = np.ascontiguousarray(solid, np.float32)
    return float(lib().bfo_wing_link_fraction(ctypes.byref(wing), _ptr(fr), np.float32(dx), _ptr(a), _ptr(b)))


def wing_build(p: Params, wing: Wing, frame, solid_prev=None, velocity_to_lattice=1.0):
    n = cells(p)
    mask = np.zeros(n, np.uint8)
    wall4 = np.zeros((n, 4), np.float32)
    fr = np.ascontiguousarray(frame, np.float32)
    prev = None if solid_prev is None else np.ascontiguousarray(solid_prev, np.uint8)
    lib().bfo_wing_build(ctypes.byref(p), ctypes.byref(wing), _ptr(fr), _ptr(prev), np.float32(velocity_to_lattice),
                         _ptr(mask), _ptr(wall4))
    return mask, wall4


def wing_links(p: Params, wing: Wing, frame, mask):
    n = cells(p)
    links = np.zeros(19 * n, np.float32)
    fr = np.ascontiguousarray(frame, np.float32)
    m = np.ascontiguousarray(mask, np.uint8)
    lib().bfo_wing_links(ctypes.byref(p), ctypes.byref(wing), _ptr(fr), _ptr(m), _ptr(links))
    return links


def ray_links(p: Params, pos, tri, half_thickness, mask, wall4, winners=False):
    """Exact link fractions of the meshed shell (bfo_ray_links): [19 * N],
    q * N + cell.  winners: also the triangle giving each link's entry
    (bfo_ray_links_winners; 0xFFFFFFFF where no entry lies in the link)."""
    n = cells(p)
    links = np.zeros(19 * n, np.float32)
    pos = np.ascontiguousarray(pos, np.float32)
    tri = np.ascontiguousarray(tri, np.uint32)
    args = (ctypes.byref(p), _ptr(pos), _ptr(tri), int(tri.shape[0]), np.float32(half_thickness),
            _ptr(np.ascontiguousarray(mask, np.uint8)), _ptr(np.ascontiguousarray(wall4, np.float32)), _ptr(links))
    if not winners:
        lib().bfo_ray_links(*args)
        return links
    win = np.zeros(19 * n, np.uint32)
    lib().bfo_ray_links_winners(*args, _ptr(win))
    return links, win


def ray_shell_f64(F, D, A, B, C, h):
    """bfo_ray_shell in FP64 over rows of [n][3] arrays."""
    arr = [np.ascontiguousarray(np.asarray(v, np.float64).reshape(-1, 3)) for v in (F, D, A, B, C)]
    out = np.empty(arr[0].shape[0], np.float64)
    lib().bfo_ray_shell_f64_n(int(out.size), *[_ptr(a) for a in arr], float(h), _ptr(out))
    return out


def ray_shell(F, D, A, B, C, h):
    """First entry in [0, 1] of F + t D into the shell of radius h around
    triangle ABC, or 2."""
    arr = [np.ascontiguousarray(v, np.float32) for v in (F, D, A, B, C)]
    return float(lib().bfo_ray_shell(*[_ptr(a) for a in arr], np.float32(h)))


def step_with_links(p: Params, fin, solid_prev, solid_cur, wall4, links, capture=False):
    """One step with the BFO_WALL_INTERPOLATED_LINKS rule and this link
    table (FP32, or FP64 when given as float64)."""
    if np.asarray(links).dtype == np.float64:
        tbl = np.ascontiguousarray(links, np.float64)
        lib().bfo_set_link_table_f64(_ptr(tbl))
        try:
            return step(p, fin, solid_prev, solid_cur, wall4, capture)
        finally:
            lib().bfo_set_link_table_f64(None)
    tbl = np.ascontiguousarray(links, np.float32)
    lib().bfo_set_link_table(_ptr(tbl))
    try:
        return step(p, fin, solid_prev, solid_cur, wall4, capture)
    finally:
        lib().bfo_set_link_table(None)
