"""Host-side solver API over the C ABI.

Mirrors the reference's BirdFlowSimulation surface
(Sources/BirdFlowMetal/BirdFlowSimulation.swift): construction from a
validated configuration (:480), advance (:687), population and field copies
(:1109, :1214), the load history, and the indexed-mesh replay entry points of
MetalIndexedBirdSurfaceReplay (MetalIndexedBirdSurfaceValidation.swift:17697).
Every call goes through libbirdflow_hip.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _abi


class BirdFlowError(RuntimeError):
    """Error carrying a bf_status code (BirdFlowError.swift:3-19)."""

    def __init__(self, status: int, message: str):
        self.status = status
        self.name = _abi.STATUS_NAMES.get(status, "unknown")
        super().__init__(f"{self.name}: {message}")


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


@dataclass(frozen=True)
class LatticeScaling:
    """LatticeScaling of SimulationConfiguration.swift:137-183 in FP32."""

    cell_size: float
    time_step: float
    lattice_reference_speed: float
    lattice_viscosity: float
    tau_plus: float
    tau_minus: float
    omega_plus: float
    omega_minus: float
    velocity_to_lattice: float
    force_to_physical: float

    @classmethod
    def make(cls, characteristic_length: float, characteristic_cells: int, reference_speed: float,
             reynolds: float, air_density: float, lattice_reference_speed: float = 0.05,
             magic: float = 3.0 / 16.0, minimum_tau_plus: float = 0.50005) -> "LatticeScaling":
        f = np.float32
        if not (characteristic_length > 0 and characteristic_cells >= 8 and reference_speed > 0
                and reynolds > 0 and air_density > 0 and lattice_reference_speed > 0 and magic > 0):
            raise BirdFlowError(_abi.ERROR_INVALID_CONFIGURATION, "all physical and lattice scales must be positive")
        mach = f(lattice_reference_speed) / f(math.sqrt(1.0 / 3.0))
        if mach > 0.15:
            raise BirdFlowError(_abi.ERROR_INVALID_CONFIGURATION, f"lattice Mach {mach} above 0.15")
        dx = f(characteristic_length) / f(characteristic_cells)
        dt = f(lattice_reference_speed) * dx / f(reference_speed)
        nu = f(lattice_reference_speed) * f(characteristic_cells) / f(reynolds)
        tau_p = f(0.5) + f(3) * nu
        if tau_p < f(minimum_tau_plus):
            raise BirdFlowError(_abi.ERROR_INVALID_CONFIGURATION, f"tau+ {tau_p} too close to 1/2")
        tau_m = f(0.5) + f(magic) / (tau_p - f(0.5))
        force = f(air_density) * f(dx) ** 4 / f(dt) ** 2
        return cls(float(dx), float(dt), float(lattice_reference_speed), float(nu), float(tau_p), float(tau_m),
                   float(f(1) / tau_p), float(f(1) / tau_m), float(dt / dx), float(force))


@dataclass
class SolverConfig:
    """Configuration of one solver context (bf_config), lattice units."""

    shape: Sequence[int]
    omega_plus: float
    omega_minus: float
    storage: int = _abi.STORAGE_FP32
    collision: int = _abi.COLLISION_TRT
    wall_rule: int = _abi.WALL_HALFWAY
    force_mode: int = 6
    load_component: int = 0
    selected_part: int = 0
    accumulate_loads: bool = False
    boundary: Sequence[Sequence[int]] = ((0, 0), (0, 0), (0, 0))
    origin: Sequence[float] = (0.0, 0.0, 0.0)
    cell_size: float = 1.0
    time_step: float = 1.0
    velocity_to_lattice: float = 1.0
    force_to_physical: float = 1.0
    sponge_strength: float = 0.0
    sponge_width: float = 4.0
    far_velocity: Sequence[float] = (0.0, 0.0, 0.0)
    far_density: float = 1.0
    body_position: Sequence[float] = (0.0, 0.0, 0.0)
    arithmetic: int = _abi.ARITHMETIC_DEFAULT

    def to_c(self) -> _abi.Config:
        c = _abi.Config()
        c.nx, c.ny, c.nz = (int(v) for v in self.shape)
        c.storage, c.collision, c.wall_rule = self.storage, self.collision, self.wall_rule
        c.force_mode, c.load_component, c.selected_part = self.force_mode, self.load_component, self.selected_part
        c.accumulate_loads = int(bool(self.accumulate_loads))
        for a in range(3):
            for s in range(2):
                c.boundary[a][s] = int(self.boundary[a][s])
        c.origin[:] = [float(v) for v in self.origin]
        c.cell_size, c.time_step = self.cell_size, self.time_step
        c.velocity_to_lattice, c.force_to_physical = self.velocity_to_lattice, self.force_to_physical
        c.omega_plus, c.omega_minus = self.omega_plus, self.omega_minus
        c.sponge_strength, c.sponge_width = self.sponge_strength, self.sponge_width
        c.far_velocity[:] = [float(v) for v in self.far_velocity]
        c.far_density = self.far_density
        c.body_position[:] = [float(v) for v in self.body_position]
        c.arithmetic = self.arithmetic
        return c

    @property
    def cells(self) -> int:
        return int(self.shape[0]) * int(self.shape[1]) * int(self.shape[2])


def device_info():
    lib = _abi.load()
    name = ctypes.create_string_buffer(256)
    total = ctypes.c_uint64(0)
    st = lib.bf_device_info(name, 256, ctypes.byref(total))
    if st != _abi.OK:
        raise BirdFlowError(st, lib.bf_status_string(st).decode())
    return name.value.decode(), int(total.value)


def copy_bandwidth(bytes_per_buffer: int = 1 << 30, iterations: int = 20) -> float:
    """Measured device copy bandwidth in GB/s (read + write)."""
    lib = _abi.load()
    out = ctypes.c_double(0.0)
    st = lib.bf_copy_bandwidth(int(bytes_per_buffer), int(iterations), ctypes.byref(out))
    if st != _abi.OK:
        raise BirdFlowError(st, lib.bf_status_string(st).decode())
    return float(out.value)


@dataclass
class BodyParameters:
    """Rigid-body parameters (bf_body_params), world units."""

    mass: float
    inertia: Sequence[float]
    gravity: Sequence[float] = (0.0, 0.0, 0.0)
    external_force: Sequence[float] = (0.0, 0.0, 0.0)
    external_torque: Sequence[float] = (0.0, 0.0, 0.0)
    safety_extent: Sequence[float] = (0.0, 0.0, 0.0)
    safety_radius: float = 0.0
    wing_speed: float = 0.0
    mach_limit: float = 0.0          # <= 0: no safety monitor
    substeps: int = 1
    free_flight: bool = True
    move_mesh: bool = False

    def to_c(self) -> _abi.BodyParams:
        b = _abi.BodyParams()
        b.mass = float(self.mass)
        for name in ("inertia", "gravity", "external_force", "external_torque", "safety_extent"):
            getattr(b, name)[:] = [float(v) for v in getattr(self, name)]
        b.safety_radius, b.wing_speed, b.mach_limit = self.safety_radius, self.wing_speed, self.mach_limit
        b.substeps = int(self.substeps)
        b.free_flight, b.move_mesh = int(bool(self.free_flight)), int(bool(self.move_mesh))
        return b


@dataclass
class FlappingWing:
    """GPUFlappingWingParameters (GPUData.swift:383-388) in grid world units
    and time units: root, mean chord, span radius, thickness, beta exponent
    (alpha - 1) and normalization, cycle, stroke amplitude, stroke and pitch
    transition durations (cycle fractions), mid-stroke pitch (downstroke,
    upstroke), maximum stroke rate (rad per cycle), reference speed."""
    root: tuple
    chord_mean: float
    radius: float
    thickness: float
    beta_exponent: float
    beta_normalization: float
    cycle: float
    stroke_amplitude: float
    stroke_duration: float
    pitch_duration: float
    pitch_low: float
    pitch_high: float
    stroke_rate_max: float
    reference_speed: float = 0.0

    @classmethod
    def li_nabawy(cls, chord_cells: float, root, cycle_steps: float, reference_speed: float = 0.0):
        """The Li & Nabawy (2022) Re 100 case of MetalFlappingWingValidation.swift:528-569:
        aspect ratio 3, thickness 0.05 c, beta shape alpha 1.48915 with the
        normalization that makes the mean chord c, stroke half amplitude 80
        deg, pitch 45 / 135 deg, acceleration and pitch durations of a
        quarter cycle; the stroke rate's plateau solves stroke continuity."""
        half = math.radians(80.0)
        return cls(root=tuple(float(v) for v in root), chord_mean=chord_cells, radius=3.0 * chord_cells,
                   thickness=0.05 * chord_cells, beta_exponent=1.4891506583557845 - 1.0,
                   beta_normalization=2.497931737168364, cycle=cycle_steps, stroke_amplitude=half,
                   stroke_duration=0.25, pitch_duration=0.25, pitch_low=math.radians(45.0),
                   pitch_high=math.radians(135.0), stroke_rate_max=2.0 * half / (0.25 + 0.5 / math.pi),
                   reference_speed=reference_speed)

    def values(self):
        return [*self.root, self.chord_mean, self.radius, self.thickness, self.beta_exponent,
                self.beta_normalization, self.cycle, self.stroke_amplitude, self.stroke_duration,
                self.pitch_duration, self.pitch_low, self.pitch_high, self.stroke_rate_max, self.reference_speed]

    def to_c(self) -> _abi.FlappingWingC:
        return _abi.FlappingWingC(*[float(v) for v in self.values()])


@dataclass
class BirdGeometry:
    """BirdParameters geometry and analytic wing kinematics
    (BirdModel.swift:129-217): metres, radians, Hz."""

    body_radii: Sequence[float]
    wing_span: float
    wing_root_chord: float
    wing_tip_chord: float
    wing_thickness: float
    wing_sweep: float
    wing_root_offset: Sequence[float]
    tail_length: float
    tail_half_width: float
    tail_thickness: float
    frequency: float
    stroke_amplitude: float
    stroke_bias: float = 0.0
    pitch_mean: float = 0.0
    pitch_amplitude: float = 0.0
    pitch_phase: float = 0.0

    @classmethod
    def demonstration(cls) -> "BirdGeometry":
        """BirdParameters.demonstration (BirdModel.swift:198-217)."""
        r = math.radians
        return cls(body_radii=(0.18, 0.065, 0.075), wing_span=0.34, wing_root_chord=0.14, wing_tip_chord=0.055,
                   wing_thickness=0.018, wing_sweep=0.055, wing_root_offset=(0.02, 0.055, 0.02), tail_length=0.17,
                   tail_half_width=0.11, tail_thickness=0.018, frequency=4.5, stroke_amplitude=r(42),
                   stroke_bias=r(2), pitch_mean=r(8), pitch_amplitude=r(24), pitch_phase=r(78))

    def to_c(self) -> _abi.BirdGeometry:
        b = _abi.BirdGeometry()
        for name, _ in _abi.BirdGeometry._fields_:
            v = getattr(self, name)
            if name in ("body_radii", "wing_root_offset"):
                getattr(b, name)[:] = [float(x) for x in v]
            else:
                setattr(b, name, float(v))
        return b


@dataclass
class Joint:
    """One articulated joint: side (+1 / -1 mirrors y), hinge of the +y side
    and the Fourier series [harmonic][sin, cos] (radians) of its stroke,
    deviation, pitch and tip twist."""

    side: float
    root: Sequence[float]
    stroke: np.ndarray
    deviation: np.ndarray
    pitch: np.ndarray
    twist: np.ndarray

    ANGLES = ("stroke", "deviation", "pitch", "twist")


@dataclass(frozen=True)
class KinematicParameter:
    """One Fourier coefficient of an articulated mesh's kinematics as an
    adjoint parameter: an offset added to coefficient [harmonic][0 for "sin",
    1 for "cos"] of series `channel` ("stroke", "deviation", "pitch",
    "twist") of every joint in `joints` (0-based indices into
    ArticulatedMesh.joints; one parameter can move both wings, its
    derivative sums over them)."""

    joints: tuple
    channel: str
    harmonic: int = 0
    term: str = "sin"

    def to_c(self) -> _abi.KinematicParameter:
        if self.channel not in Joint.ANGLES:
            raise ValueError(f"channel is one of {Joint.ANGLES}")
        if self.term not in ("sin", "cos"):
            raise ValueError("term is 'sin' or 'cos'")
        mask = 0
        for j in self.joints:
            mask |= 1 << int(j)
        return _abi.KinematicParameter(mask, Joint.ANGLES.index(self.channel), int(self.harmonic),
                                       0 if self.term == "sin" else 1)


@dataclass
class ArticulatedMesh:
    """An articulated mesh for bf_mesh_upload_articulated: rest positions
    relative to `center`, each vertex's joint (0 rigid, k = joints[k - 1]),
    the twist weight of each vertex, the triangles and their part ids, and
    the Fourier kinematics of the joints (period in time units, one amplitude
    scale for every angle)."""

    rest: np.ndarray
    vertex_joint: np.ndarray
    span_fraction: np.ndarray
    triangles: np.ndarray
    parts: np.ndarray
    half_thickness: float
    period: float
    amplitude_scale: float
    center: Sequence[float]
    joints: list
    harmonics: int = 3

    def to_c(self) -> _abi.Kinematics:
        k = _abi.Kinematics()
        k.period, k.amplitude_scale = float(self.period), float(self.amplitude_scale)
        k.harmonics, k.joint_count = int(self.harmonics), len(self.joints)
        k.center[:] = [float(v) for v in self.center]
        for i, j in enumerate(self.joints):
            J = k.joints[i]
            J.side = float(j.side)
            J.root[:] = [float(v) for v in j.root]
            for name in Joint.ANGLES:
                c = np.asarray(getattr(j, name), np.float32).reshape(-1, 2)
                arr = getattr(J, name)
                for h in range(c.shape[0]):
                    arr[h][0], arr[h][1] = float(c[h, 0]), float(c[h, 1])
        return k

    def kinematics_dict(self) -> dict:
        """The kinematics as plain values (checkpoint manifests), each rounded
        to the FP32 value the library uses, so that a restored mesh writes
        the same manifest again."""
        f32 = lambda v: float(np.float32(v))
        return {"period": f32(self.period), "amplitude_scale": f32(self.amplitude_scale),
                "harmonics": int(self.harmonics), "center": [f32(v) for v in self.center],
                "half_thickness": f32(self.half_thickness),
                "joints": [{"side": f32(j.side), "root": [f32(v) for v in j.root],
                            **{n: np.asarray(getattr(j, n), np.float32).astype(np.float64).tolist()
                               for n in Joint.ANGLES}}
                           for j in self.joints]}

    def perturbed(self, parameters, offsets) -> "ArticulatedMesh":
        """A copy with offsets[k] added to KinematicParameter k's coefficient
        in every joint it names (kept in FP64 until the upload rounds them):
        the function whose derivative at offsets = 0 the adjoint returns."""
        joints = [dataclasses.replace(j, **{n: np.array(getattr(j, n), np.float64).reshape(-1, 2)
                                            for n in Joint.ANGLES}) for j in self.joints]
        for q, d in zip(parameters, np.asarray(offsets, np.float64)):
            for j in q.joints:
                getattr(joints[j], q.channel)[q.harmonic, 0 if q.term == "sin" else 1] += d
        return dataclasses.replace(self, joints=joints)

    @classmethod
    def from_arrays(cls, arrays: dict, kin: dict) -> "ArticulatedMesh":
        joints = [Joint(j["side"], j["root"], *(np.asarray(j[n], np.float32) for n in Joint.ANGLES))
                  for j in kin["joints"]]
        return cls(arrays["rest"], arrays["vertex_joint"], arrays["span_fraction"], arrays["triangles"],
                   arrays["triangle_parts"], kin["half_thickness"], kin["period"], kin["amplitude_scale"],
                   kin["center"], joints, kin["harmonics"])


def _body_to_c(state) -> _abi.BodyState:
    """13-vector (position, orientation x y z w, linear velocity, body angular velocity)."""
    v = np.asarray(state, np.float64).ravel()
    if v.size != 13:
        raise ValueError("a body state has 13 components")
    b = _abi.BodyState()
    b.position[:] = v[0:3].tolist()
    b.orientation[:] = v[3:7].tolist()
    b.linear_velocity[:] = v[7:10].tolist()
    b.angular_velocity_body[:] = v[10:13].tolist()
    return b


def _body_from_c(b) -> np.ndarray:
    return np.array(list(b.position) + list(b.orientation) + list(b.linear_velocity)
                    + list(b.angular_velocity_body), np.float64)


class BirdFlowSimulation:
    """One solver context on the current HIP device."""

    def __init__(self, config: SolverConfig):
        self.config = config
        self._lib = _abi.load()
        self._cfg = config.to_c()
        handle = ctypes.c_void_p()
        st = self._lib.bf_create(ctypes.byref(self._cfg), ctypes.byref(handle))
        if st != _abi.OK:
            raise BirdFlowError(st, self._lib.bf_status_string(st).decode())
        self._ctx = handle
        self.n = config.cells
        # inputs of the geometry source and body parameters, kept for checkpoints
        self.geometry_source = None
        self.body_parameters = None

    # -- lifetime --------------------------------------------------------
    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.bf_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st):
        if st != _abi.OK:
            msg = self._lib.bf_last_error(self._ctx)
            raise BirdFlowError(st, (msg or b"").decode() or self._lib.bf_status_string(st).decode())

    # -- state -----------------------------------------------------------
    def init_equilibrium(self):
        self._check(self._lib.bf_init_equilibrium(self._ctx))

    def init_shear_wave(self, amplitude: float):
        self._check(self._lib.bf_init_shear_wave(self._ctx, float(amplitude)))

    def write_populations(self, f: np.ndarray):
        f = np.ascontiguousarray(f, np.float32).ravel()
        self._check(self._lib.bf_write_populations(self._ctx, _ptr(f), f.size))

    def read_populations(self) -> np.ndarray:
        out = np.empty(19 * self.n, np.float32)
        self._check(self._lib.bf_read_populations(self._ctx, _ptr(out), out.size))
        return out

    @property
    def storage_dtype(self):
        """Element type of the stored form: float32 populations, or the
        float16 / bfloat16 bits (uint16) of the offsets f_q - w_q."""
        return np.float32 if self.config.storage == _abi.STORAGE_FP32 else np.uint16

    def write_population_storage(self, raw: np.ndarray):
        """The stored form itself, direction-major (bit-exact for every storage)."""
        r = np.ascontiguousarray(raw, self.storage_dtype).ravel()
        self._check(self._lib.bf_write_population_storage(self._ctx, _ptr(r), r.nbytes))

    def read_population_storage(self) -> np.ndarray:
        out = np.empty(19 * self.n, self.storage_dtype)
        self._check(self._lib.bf_read_population_storage(self._ctx, _ptr(out), out.nbytes))
        return out

    def restart_state(self):
        """(completed step, mesh clock origin): mesh time of step s is
        origin + s * time_step."""
        step = ctypes.c_uint64(0)
        origin = ctypes.c_float(0.0)
        self._check(self._lib.bf_read_restart_state(self._ctx, ctypes.byref(step), ctypes.byref(origin)))
        return int(step.value), float(origin.value)

    def set_restart_state(self, step: int, mesh_time_origin: float, last_load=None, solid_mask=None):
        """Renumbers the completed step, sets the mesh clock origin, the last
        reduced load (8 values) and the solid mask the next step starts from,
        and rebuilds the current step's geometry.  Call after the geometry
        source and body are set, before the populations are written."""
        load = None if last_load is None else np.ascontiguousarray(last_load, np.float64).ravel()
        if load is not None and load.size != 8:
            raise ValueError("last_load has 8 components")
        mask = None if solid_mask is None else np.ascontiguousarray(solid_mask, np.uint8).ravel()
        if mask is not None and mask.size != self.n:
            raise ValueError("solid_mask has one byte per cell")
        self._check(self._lib.bf_set_restart_state(self._ctx, int(step), float(mesh_time_origin), _ptr(load),
                                                   _ptr(mask)))

    def copy_macroscopic_fields(self):
        rho = np.empty(self.n, np.float32)
        u = np.empty((self.n, 3), np.float32)
        self._check(self._lib.bf_read_fields(self._ctx, _ptr(rho), _ptr(u)))
        return rho, u

    @property
    def step_index(self) -> int:
        return int(self._lib.bf_step_index(self._ctx))

    @property
    def allocated_bytes(self) -> int:
        return int(self._lib.bf_allocated_bytes(self._ctx))

    # -- geometry --------------------------------------------------------
    def set_geometry(self, mask_cur, wall4, mask_prev=None):
        mc = np.ascontiguousarray(mask_cur, np.uint8).ravel()
        w = np.ascontiguousarray(wall4, np.float32).reshape(-1, 4)
        mp = None if mask_prev is None else np.ascontiguousarray(mask_prev, np.uint8).ravel()
        self._check(self._lib.bf_set_geometry(self._ctx, _ptr(mp), _ptr(mc), _ptr(w)))
        # the current mask and wall are read back from the device when saved;
        # a distinct previous mask matters only until the next step
        self.geometry_source = {"kind": "static", "mask_prev": None if mp is None else mp.copy(),
                                "set_at_step": self.step_index}

    def read_geometry(self):
        mask = np.empty(self.n, np.uint8)
        wall = np.empty((self.n, 4), np.float32)
        self._check(self._lib.bf_read_geometry(self._ctx, _ptr(mask), _ptr(wall)))
        return mask, wall

    def upload_mesh(self, points, frame_times, triangles, triangle_parts, half_thickness,
                    translation=(0.0, 0.0, 0.0), time_start: Optional[float] = None):
        pts = np.ascontiguousarray(points, np.float32)
        frames, verts = pts.shape[0], pts.shape[1]
        times = np.ascontiguousarray(frame_times, np.float32)
        tri = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
        parts = np.ascontiguousarray(triangle_parts, np.uint8)
        tr = np.ascontiguousarray(translation, np.float32)
        t0 = float(times[0]) if time_start is None else float(time_start)
        self._check(self._lib.bf_mesh_upload(self._ctx, _ptr(pts), _ptr(times), frames, verts, _ptr(tri), _ptr(parts),
                                             tri.shape[0], float(half_thickness), _ptr(tr), t0))
        self._parameters = 0          # the upload clears the adjoint's tangents and parameters
        self.geometry_source = {"kind": "mesh", "points": pts.copy(), "frame_times": times.copy(),
                                "triangles": tri.copy(), "triangle_parts": parts.copy(),
                                "half_thickness": float(half_thickness), "translation": tr.copy()}

    def build_mesh_geometry(self, time: float):
        self._check(self._lib.bf_mesh_build(self._ctx, float(time)))

    def upload_articulated(self, mesh: ArticulatedMesh, time_start: float = 0.0):
        """Articulated mesh with on-device Fourier kinematics
        (bf_mesh_upload_articulated): no frames; each step prepares the
        vertices at its own time."""
        rest = np.ascontiguousarray(mesh.rest, np.float32).reshape(-1, 3)
        joint = np.ascontiguousarray(mesh.vertex_joint, np.uint8)
        span = np.ascontiguousarray(mesh.span_fraction, np.float32)
        tri = np.ascontiguousarray(mesh.triangles, np.uint32).reshape(-1, 3)
        parts = np.ascontiguousarray(mesh.parts, np.uint8)
        k = mesh.to_c()
        self._check(self._lib.bf_mesh_upload_articulated(self._ctx, _ptr(rest), _ptr(joint), _ptr(span), rest.shape[0],
                                                         _ptr(tri), _ptr(parts), tri.shape[0],
                                                         float(mesh.half_thickness), ctypes.byref(k),
                                                         float(time_start)))
        self._parameters = 0
        self.geometry_source = {"kind": "articulated", "rest": rest.copy(), "vertex_joint": joint.copy(),
                                "span_fraction": span.copy(), "triangles": tri.copy(),
                                "triangle_parts": parts.copy(), "kinematics": mesh.kinematics_dict()}

    def mesh_vertices_at(self, time: float, vertex_count: int):
        """Prepared vertex positions and lattice velocities [V][3] at `time`."""
        pos = np.zeros((vertex_count, 3), np.float32)
        vel = np.zeros((vertex_count, 3), np.float32)
        self._check(self._lib.bf_mesh_vertices_at(self._ctx, float(time), _ptr(pos), _ptr(vel), vertex_count))
        return pos, vel

    def set_translating_sphere(self, center_radius, geometry_velocity, wall_velocity):
        cr = np.ascontiguousarray(center_radius, np.float32)
        gv = np.ascontiguousarray(geometry_velocity, np.float32)
        wv = np.ascontiguousarray(wall_velocity, np.float32)
        self._check(self._lib.bf_set_translating_sphere(self._ctx, _ptr(cr), _ptr(gv), _ptr(wv)))
        self.geometry_source = {"kind": "sphere", "center_radius": cr.tolist(), "geometry_velocity": gv.tolist(),
                                "wall_velocity": wv.tolist()}

    def set_bird(self, bird: "BirdGeometry", keyframes=None):
        """Analytic signed-distance bird (BirdFlow.metal:686-938) flying on the
        body pose; builds the current step.  keyframes: measured wing
        kinematics, [count, 17] (phase, left angles stroke / deviation / pitch
        / tip twist, left rates, right angles, right rates), or None for the
        analytic sinusoid."""
        kf = None if keyframes is None else np.ascontiguousarray(keyframes, np.float32).reshape(-1, 17)
        self._check(self._lib.bf_set_bird_keyframes(self._ctx, _ptr(kf), 0 if kf is None else kf.shape[0]))
        cb = bird.to_c()
        self._check(self._lib.bf_set_bird(self._ctx, ctypes.byref(cb)))
        self.geometry_source = {"kind": "bird", "bird": dataclasses.asdict(bird),
                                "keyframes": None if kf is None else kf.copy()}

    def set_flapping_wing(self, wing: "FlappingWing"):
        """Prescribed beta-planform wing (BirdFlow.metal:940-1297) with its
        link table; builds the current step.  Pair it with wall_rule
        WALL_INTERPOLATED_LINKS."""
        cw = wing.to_c()
        self._check(self._lib.bf_set_flapping_wing(self._ctx, ctypes.byref(cw)))
        self.geometry_source = {"kind": "wing", "wing": dataclasses.asdict(wing)}

    def wing_frame(self) -> np.ndarray:
        """Prepared frame of the current step, [6, 4] (GPUPreparedFlappingWing rows)."""
        out = np.zeros((6, 4), np.float32)
        self._check(self._lib.bf_read_wing_frame(self._ctx, _ptr(out)))
        return out

    def wing_links(self) -> np.ndarray:
        """The current step's link table, [19, N]: the fluid-side fraction of
        each fluid cell whose source is on the wing, 0 elsewhere."""
        out = np.zeros(19 * self.n, np.float32)
        self._check(self._lib.bf_read_wing_links(self._ctx, _ptr(out), out.size))
        return out.reshape(19, -1)

    def ray_links(self) -> np.ndarray:
        """The current step's exact-root link table (wall rule
        WALL_INTERPOLATED_RAY with a mesh), [19, N]: the clamped fraction of
        each fluid cell's link whose source is solid, 0 elsewhere."""
        out = np.zeros(19 * self.n, np.float32)
        self._check(self._lib.bf_read_ray_links(self._ctx, _ptr(out), out.size))
        return out.reshape(19, -1)

    def population_minimum(self):
        """(value, direction-major index, non_finite) of the smallest current
        population, a non-finite value ranking first (reducePopulationMinimum)."""
        v = ctypes.c_float()
        i = ctypes.c_uint64()
        nf = ctypes.c_uint32()
        self._check(self._lib.bf_population_minimum(self._ctx, ctypes.byref(v), ctypes.byref(i), ctypes.byref(nf)))
        return float(v.value), int(i.value), bool(nf.value)

    def inject_population(self, index: int, value: float):
        """Fault injection: population `index` (q N + cell) set to `value`."""
        self._check(self._lib.bf_inject_population(self._ctx, int(index), float(value)))

    def set_wing_dynamics(self, left=None, right=None):
        """Prescribed-wing dynamics of the analytic bird: each wing a
        (mass, centre of mass from the hinge along chord / span / normal,
        principal inertia) triple, or both None to clear."""
        if (left is None) != (right is None):
            raise ValueError("give both wings' dynamics, or neither")

        def c(w):
            d = _abi.WingDynamicsC()
            d.mass = float(w[0])
            d.center_of_mass[:] = [float(v) for v in w[1]]
            d.inertia[:] = [float(v) for v in w[2]]
            return d
        if left is None:
            self._check(self._lib.bf_set_bird_wing_dynamics(self._ctx, None, None))
        else:
            cl, cr = c(left), c(right)
            self._check(self._lib.bf_set_bird_wing_dynamics(self._ctx, ctypes.byref(cl), ctypes.byref(cr)))

    def wing_reaction(self) -> np.ndarray:
        """The current step's hinge reactions [4, 4]: left force, left torque,
        right force, right torque."""
        out = np.zeros((4, 4), np.float32)
        self._check(self._lib.bf_read_wing_reaction(self._ctx, _ptr(out)))
        return out

    def bird_frame(self) -> np.ndarray:
        """Prepared frame of the current step, [14, 4] (GPUPreparedBirdGeometry rows)."""
        out = np.zeros((14, 4), np.float32)
        self._check(self._lib.bf_read_bird_frame(self._ctx, _ptr(out)))
        return out

    # -- momentum ledger -------------------------------------------------
    def fluid_momentum(self) -> np.ndarray:
        """Mass and momentum (x, y, z) of the fluid populations, lattice units, FP64."""
        out = np.zeros(4, np.float64)
        self._check(self._lib.bf_ledger_fluid(self._ctx, _ptr(out)))
        return out

    def step_momentum_sources(self, allocate: bool = True) -> dict:
        """Momentum sums of the step just completed (lattice units, FP64):
        fluid change, open-boundary, sponge and persistent-link sources as
        (mass, px, py, pz), and the link / cell counts.  allocate=False
        goes through bf_ledger_step_sources_noalloc (the same sums from the
        library's fixed device scratch, no device allocation)."""
        sums = np.zeros(16, np.float64)
        counts = np.zeros(4, np.uint64)
        fn = self._lib.bf_ledger_step_sources if allocate else self._lib.bf_ledger_step_sources_noalloc
        self._check(fn(self._ctx, _ptr(sums), _ptr(counts)))
        return {"change": sums[0:4].copy(), "open_boundary": sums[4:8].copy(), "sponge": sums[8:12].copy(),
                "persistent_link": sums[12:16].copy(), "open_boundary_links": int(counts[0]),
                "sponge_cells": int(counts[1]), "persistent_links": int(counts[2]),
                "topology_transitions": int(counts[3])}

    # -- stepping --------------------------------------------------------
    def advance(self, steps: int, synchronize: bool = True):
        if steps < 0:
            raise BirdFlowError(_abi.ERROR_INVALID_ADVANCE_REQUEST, f"steps must be >= 0, got {steps}")
        self._check(self._lib.bf_step(self._ctx, int(steps)))
        if synchronize:
            self.synchronize()

    def synchronize(self):
        self._check(self._lib.bf_synchronize(self._ctx))

    def read_loads(self) -> np.ndarray:
        out = np.zeros(8, np.float64)
        self._check(self._lib.bf_read_loads(self._ctx, _ptr(out)))
        return out

    def load_history(self, max_records: int = 1 << 16) -> np.ndarray:
        out = np.zeros((max_records, 8), np.float64)
        n = ctypes.c_uint32(0)
        self._check(self._lib.bf_read_load_history(self._ctx, _ptr(out), max_records, ctypes.byref(n)))
        return out[: n.value].copy()

    # -- rigid body ------------------------------------------------------
    def set_body(self, state=None, params: Optional[BodyParameters] = None):
        """Body state as a 13-vector (position, orientation x y z w, linear
        velocity, body angular velocity) and/or parameters."""
        cs = _body_to_c(state) if state is not None else None
        cp = params.to_c() if params is not None else None
        self._check(self._lib.bf_body_set(self._ctx, ctypes.byref(cs) if cs is not None else None,
                                          ctypes.byref(cp) if cp is not None else None))
        if params is not None:
            self.body_parameters = params

    def body_state(self) -> np.ndarray:
        b = _abi.BodyState()
        self._check(self._lib.bf_body_get(self._ctx, ctypes.byref(b)))
        return _body_from_c(b)

    def body_history(self, max_records: int = 1 << 16) -> np.ndarray:
        arr = (_abi.BodyState * max_records)()
        n = ctypes.c_uint32(0)
        self._check(self._lib.bf_read_body_history(self._ctx, arr, max_records, ctypes.byref(n)))
        return np.array([_body_from_c(arr[i]) for i in range(n.value)], np.float64).reshape(-1, 13)

    def safety(self):
        """(metrics [max Mach, min clearance, Mach, clearance at the first event], event step, flags)."""
        m = np.zeros(4, np.float64)
        step = ctypes.c_uint64(0)
        flags = ctypes.c_uint32(0)
        self._check(self._lib.bf_read_safety(self._ctx, _ptr(m), ctypes.byref(step), ctypes.byref(flags)))
        return m, int(step.value), int(flags.value)

    # -- adjoint ---------------------------------------------------------
    def adjoint_set_tangents(self, tangents):
        """d(mesh frame vertex positions)/d theta, [K, frames, vertices, 3] (K <= 8)."""
        t = np.ascontiguousarray(tangents, np.float32)
        k = 0 if t.size == 0 else t.shape[0]
        self._check(self._lib.bf_adjoint_set_tangents(self._ctx, k, _ptr(t) if k else None))
        self._parameters = k

    def adjoint_set_kinematic_parameters(self, parameters):
        """Fourier coefficients of the uploaded articulated mesh as the
        adjoint's parameters (a sequence of KinematicParameter, at most 8; the
        device differentiates the kinematics, no tangents are uploaded)."""
        ps = list(parameters)
        arr = (_abi.KinematicParameter * max(len(ps), 1))(*[q.to_c() for q in ps])
        self._check(self._lib.bf_adjoint_set_kinematic_parameters(self._ctx, len(ps), arr if ps else None))
        self._parameters = len(ps)

    def adjoint_gradient(self, steps: int, load_weights, checkpoint_interval: int = 100):
        """Runs `steps` steps and the reverse sweep for J = sum_s w[s] . L_s
        (load_weights [steps, 6]: force xyz, torque xyz).  Returns
        (J, dJ/dtheta [K], {forward_ms, reverse_ms})."""
        w = np.ascontiguousarray(load_weights, np.float64).reshape(int(steps), 6)
        k = getattr(self, "_parameters", 0)
        grad = np.zeros(max(k, 1), np.float64)
        timings = np.zeros(2, np.float64)
        J = ctypes.c_double(0.0)
        self._check(self._lib.bf_adjoint_gradient(self._ctx, int(steps), _ptr(w), int(checkpoint_interval),
                                                  ctypes.byref(J), _ptr(grad), _ptr(timings)))
        return J.value, grad[:k].copy(), {"forward_ms": float(timings[0]), "reverse_ms": float(timings[1])}

    def adjoint_sensitivity(self) -> np.ndarray:
        """dJ/df at the state the last adjoint sweep started from (19 x cells)."""
        out = np.empty(19 * self.config.cells, np.float32)
        self._check(self._lib.bf_adjoint_read_sensitivity(self._ctx, _ptr(out), out.size))
        return out

    def freeze_masks(self, masks):
        """Replay recorded part masks ([steps, cells] uint8) from the current
        step on; None ends the replay."""
        if masks is None:
            self._check(self._lib.bf_mesh_freeze_masks(self._ctx, None, 0))
            return
        m = np.ascontiguousarray(masks, np.uint8).reshape(-1, self.config.cells)
        self._check(self._lib.bf_mesh_freeze_masks(self._ctx, _ptr(m), m.shape[0]))

    def link_census(self):
        out = np.zeros(21, np.uint64)
        self._check(self._lib.bf_link_census(self._ctx, _ptr(out)))
        return out[:19].copy(), int(out[19]), int(out[20])

    def set_profiling(self, enabled=True, level=None):
        """Per-step HIP events: level 1 around the fast-fluid kernel only, 2
        also around the general kernel and the geometry (default 2 when
        enabled)."""
        lvl = 0 if not enabled else (2 if level is None else int(level))
        self._check(self._lib.bf_set_profiling(self._ctx, lvl))

    def read_profile(self):
        out = np.zeros(6, np.float64)
        self._check(self._lib.bf_read_profile(self._ctx, _ptr(out)))
        return {"geometry_ms": out[0], "fast_ms": out[1], "general_ms": out[2], "steps": int(out[3]),
                "general_tiles": int(out[4]), "tiles": int(out[5])}
