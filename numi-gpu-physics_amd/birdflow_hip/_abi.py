"""ctypes declarations of libbirdflow_hip.so (include/birdflow_hip.h)."""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "libbirdflow_hip.so")

# bf_status (BirdFlowError.swift:3-19)
OK = 0
STATUS_NAMES = {
    0: "ok",
    1: "deviceUnavailable",
    2: "invalidConfiguration",
    3: "allocationFailed",
    4: "bufferExceedsDeviceLimit",
    5: "workingSetExceedsMemory",
    6: "invalidAdvanceRequest",
    7: "invalidArgument",
    8: "launchFailed",
    9: "runtimeSafetyViolation",
    10: "simulationStateInvalidated",
    11: "notSupported",
}
ERROR_DEVICE_UNAVAILABLE = 1
ERROR_INVALID_CONFIGURATION = 2
ERROR_INVALID_ADVANCE_REQUEST = 6
ERROR_INVALID_ARGUMENT = 7
ERROR_RUNTIME_SAFETY_VIOLATION = 9
ERROR_STATE_INVALIDATED = 10

BC_PERIODIC, BC_FARFIELD, BC_OUTLET = 0, 1, 2
COLLISION_TRT, COLLISION_TRT_SYMMETRIC_LIMITER, COLLISION_REGULARIZED, COLLISION_RR3 = 0, 1, 2, 3
WALL_HALFWAY, WALL_INTERPOLATED_SDF, WALL_INTERPOLATED_HALF, WALL_INTERPOLATED_LINKS, WALL_INTERPOLATED_RAY = 0, 1, 2, 3, 4
STORAGE_FP32, STORAGE_FP16_OFFSET, STORAGE_BF16_OFFSET = 0, 1, 2
GEOMETRY_STATIC, GEOMETRY_MESH, GEOMETRY_SPHERE, GEOMETRY_BIRD, GEOMETRY_WING = 0, 1, 2, 3, 4
ARITHMETIC_DEFAULT, ARITHMETIC_STRICT, ARITHMETIC_FUSED = 0, 1, 2

EXPORTED_FUNCTIONS = (
    "bf_create", "bf_destroy", "bf_last_error", "bf_status_string", "bf_abi_version", "bf_device_info",
    "bf_allocated_bytes", "bf_init_equilibrium", "bf_init_shear_wave", "bf_write_populations",
    "bf_read_populations", "bf_write_population_storage", "bf_read_population_storage", "bf_read_fields",
    "bf_step_index", "bf_read_restart_state", "bf_set_restart_state",
    "bf_set_geometry", "bf_read_geometry",
    "bf_mesh_upload", "bf_mesh_build", "bf_mesh_upload_articulated", "bf_mesh_vertices_at",
    "bf_set_translating_sphere", "bf_step", "bf_synchronize",
    "bf_read_loads", "bf_read_load_history", "bf_link_census", "bf_set_profiling", "bf_read_profile",
    "bf_copy_bandwidth", "bf_body_set", "bf_body_get", "bf_read_body_history", "bf_read_safety",
    "bf_population_minimum", "bf_inject_population",
    "bf_adjoint_set_tangents", "bf_adjoint_set_kinematic_parameters", "bf_adjoint_gradient", "bf_adjoint_read_sensitivity", "bf_mesh_freeze_masks",
    "bf_set_bird", "bf_set_bird_keyframes", "bf_read_bird_frame", "bf_set_bird_wing_dynamics",
    "bf_read_wing_reaction", "bf_ledger_fluid", "bf_ledger_step_sources",
    "bf_ledger_step_sources_noalloc", "bf_set_flapping_wing", "bf_read_wing_frame", "bf_read_wing_links",
    "bf_read_ray_links",
)


class Joint(ctypes.Structure):
    """bf_joint: side, hinge, Fourier series [harmonic][sin, cos] per angle."""
    _fields_ = [("side", ctypes.c_float), ("root", ctypes.c_float * 3)] + [
        (name, (ctypes.c_float * 2) * 4) for name in ("stroke", "deviation", "pitch", "twist")]


class Kinematics(ctypes.Structure):
    """bf_kinematics: period, amplitude scale, harmonics, centre, joints."""
    _fields_ = [("period", ctypes.c_float), ("amplitude_scale", ctypes.c_float), ("harmonics", ctypes.c_uint32),
                ("joint_count", ctypes.c_uint32), ("center", ctypes.c_float * 3), ("reserved", ctypes.c_float),
                ("joints", Joint * 8)]


class KinematicParameter(ctypes.Structure):
    """bf_kinematic_parameter: joint bitmask, series (0 stroke .. 3 twist),
    harmonic, component (0 sin, 1 cos)."""
    _fields_ = [("joint_mask", ctypes.c_uint32), ("channel", ctypes.c_uint32), ("harmonic", ctypes.c_uint32),
                ("component", ctypes.c_uint32)]


class Config(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_uint32), ("ny", ctypes.c_uint32), ("nz", ctypes.c_uint32),
        ("storage", ctypes.c_uint32),
        ("collision", ctypes.c_uint32),
        ("wall_rule", ctypes.c_uint32),
        ("force_mode", ctypes.c_uint32),
        ("load_component", ctypes.c_uint32),
        ("selected_part", ctypes.c_uint32),
        ("accumulate_loads", ctypes.c_uint32),
        ("boundary", (ctypes.c_uint32 * 2) * 3),
        ("origin", ctypes.c_float * 3),
        ("cell_size", ctypes.c_float),
        ("time_step", ctypes.c_float),
        ("velocity_to_lattice", ctypes.c_float),
        ("force_to_physical", ctypes.c_float),
        ("omega_plus", ctypes.c_float), ("omega_minus", ctypes.c_float),
        ("sponge_strength", ctypes.c_float),
        ("sponge_width", ctypes.c_float),
        ("far_velocity", ctypes.c_float * 3),
        ("far_density", ctypes.c_float),
        ("body_position", ctypes.c_float * 3),
        ("arithmetic", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32 * 7),
    ]


class BodyState(ctypes.Structure):
    """bf_body_state: FP64 rigid-body state (world units)."""
    _fields_ = [
        ("position", ctypes.c_double * 3),
        ("orientation", ctypes.c_double * 4),
        ("linear_velocity", ctypes.c_double * 3),
        ("angular_velocity_body", ctypes.c_double * 3),
    ]


class BodyParams(ctypes.Structure):
    """bf_body_params."""
    _fields_ = [
        ("mass", ctypes.c_double),
        ("inertia", ctypes.c_double * 3),
        ("gravity", ctypes.c_double * 3),
        ("external_force", ctypes.c_double * 3),
        ("external_torque", ctypes.c_double * 3),
        ("safety_extent", ctypes.c_double * 3),
        ("safety_radius", ctypes.c_double),
        ("wing_speed", ctypes.c_double),
        ("mach_limit", ctypes.c_double),
        ("substeps", ctypes.c_uint32),
        ("free_flight", ctypes.c_uint32),
        ("move_mesh", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class BirdGeometry(ctypes.Structure):
    """bf_bird_geometry (metres, radians, Hz)."""
    _fields_ = [
        ("body_radii", ctypes.c_float * 3),
        ("wing_span", ctypes.c_float), ("wing_root_chord", ctypes.c_float), ("wing_tip_chord", ctypes.c_float),
        ("wing_thickness", ctypes.c_float), ("wing_sweep", ctypes.c_float),
        ("wing_root_offset", ctypes.c_float * 3),
        ("tail_length", ctypes.c_float), ("tail_half_width", ctypes.c_float), ("tail_thickness", ctypes.c_float),
        ("frequency", ctypes.c_float), ("stroke_amplitude", ctypes.c_float), ("stroke_bias", ctypes.c_float),
        ("pitch_mean", ctypes.c_float), ("pitch_amplitude", ctypes.c_float), ("pitch_phase", ctypes.c_float),
    ]


class WingDynamicsC(ctypes.Structure):
    """bf_wing_dynamics (kg, metres, kg m^2)."""
    _fields_ = [("mass", ctypes.c_float), ("center_of_mass", ctypes.c_float * 3), ("inertia", ctypes.c_float * 3)]


WING_FIELDS = ("root_x", "root_y", "root_z", "chord_mean", "radius", "thickness", "beta_exponent",
               "beta_normalization", "cycle", "stroke_amplitude", "stroke_duration", "pitch_duration", "pitch_low",
               "pitch_high", "stroke_rate_max", "reference_speed")


class FlappingWingC(ctypes.Structure):
    """bf_flapping_wing (grid world units, time units, radians)."""
    _fields_ = [(name, ctypes.c_float) for name in WING_FIELDS]


_lib = None


class LibraryMissing(RuntimeError):
    pass


def load():
    """Load the in-tree library.  There is no fallback: a missing library is an error."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY_PATH):
            raise LibraryMissing(
                f"{LIBRARY_PATH} is not built; run __graft_entry__.build() or make -C numi-gpu-physics_amd/csrc")
        lib = ctypes.CDLL(LIBRARY_PATH)
        _declare(lib)
        _lib = lib
    return _lib


def _declare(l):
    vp, u32, u64, f32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float
    st = ctypes.c_int
    sig = {
        "bf_create": (st, [ctypes.POINTER(Config), ctypes.POINTER(vp)]),
        "bf_destroy": (None, [vp]),
        "bf_last_error": (ctypes.c_char_p, [vp]),
        "bf_status_string": (ctypes.c_char_p, [st]),
        "bf_abi_version": (u32, []),
        "bf_device_info": (st, [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(u64)]),
        "bf_allocated_bytes": (u64, [vp]),
        "bf_init_equilibrium": (st, [vp]),
        "bf_init_shear_wave": (st, [vp, f32]),
        "bf_write_populations": (st, [vp, vp, u64]),
        "bf_read_populations": (st, [vp, vp, u64]),
        "bf_write_population_storage": (st, [vp, vp, u64]),
        "bf_read_population_storage": (st, [vp, vp, u64]),
        "bf_read_fields": (st, [vp, vp, vp]),
        "bf_step_index": (u64, [vp]),
        "bf_read_restart_state": (st, [vp, vp, vp]),
        "bf_set_restart_state": (st, [vp, u64, f32, vp, vp]),
        "bf_set_geometry": (st, [vp, vp, vp, vp]),
        "bf_read_geometry": (st, [vp, vp, vp]),
        "bf_mesh_upload": (st, [vp, vp, vp, u32, u32, vp, vp, u32, f32, vp, f32]),
        "bf_mesh_build": (st, [vp, f32]),
        "bf_mesh_upload_articulated": (st, [vp, vp, vp, vp, u32, vp, vp, u32, f32, vp, f32]),
        "bf_mesh_vertices_at": (st, [vp, f32, vp, vp, u32]),
        "bf_set_translating_sphere": (st, [vp, vp, vp, vp]),
        "bf_step": (st, [vp, u32]),
        "bf_synchronize": (st, [vp]),
        "bf_read_loads": (st, [vp, vp]),
        "bf_read_load_history": (st, [vp, vp, u32, ctypes.POINTER(u32)]),
        "bf_link_census": (st, [vp, vp]),
        "bf_set_profiling": (st, [vp, ctypes.c_int]),
        "bf_read_profile": (st, [vp, vp]),
        "bf_copy_bandwidth": (st, [u64, u32, ctypes.POINTER(ctypes.c_double)]),
        "bf_body_set": (st, [vp, ctypes.POINTER(BodyState), ctypes.POINTER(BodyParams)]),
        "bf_body_get": (st, [vp, ctypes.POINTER(BodyState)]),
        "bf_read_body_history": (st, [vp, vp, u32, ctypes.POINTER(u32)]),
        "bf_read_safety": (st, [vp, vp, ctypes.POINTER(u64), ctypes.POINTER(u32)]),
        "bf_adjoint_set_tangents": (st, [vp, u32, vp]),
        "bf_adjoint_set_kinematic_parameters": (st, [vp, u32, vp]),
        "bf_adjoint_gradient": (st, [vp, u32, vp, u32, ctypes.POINTER(ctypes.c_double), vp, vp]),
        "bf_adjoint_read_sensitivity": (st, [vp, vp, u64]),
        "bf_mesh_freeze_masks": (st, [vp, vp, u32]),
        "bf_set_bird": (st, [vp, ctypes.POINTER(BirdGeometry)]),
        "bf_read_bird_frame": (st, [vp, vp]),
        "bf_set_bird_wing_dynamics": (st, [vp, ctypes.POINTER(WingDynamicsC), ctypes.POINTER(WingDynamicsC)]),
        "bf_read_wing_reaction": (st, [vp, vp]),
        "bf_population_minimum": (st, [vp, vp, vp, vp]),
        "bf_inject_population": (st, [vp, ctypes.c_uint64, ctypes.c_float]),
        "bf_set_bird_keyframes": (st, [vp, vp, u32]),
        "bf_set_flapping_wing": (st, [vp, ctypes.POINTER(FlappingWingC)]),
        "bf_read_wing_frame": (st, [vp, vp]),
        "bf_read_wing_links": (st, [vp, vp, ctypes.c_uint64]),
        "bf_read_ray_links": (st, [vp, vp, ctypes.c_uint64]),
        "bf_ledger_fluid": (st, [vp, vp]),
        "bf_ledger_step_sources": (st, [vp, vp, vp]),
        "bf_ledger_step_sources_noalloc": (st, [vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(l, name)
        fn.restype = res
        fn.argtypes = args
