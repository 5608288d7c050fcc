/*
 * birdflow_hip.h — C ABI of libbirdflow_hip.so, the MI355X (gfx950) compute
 * backend for the D3Q19 moving-boundary step.
 *
 * It replaces the Metal compute backend of the reference:
 *   Sources/BirdFlowMetal/MetalBackend.swift:7-186        (device, pipelines, buffers, dispatch)
 *   Sources/BirdFlowMetal/BirdFlowSimulation.swift:1300-1727 (encode* step graph)
 *   Sources/BirdFlowMetal/GPUData.swift:4-80               (GPUUniforms lane packing)
 * and the indexed-mesh replay driver
 *   Sources/BirdFlowMetal/MetalIndexedBirdSurfaceValidation.swift:17697-18602.
 * Host code keeps the solver API of BirdFlowSimulation (init, advance,
 * snapshot, field copies, mesh upload, load readback, body state) on top of
 * these calls.
 *
 * Conventions
 *   - Plain C types only; the library owns every device allocation and copies
 *     caller arrays in and out.  No callbacks.
 *   - One context = one HIP stream on the current device.  Calls on one
 *     context must be serialized by the caller.
 *   - Populations are exchanged as FP32, direction-major SoA:
 *     f[q * N + x + nx * (y + ny * z)], N = nx * ny * nz, q in the reference's
 *     D3Q19 order (BirdFlow.metal:7-44).
 *   - Status codes map one-to-one onto BirdFlowError
 *     (Sources/BirdFlowMetal/BirdFlowError.swift:3-19).  After a failed step
 *     the context is invalidated: every later call except bf_destroy and
 *     bf_last_error returns BF_ERROR_STATE_INVALIDATED
 *     (BirdFlowSimulation.swift:2059-2087).
 *   - There is no CPU fallback: bf_create fails with
 *     BF_ERROR_DEVICE_UNAVAILABLE when no gfx950 device is present.
 */
#ifndef BIRDFLOW_HIP_H
#define BIRDFLOW_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BF_ABI_VERSION 1u

typedef enum bf_status {
    BF_OK = 0,
    BF_ERROR_DEVICE_UNAVAILABLE = 1,          /* metalUnavailable / deviceUnavailable */
    BF_ERROR_INVALID_CONFIGURATION = 2,       /* BirdFlowConfigurationError (grid, tau, Mach) */
    BF_ERROR_ALLOCATION_FAILED = 3,           /* allocationFailed(bytes:) */
    BF_ERROR_BUFFER_EXCEEDS_DEVICE_LIMIT = 4, /* bufferExceedsDeviceLimit */
    BF_ERROR_WORKING_SET_EXCEEDS_MEMORY = 5,  /* workingSetExceedsRecommendation */
    BF_ERROR_INVALID_ADVANCE_REQUEST = 6,     /* invalidAdvanceRequest(steps:batchSize:) */
    BF_ERROR_INVALID_ARGUMENT = 7,            /* resourceMissing / malformed caller arrays */
    BF_ERROR_LAUNCH_FAILED = 8,               /* commandBufferFailed */
    BF_ERROR_RUNTIME_SAFETY_VIOLATION = 9,    /* runtimeSafetyViolation */
    BF_ERROR_STATE_INVALIDATED = 10,          /* simulationStateInvalidated */
    BF_ERROR_NOT_SUPPORTED = 11
} bf_status;

/* Face treatment (per axis, low and high side).  The reference offers
 * all-periodic wrapping or all-faces far-field equilibrium
 * (BirdFlow.metal:2696-2725); OUTLET (zero-gradient: the pull source is
 * clamped into the domain) is an addition. */
enum { BF_BC_PERIODIC = 0, BF_BC_FARFIELD = 1, BF_BC_OUTLET = 2 };

/* Collision operator, replacing the caseParameters.w thresholds
 * (BirdFlow.metal:2906-2912). */
enum { BF_COLLISION_TRT = 0, BF_COLLISION_TRT_SYMMETRIC_LIMITER = 1,
       BF_COLLISION_REGULARIZED = 2, BF_COLLISION_RR3 = 3 };

/* Wall rule: halfway bounce-back, or Bouzidi interpolated bounce-back with the
 * link fraction from the signed-distance ratio (BirdFlow.metal:2044-2076,
 * 2729-2806), or the interpolated path at q = 1/2 (translating-sphere
 * canonical, BirdFlow.metal:2340-2345), or a link table (the prescribed
 * wing's bisected fractions, BirdFlow.metal:2738-2745).
 * BF_WALL_INTERPOLATED_RAY (meshes): the exact fraction of each link, the
 * first entry of the link from the fluid node toward its solid source into
 * the meshed shell {dist(p, triangle) <= half thickness}, in closed form
 * (offset faces, edge cylinders, vertex spheres); the reference computes
 * these roots only in its CPU audits (MetalIndexedBirdSurfaceValidation.swift:
 * 14245-14444).  A link with no entry within it keeps the signed-distance
 * ratio.  Masks, wall velocities and link counts are those of the SDF rule. */
enum { BF_WALL_HALFWAY = 0, BF_WALL_INTERPOLATED_SDF = 1, BF_WALL_INTERPOLATED_HALF = 2,
       BF_WALL_INTERPOLATED_LINKS = 3, BF_WALL_INTERPOLATED_RAY = 4 };

/* Population storage: FP32, or FP16 / BF16 offsets f - w_q with FP32
 * arithmetic. */
enum { BF_STORAGE_FP32 = 0, BF_STORAGE_FP16_OFFSET = 1, BF_STORAGE_BF16_OFFSET = 2 };

/* Collision arithmetic.  STRICT evaluates the reference's expressions in
 * the reference's order without contraction (FP32 storage only): the step is
 * reproducible operation for operation on any IEEE FP32 machine.  FUSED
 * evaluates the same operators over opposite-direction pairs with fused
 * multiply-adds, about a third of the instructions; the two differ by
 * rounding.  DEFAULT (0, a zero-initialised config) is STRICT for FP32
 * storage, where the step is bandwidth-bound either way, and FUSED for
 * FP16 / BF16 storage, where the arithmetic would otherwise bound it. */
enum { BF_ARITHMETIC_DEFAULT = 0, BF_ARITHMETIC_STRICT = 1, BF_ARITHMETIC_FUSED = 2 };

/* Where each step's geometry comes from. */
enum { BF_GEOMETRY_STATIC = 0, BF_GEOMETRY_MESH = 1, BF_GEOMETRY_SPHERE = 2, BF_GEOMETRY_BIRD = 3,
       BF_GEOMETRY_WING = 4 };

/* Solver configuration in lattice units (the host converts physical units
 * with the reference's LatticeScaling, SimulationConfiguration.swift:137-183).
 * Replaces GPUUniforms (GPUData.swift:4-80). */
typedef struct bf_config {
    uint32_t nx, ny, nz;          /* each >= 1; nx*ny*nz < 2^32 */
    uint32_t storage;             /* BF_STORAGE_* */
    uint32_t collision;           /* BF_COLLISION_* */
    uint32_t wall_rule;           /* BF_WALL_* */
    uint32_t force_mode;          /* 0..6 momentum-exchange estimator; 6 = conservative moving-domain */
    uint32_t load_component;      /* 0 all, 1 link exchange only, 2 topology impulse only */
    uint32_t selected_part;       /* 0 = all parts */
    uint32_t accumulate_loads;    /* gravity.w */
    uint32_t boundary[3][2];      /* BF_BC_* per axis, [low, high] */
    float origin[3];              /* world position of cell (0,0,0)'s corner */
    float cell_size;              /* dx (world units) */
    float time_step;              /* dt (world units) used for geometry time */
    float velocity_to_lattice;    /* dt / dx */
    float force_to_physical;      /* rho dx^4 / dt^2 */
    float omega_plus, omega_minus;
    float sponge_strength;        /* 0..1 */
    float sponge_width;           /* cells */
    float far_velocity[3];        /* lattice units */
    float far_density;
    float body_position[3];       /* torque reference point (world units) */
    uint32_t arithmetic;          /* BF_ARITHMETIC_*; 0 = default for the storage */
    uint32_t reserved[7];
} bf_config;

typedef struct bf_context bf_context;

/* MetalBackend.init + BirdFlowSimulation.init buffer plan
 * (MetalBackend.swift:14-46, 77-106; BirdFlowSimulation.swift:480-625).
 * The populations start at the far-field equilibrium of the empty geometry;
 * bf_init_equilibrium re-initializes them over the geometry set since. */
bf_status bf_create(const bf_config* config, bf_context** out);
void bf_destroy(bf_context* ctx);
const char* bf_last_error(const bf_context* ctx);
const char* bf_status_string(bf_status status);
uint32_t bf_abi_version(void);
/* Device name and memory of the current device (MetalBackend.device). */
bf_status bf_device_info(char* name, size_t name_capacity, uint64_t* total_bytes);
/* Bytes of device memory the context holds. */
uint64_t bf_allocated_bytes(const bf_context* ctx);

/* ---- state ---- */
/* initializePopulations (BirdFlow.metal:2078-2103): solid cells at their wall
 * velocity, fluid at the far-field velocity, rho = 1; resets the step index. */
bf_status bf_init_equilibrium(bf_context* ctx);
/* initializeShearWave (BirdFlow.metal:2132-2162). */
bf_status bf_init_shear_wave(bf_context* ctx, float amplitude);
/* Blit copies of the population lattice (MetalShearWaveSimulation.copyPopulations,
 * BirdFlowSimulation.swift:1214-1298 checkpoint read-back).  count = 19 * N. */
bf_status bf_write_populations(bf_context* ctx, const float* host, uint64_t count);
bf_status bf_read_populations(bf_context* ctx, float* host, uint64_t count);
/* The stored form itself, unchanged (the raw buffer copies of the
 * checkpoint, BirdFlowSimulation.swift:1214-1298 and 1920-1978): 19 * N
 * elements in the direction-major order above, FP32 populations or the FP16 /
 * BF16 bits of the offsets f_q - w_q; bytes = 19 * N * (4 or 2).  A
 * write / read round trip through this form is bit-exact for every storage
 * (through FP32 values it is not for FP16 / BF16). */
bf_status bf_write_population_storage(bf_context* ctx, const void* host, uint64_t bytes);
bf_status bf_read_population_storage(bf_context* ctx, void* host, uint64_t bytes);
/* extractMacroscopicFields (BirdFlow.metal:2107-2127) into host arrays:
 * density[N], velocity[3N] interleaved. */
bf_status bf_read_fields(bf_context* ctx, float* density, float* velocity3);
uint64_t bf_step_index(const bf_context* ctx);
/* Restart state beside the populations (no single counterpart in the
 * reference, whose restore sets the step, clock, last load and solid mask
 * inside restoreCheckpointState, BirdFlowSimulation.swift:1920-1990, behind
 * init(checkpointURL:), BirdFlowSimulation.swift:661-680).
 * bf_read_restart_state: the completed step index and the mesh clock origin
 * (mesh time of step s = origin + s * time_step).
 * bf_set_restart_state: renumbers the completed step, sets the clock origin,
 * optionally the last reduced load (8 doubles as bf_read_loads reports) and
 * the solid mask the next step starts from (N bytes, x fastest), and rebuilds
 * the current step's mesh or sphere geometry at the new time.  Call it after
 * the geometry source and body are set and before bf_write_populations.
 * Load and body histories restart empty.  A mask that differs from a static
 * geometry is BF_ERROR_INVALID_ARGUMENT. */
bf_status bf_read_restart_state(bf_context* ctx, uint64_t* step, float* mesh_time_origin);
bf_status bf_set_restart_state(bf_context* ctx, uint64_t step, float mesh_time_origin, const double last_load[8],
                               const uint8_t* solid_mask);

/* ---- geometry ---- */
/* Static geometry written by the caller: part masks (0 fluid, 1..255 part id)
 * of the previous and current step and wall[N] = (ux, uy, uz, signed
 * distance / dx).  mask_prev may be NULL (= mask_cur).  Selects
 * BF_GEOMETRY_STATIC. */
bf_status bf_set_geometry(bf_context* ctx, const uint8_t* mask_prev, const uint8_t* mask_cur, const float* wall4);
/* Current geometry read-back (integer flags for bit-exact parity). */
bf_status bf_read_geometry(bf_context* ctx, uint8_t* mask, float* wall4);

/* Indexed mesh upload (MeasuredBirdSurfaceSequence.swift:120-204,
 * MetalIndexedBirdSurfaceValidation.swift:17746-17915): frame-major points
 * [frame_count][vertex_count][3] (world units), strictly increasing
 * frame_times, triangles [triangle_count][3] vertex indices, triangle part ids
 * (1..255), shell half thickness (world units), rigid translation added to
 * every vertex.  No triangle limit (keys are 64-bit).  Builds the geometry at
 * time_start into the current step and selects BF_GEOMETRY_MESH; each later
 * step builds the geometry at time_start + step * time_step. */
bf_status bf_mesh_upload(bf_context* ctx, const float* points, const float* frame_times, uint32_t frame_count,
                         uint32_t vertex_count, const uint32_t* triangles, const uint8_t* triangle_parts,
                         uint32_t triangle_count, float half_thickness, const float translation[3],
                         float time_start);
/* Rebuild the current step's geometry at an arbitrary time (geometry audit,
 * MetalIndexedBirdSurfaceReplay.snapshot, :17916-17992). */
bf_status bf_mesh_build(bf_context* ctx, float time);

/* On-device kinematics (bf_set_kinematics of the boundary plan): an
 * articulated mesh whose vertices ride on rigid joints driven by Fourier
 * series, evaluated on the device each step instead of uploading one frame
 * per step.  The reference drives its bird's wings the same way, by stroke,
 * deviation, pitch and tip twist angles about a hinge (makeWingFrame,
 * prepareBirdGeometry, BirdFlow.metal:611-814); its indexed meshes carry
 * only recorded frames.  A vertex of joint 0 is rigid: position = center +
 * rest.  A vertex of joint j >= 1 (side s = joints[j-1].side, +1 or -1):
 *   l   = (rest.x, s rest.y, rest.z)
 *   a   = pitch + twist * span_fraction  (radians)
 *   u3  = Ry(a) l,  Ry(a) (x, y, z) = (cos a x + sin a z, y, -sin a x + cos a z)
 *   u2  = Rz(s deviation) u3,  u1 = Rx(s stroke) u2  (right-handed rotations)
 *   pos = center + (root.x, s root.y, root.z) + u1
 * with each angle = amplitude_scale * sum_h (c[h][0] sin(h phi) + c[h][1]
 * cos(h phi)), phi = 2 pi (t mod period) / period, h = 1..harmonics; the
 * velocity is the time derivative of pos (rates from the same series) times
 * velocity_to_lattice.  FP32 with this library's own sine and cosine.
 * Triangles whose three vertices are rigid keep their raster keys from step
 * to step.  Builds the geometry at time_start like bf_mesh_upload and selects
 * BF_GEOMETRY_MESH (the adjoint, which differentiates frame tangents, refuses
 * an articulated mesh). */
enum { BF_MAX_JOINTS = 8, BF_MAX_HARMONICS = 4 };
typedef struct bf_joint {
    float side;                          /* +1 or -1 */
    float root[3];                       /* hinge of the +y side (world units) */
    float stroke[BF_MAX_HARMONICS][2];   /* [harmonic 1..4][sin, cos] radians */
    float deviation[BF_MAX_HARMONICS][2];
    float pitch[BF_MAX_HARMONICS][2];
    float twist[BF_MAX_HARMONICS][2];
} bf_joint;
typedef struct bf_kinematics {
    float period;                        /* time units, > 0 */
    float amplitude_scale;               /* multiplies every angle */
    uint32_t harmonics;                  /* 1..BF_MAX_HARMONICS */
    uint32_t joint_count;                /* 0..BF_MAX_JOINTS */
    float center[3];                     /* world position of the rest frame's origin */
    float reserved;
    bf_joint joints[BF_MAX_JOINTS];
} bf_kinematics;
/* rest: [vertex_count][3] (world units, relative to center); vertex_joint:
 * [vertex_count] in 0..joint_count; span_fraction: [vertex_count] (the
 * twist's weight); triangles and parts as in bf_mesh_upload. */
bf_status bf_mesh_upload_articulated(bf_context* ctx, const float* rest, const uint8_t* vertex_joint,
                                     const float* span_fraction, uint32_t vertex_count, const uint32_t* triangles,
                                     const uint8_t* triangle_parts, uint32_t triangle_count, float half_thickness,
                                     const bf_kinematics* kinematics, float time_start);
/* Positions and lattice velocities [vertex_count][3] of the uploaded mesh
 * (either form) prepared at `time` exactly as a step at that time prepares
 * them (the vertex-prepare pass, BirdFlow.metal:1589-1622). */
bf_status bf_mesh_vertices_at(bf_context* ctx, float time, float* position3, float* velocity3,
                              uint32_t vertex_count);

/* Translating-sphere canonical (BirdFlow.metal:2263-2350): center(t) =
 * center_radius.xyz + geometry_velocity * t, wall_velocity.w selects uniform
 * (0), tangential (1) or normal (2) wall velocity.  Builds the geometry at
 * t = step_index * time_step and selects BF_GEOMETRY_SPHERE. */
bf_status bf_set_translating_sphere(bf_context* ctx, const float center_radius[4], const float geometry_velocity[3],
                                    const float wall_velocity[4]);

/* Analytic signed-distance bird: BirdParameters geometry and analytic wing
 * kinematics as GPUBirdParameters packs them (GPUData.swift:82-140, from
 * BirdModel.swift:129-217): metres, radians, Hz. */
typedef struct bf_bird_geometry {
    float body_radii[3];
    float wing_span, wing_root_chord, wing_tip_chord, wing_thickness, wing_sweep;
    float wing_root_offset[3];
    float tail_length, tail_half_width, tail_thickness;
    float frequency, stroke_amplitude, stroke_bias, pitch_mean, pitch_amplitude, pitch_phase;
} bf_bird_geometry;

/* The reference's product geometry (prepareBirdGeometry + buildBirdGeometry,
 * BirdFlow.metal:686-938; encodeGeometryPreparation / encodeGeometry,
 * BirdFlowSimulation.swift:1451-1510): each step's frame from the device body
 * pose at t = step * time_step, then part masks (1 body, 2 left wing, 3 right
 * wing, 4 tail) and wall = (velocity * velocity_to_lattice, 0) in world
 * units (origin, cell_size).  Builds the current step (no previous geometry,
 * as the reference's initialization) and selects BF_GEOMETRY_BIRD; in free
 * flight every step's geometry follows the integrated pose. */
bf_status bf_set_bird(bf_context* ctx, const bf_bird_geometry* bird);
/* Measured wing kinematics (BirdParameters.measuredWingKinematics,
 * MeasuredBirdDataset.swift:105-160; sampled by measuredWingKinematics,
 * BirdFlow.metal:550-609, cubic Hermite over the cycle phase time *
 * frequency): `count` keyframes of 17 floats (phase, left stroke / deviation /
 * pitch / tip twist angles, their rates per second, right angles, right
 * rates); 4...4096 keyframes, phases strictly increasing in [0, 1) from 0.
 * count 0 returns to the analytic sinusoid.  With the bird set, the current
 * step's geometry is rebuilt. */
bf_status bf_set_bird_keyframes(bf_context* ctx, const float* keyframes, uint32_t count);
/* Prescribed-wing dynamics of the analytic bird (BirdParameters.
 * prescribedWingDynamics, BirdModel.swift:98-128, as GPUBirdParameters packs
 * it, GPUData.swift:150-176): per wing, mass (kg), centre of mass from the
 * hinge along the wing's chord, span and normal axes (m), principal inertia
 * about those axes through the centre of mass (kg m^2). */
typedef struct bf_wing_dynamics {
    float mass;
    float center_of_mass[3];
    float inertia[3];
} bf_wing_dynamics;
/* Sets (both non-NULL) or clears (both NULL) the wings' dynamics: each step,
 * after the frame is prepared, updateWingInertialReaction
 * (BirdFlow.metal:6493-6572) forms each wing's hinge reaction -(momentum -
 * previous momentum) / time_step, and in free flight integrateBirdBody
 * (:6574-6627) adds both to the fluid force and torque.  With the bird set,
 * the current step's frame is rebuilt and its momenta start the history
 * (initializeOnly: no reaction on it). */
bf_status bf_set_bird_wing_dynamics(bf_context* ctx, const bf_wing_dynamics* left, const bf_wing_dynamics* right);
/* The current step's hinge reactions (GPUWingInertialReaction, 4 float4: left
 * force, left torque, right force, right torque; world units, w = 0). */
bf_status bf_read_wing_reaction(bf_context* ctx, float reaction[16]);
/* Prepared frame of the current step (GPUPreparedBirdGeometry, 14 float4:
 * body position, orientation x y z w, linear velocity, world angular
 * velocity + geometry radius, then left and right wing root, chord (w: tip
 * twist), span, normal, angular velocity (w: tip twist rate)). */
bf_status bf_read_bird_frame(bf_context* ctx, float frame[56]);

/* Prescribed beta-planform flapping wing (GPUFlappingWingParameters,
 * GPUData.swift:383-388; BirdFlow.metal:940-1297): root (grid world units),
 * mean chord, span radius, thickness, beta exponent and normalization, cycle
 * (time units of the grid's time_step), stroke amplitude (rad), stroke and
 * pitch transition durations (fractions of a cycle), pitch at mid-stroke
 * (downstroke, upstroke; rad), maximum stroke rate (rad per cycle) and the
 * reference speed, in that order. */
typedef struct bf_flapping_wing {
    float root[3], chord_mean;
    float radius, thickness, beta_exponent, beta_normalization;
    float cycle, stroke_amplitude, stroke_duration, pitch_duration;
    float pitch_low, pitch_high, stroke_rate_max, reference_speed;
} bf_flapping_wing;

/* preparePrescribedFlappingWing + buildPrescribedFlappingWing +
 * the link-fraction pass (BirdFlow.metal:1040-1297, encoded by
 * MetalFlappingWingValidation.swift): each step's frame at t = step *
 * time_step, mask 1 on the wing, wall = (omega x r * velocity_to_lattice,
 * -1 on the wing, +1 near it), and the bisected fluid-side link fraction of
 * every fluid cell whose source is on the wing, read by the wall rule
 * BF_WALL_INTERPOLATED_LINKS.  Builds the current step and selects
 * BF_GEOMETRY_WING. */
bf_status bf_set_flapping_wing(bf_context* ctx, const bf_flapping_wing* wing);
/* Prepared frame of the current step (GPUPreparedFlappingWing, 6 float4:
 * root (w: mean chord), chord, span, normal, angular velocity, state =
 * stroke, pitch and their rates per time unit). */
bf_status bf_read_wing_frame(bf_context* ctx, float frame[24]);
/* The current step's link table as links[q * N + cell] (count = 19 N): the
 * fraction for a fluid cell whose source cell - c_q is on the wing, 0
 * elsewhere. */
bf_status bf_read_wing_links(bf_context* ctx, float* links, uint64_t count);
/* The current step's exact-root table (BF_WALL_INTERPOLATED_RAY with a mesh)
 * as links[q * N + cell] (count = 19 N): the clamped fraction of every fluid
 * cell's link whose source is solid, 0 elsewhere.  Replaces the per-link
 * roots of the reference's link-ray-root audit
 * (MetalIndexedBirdSurfaceValidation.swift:14245-14444). */
bf_status bf_read_ray_links(bf_context* ctx, float* links, uint64_t count);

/* ---- momentum ledger (FP64 on the device) ----
 * The device sums of BirdFlowSimulation.advanceWithCoupledMomentumLedger
 * (BirdFlowSimulation.swift:805-1085; kernels BirdFlow.metal:4272-4526),
 * formed per cell in FP64 and summed in a fixed order.  Lattice units.
 * bf_ledger_fluid: mass and momentum (x, y, z) of the current populations
 * over the fluid cells of the current mask (measureFluidMassMomentum).
 * bf_ledger_step_sources: of the step just completed (call it after bf_step
 * and before any other population write): sums[0..3] fluid change (after
 * over the current mask minus before over the previous mask), [4..7]
 * open-boundary sources (far-field and outlet links), [8..11] sponge
 * sources, [12..15] persistent-wall-link exchange (halfway form, wall density
 * far_rho); counts: open-boundary links, sponge cells, persistent links,
 * cells whose occupancy changed.  Otherwise BF_ERROR_INVALID_ARGUMENT. */
bf_status bf_ledger_fluid(bf_context* ctx, double out[4]);
bf_status bf_ledger_step_sources(bf_context* ctx, double sums[16], uint64_t counts[4]);
/* bf_ledger_step_sources without allocating.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * same kernels and block count, so the same sums bit for bit; the block
 * partials (at most 4,096 blocks of 16 + 4 values) and the sums live in a
 * file-scope device array of 655,616 bytes that each GPU holds until the process
 * ends, instead of buffers allocated on the first ledger call.  Every grid
 * fits (the kernels stride over cells); a block count above the scratch is
 * BF_ERROR_INVALID_ARGUMENT.  All contexts on one GPU share the array: calls
 * that overlap on one GPU race, and a call overwrites the previous call's
 * contents. */
bf_status bf_ledger_step_sources_noalloc(bf_context* ctx, double sums[16], uint64_t counts[4]);

/* ---- stepping ---- */
/* advance(steps:) (BirdFlowSimulation.swift:687-799): per step, geometry at
 * t_{n+1} into the next mask, fluid step (current -> next lattice, previous /
 * current masks), fixed-order FP64 load reduction into the history, swap.
 * Asynchronous: returns after enqueueing; errors of earlier work surface here
 * or at bf_synchronize. */
bf_status bf_step(bf_context* ctx, uint32_t steps);
bf_status bf_synchronize(bf_context* ctx);
/* Loads of the last step: {Fx, Fy, Fz, Tx, Ty, Tz, limiter activations,
 * max limiter restriction} (GPUForceTorque, BirdFlow.metal:197-200), FP64. */
bf_status bf_read_loads(bf_context* ctx, double out[8]);
/* Per-step load history (storeForceTorqueSample, BirdFlow.metal:5392-5401):
 * up to max_records most recent records of 8 doubles, oldest first. */
bf_status bf_read_load_history(bf_context* ctx, double* out, uint32_t max_records, uint32_t* records);
/* Link census of the current step: counts[q] = fluid cells whose pull source
 * along q is solid, counts[19] newly covered, counts[20] newly uncovered. */
bf_status bf_link_census(bf_context* ctx, uint64_t counts[21]);

/* ---- measurement ---- */
/* enabled = 1: every step records HIP events around the fast-fluid launch on
 * its stream; 2: also around the general-fluid launch and the next step's
 * geometry (each event pair adds latency to the step); 0: off. */
bf_status bf_set_profiling(bf_context* ctx, int enabled);
/* {geometry ms, fast fluid ms, general fluid ms (0 unless level 2), steps
 * profiled, tiles on the general path in the last step, total tiles (64 x 4
 * cells)} summed over the steps profiled since the last reset; resets the
 * counters. */
bf_status bf_read_profile(bf_context* ctx, double out[6]);

/* Device copy bandwidth on the current device: a plain 16-byte-per-lane copy
 * between two fresh buffers of `bytes` each, repeated `iterations` times;
 * returns (read + write bytes) / time in GB/s.  Calibrates achieved-bandwidth
 * figures and hardware-counter byte counts. */
bf_status bf_copy_bandwidth(uint64_t bytes, uint32_t iterations, double* gbs);

/* ---- Rigid body (six degrees of freedom, FP64 on the device) ----
 * Replaces integrateBirdBody / monitorBirdRuntimeSafety (BirdFlow.metal:
 * 6574-6707), RigidBodyIntegrator.integrate (RigidBody.swift:85-126) and the
 * host side of freeFlight stepping (BirdFlowSimulation.swift:1412-1424).
 * Units are world units (those of origin, cell_size and time_step).  The
 * fluid kernels take the torque reference point from the device body
 * position; bf_create places the body at config.body_position, at rest. */
typedef struct bf_body_state {
    double position[3];
    double orientation[4];            /* body to world, quaternion x, y, z, w */
    double linear_velocity[3];
    double angular_velocity_body[3];  /* body frame */
} bf_body_state;

typedef struct bf_body_params {
    double mass;                      /* bodyRadiiAndMass.w */
    double inertia[3];                /* principal moments, body frame */
    double gravity[3];                /* world acceleration (uniforms.gravity.xyz) */
    double external_force[3];         /* added to the fluid load every step (world): */
    double external_torque[3];        /*   e.g. wing hinge reactions (BirdFlow.metal:6586-6591) */
    double safety_extent[3];          /* body-frame half extents (safetyGeometry.xyz) */
    double safety_radius;             /* rotation radius (safetyGeometry.w) */
    double wing_speed;                /* added surface speed (safetyLimits.x) */
    double mach_limit;                /* safetyLimits.y; <= 0 turns the monitor off */
    uint32_t substeps;                /* integration.x, >= 1 */
    uint32_t free_flight;             /* flags.x: integrate after every step (needs accumulate_loads) */
    uint32_t move_mesh;               /* mesh frames are body-frame; the body pose places them each step */
    uint32_t reserved;
} bf_body_params;

/* Set the body state and/or parameters (either may be NULL).  Resets the
 * safety record and the body history.  With move_mesh, call it before
 * bf_mesh_upload / bf_mesh_build so that the first geometry uses the pose. */
bf_status bf_body_set(bf_context* ctx, const bf_body_state* state, const bf_body_params* params);
bf_status bf_body_get(bf_context* ctx, bf_body_state* state);
/* Body state after each of the last min(max_records, steps) free-flight
 * steps, oldest first. */
bf_status bf_read_body_history(bf_context* ctx, bf_body_state* out, uint32_t max_records, uint32_t* records);
/* Safety record: metrics = max Mach, min clearance, Mach and clearance at
 * the first event; flags = 1 Mach above the limit, 2 clearance below zero,
 * 4 non-finite state (GPURuntimeSafetyRecord, GPUData.swift:234-245), and 8
 * non-finite populations: every fluid kernel checks that each cell's stored
 * populations sum to a finite density and records the first step that fails
 * (event_step when no body event came first; reset by a new population
 * state).  bf_synchronize returns BF_ERROR_RUNTIME_SAFETY_VIOLATION once when
 * a new event is recorded; the context stays usable. */
bf_status bf_read_safety(bf_context* ctx, double metrics[4], uint64_t* event_step, uint32_t* flags);
/* reducePopulationMinimum (BirdFlow.metal:6404-6459, read back by
 * MetalIndexedBirdSurfaceValidation.swift:20067-20087): the smallest current
 * population (FP16 / BF16 offsets decoded), a non-finite value comparing as
 * -infinity, ties to the smallest direction-major index q N + cell; the
 * value, that index and whether it is non-finite. */
bf_status bf_population_minimum(bf_context* ctx, float* value, uint64_t* index, uint32_t* non_finite);
/* Fault injection for the safety checks: sets population `index`
 * (direction-major, q N + cell) to `value` in the stored form. */
bf_status bf_inject_population(bf_context* ctx, uint64_t index, float value);

/* ---- Adjoint (reverse-mode gradient of a load functional) ----
 * J = sum over the next `steps` steps s of sum_i w[s][i] L_s[i], with L_s the
 * step's force (i = 0..2) and torque (3..5) as bf_read_loads reports them.
 * dJ/dtheta for up to 8 parameters theta_k that move the mesh: the caller
 * supplies d(frame vertex positions)/d theta_k; wall velocities and signed
 * distances then follow from the same frame interpolation and closest-point
 * geometry.  Each step's link set (masks) is held fixed, as it is
 * discontinuous in theta; link fractions and wall velocities are
 * differentiated.  No counterpart exists in the reference.  Supported: FP32
 * and FP16 / BF16 offset storage (populations decoded on load, the storage
 * rounding taken as the identity; dJ/df stays FP32), every collision
 * (the limiters differentiated through the direction that sets their
 * scale), periodic and far-field faces and outlets on the x faces, force
 * mode 6, mesh (frames or articulated) or static geometry with the
 * signed-distance, half-way, interpolated-half or exact-root wall rule (each
 * exact root differentiated through the triangle whose shell gives it), a
 * fixed body pose (move_mesh placements included), no free flight.
 * Populations are snapshotted every checkpoint_interval steps in the forward
 * pass and recomputed segment by segment in the reverse sweep (memory:
 * about (segments + interval) x 19 x the storage's bytes and interval x 21 B
 * per cell). */
/* point_tangents: [parameter_count][frame_count][vertex_count][3]; 0 clears.
 * A new bf_mesh_upload clears the tangents. */
bf_status bf_adjoint_set_tangents(bf_context* ctx, uint32_t parameter_count, const float* point_tangents);
/* The parameters of an articulated mesh (bf_mesh_upload_articulated) are
 * its Fourier coefficients: parameter k is an offset added to coefficient
 * [harmonic][component] (component 0 multiplies sin((h+1) phi), 1
 * cos((h+1) phi)) of series `channel` (0 stroke, 1 deviation, 2 pitch,
 * 3 twist) of every joint whose bit is set in joint_mask (bit j: joints[j];
 * one parameter can move both wings: its derivative is the sum of the
 * joints' derivatives).  The device differentiates the kinematics of each
 * step's time
 * (forward-mode dual numbers), no tangents are uploaded.  Replaces frame
 * tangents; 0 clears.  A new mesh upload clears the parameters. */
typedef struct bf_kinematic_parameter {
    uint32_t joint_mask;
    uint32_t channel;
    uint32_t harmonic;
    uint32_t component;
} bf_kinematic_parameter;
bf_status bf_adjoint_set_kinematic_parameters(bf_context* ctx, uint32_t parameter_count,
                                              const bf_kinematic_parameter* parameters);
/* Runs `steps` forward steps from the current state (the context advances
 * as with bf_step), then the reverse sweep.  load_weights: [steps][6].
 * objective <- J; gradient[parameter_count] <- dJ/dtheta (FP64 sums);
 * timings_ms (may be NULL) <- {forward pass, reverse sweep}. */
bf_status bf_adjoint_gradient(bf_context* ctx, uint32_t steps, const double* load_weights,
                              uint32_t checkpoint_interval, double* objective, double* gradient,
                              double timings_ms[2]);
/* dJ/df at the state the last sweep started from (19 x cells, FP32). */
bf_status bf_adjoint_read_sensitivity(bf_context* ctx, float* out, uint64_t count);
/* Replays recorded part masks: masks[k] (cells bytes) replaces the part mask
 * of step (current step + k) for k < count, the current step's at once; wall
 * fields still follow the mesh.  Finite differences of a load functional with
 * the link set held fixed (the derivative the adjoint computes) re-step with
 * the masks of the unperturbed run.  count = 0 ends the replay. */
bf_status bf_mesh_freeze_masks(bf_context* ctx, const uint8_t* masks, uint32_t count);

#ifdef __cplusplus
}
#endif
#endif
