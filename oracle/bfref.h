/*
 * bfref — CPU oracle for the D3Q19 moving-boundary step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in numi-gpu-physics_amd/csrc.  Only tests/, the smoke entry
 * point and the CPU-baseline leg of bench.py may load it; the product path
 * never calls it and has no CPU fallback.
 *
 * It restates, in plain C, the arithmetic of the reference solver:
 *   - D3Q19 tables and second-order equilibrium
 *       Sources/BirdFlowCore/D3Q19.swift:3-58, Sources/BirdFlowMetal/Metal/BirdFlow.metal:4-44,478-484
 *   - the fused pull-stream / boundary / collide / load step
 *       Sources/BirdFlowMetal/Metal/BirdFlow.metal:2507-3236 (stepFluidTRT)
 *   - sponge factor  BirdFlow.metal:2427-2446
 *   - interpolated bounce-back and cover/uncover impulses
 *       Sources/BirdFlowMetal/InterpolatedBounceBackReference.swift:12-61
 *   - link fraction from signed distances  BirdFlow.metal:2044-2076
 *   - indexed-mesh vertex preparation, raster, resolve
 *       BirdFlow.metal:1529-1583 (closest point), 1589-1622 (prepare),
 *       1638-1686 (raster), 1697-1738 (resolve), 1898-1911 (clear);
 *       CPU twin Sources/BirdFlowMetal/MetalIndexedBirdSurfaceValidation.swift:18005-18110,20302-20359
 *   - frame interval search  Sources/BirdFlowMetal/MeasuredBirdSurfaceSequence.swift:174-204
 *   - shear-wave initial state  BirdFlow.metal:2132-2162 and
 *       Sources/BirdFlowMetal/MetalShearWaveValidation.swift:666-755
 *   - translating-sphere topology canonical  BirdFlow.metal:183-195,2263-2350
 *   - six-degree-of-freedom body update  BirdFlow.metal:6574-6627,
 *   - runtime safety monitor             BirdFlow.metal:6629-6707,
 *   - analytic signed-distance bird      BirdFlow.metal:486-533,550-938
 *   - prescribed beta-planform wing      BirdFlow.metal:940-1297
 *       Sources/BirdFlowCore/RigidBody.swift:85-127
 *
 * Fluid functions exist in two precisions: *_f32 (same operation order as the
 * HIP kernels, built without floating-point contraction) and *_f64 (the FP64
 * oracle used to state the error of the FP32 and reduced-storage forms).
 * Geometry is FP32 only: its integer flags are defined by FP32 arithmetic.
 *
 * Population layout: SoA, direction-major, f[q * N + x + nx * (y + ny * z)].
 */
#ifndef BFREF_H
#define BFREF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary treatment for a domain face (per axis, low and high side). */
enum {
    BFO_BC_PERIODIC = 0,   /* wrap to the opposite face */
    BFO_BC_FARFIELD = 1,   /* equilibrium at (far_rho, far_u) */
    BFO_BC_OUTLET = 2      /* zero-gradient: source clamped into the domain */
};

enum {
    BFO_COLLISION_TRT = 0,
    BFO_COLLISION_TRT_SYMMETRIC_LIMITER = 1,
    BFO_COLLISION_REGULARIZED = 2,
    BFO_COLLISION_RR3 = 3
};

enum {
    BFO_WALL_HALFWAY = 0,            /* halfway bounce-back, wall velocity of the solid source */
    BFO_WALL_INTERPOLATED_SDF = 1,   /* Bouzidi with q from the signed-distance ratio */
    BFO_WALL_INTERPOLATED_HALF = 2,  /* Bouzidi path with q fixed at 1/2 */
    BFO_WALL_INTERPOLATED_LINKS = 3  /* Bouzidi with q from a link table (bfo_set_link_table) */
};

typedef struct bfo_params {
    int32_t nx, ny, nz;
    int32_t bc[3][2];              /* [axis][0=low,1=high] */
    int32_t collision;
    int32_t wall_rule;
    int32_t force_mode;            /* 0..6 momentum-exchange estimator */
    int32_t load_component;        /* 0 all, 1 link exchange only, 2 topology impulse only */
    int32_t selected_part;         /* 0 = every part */
    int32_t accumulate_loads;
    int32_t capture_fields;
    int32_t pad0;
    float omega_plus, omega_minus;
    float sponge_strength, sponge_width;
    float far_u[3], far_rho;
    float origin[3], dx;
    float force_scale;             /* lattice -> physical force */
    float body_position[3];
} bfo_params;

/* ---- tables ---- */
extern const int bfo_c[19][3];
extern const int bfo_opp[19];
extern const double bfo_w[19];

/* ---- fluid ---- */
float bfo_equilibrium_f32(int q, float rho, float ux, float uy, float uz);
double bfo_equilibrium_f64(int q, double rho, double ux, double uy, double uz);

/* One fused step fin -> fout.  loads (may be NULL) receives
 * {Fx, Fy, Fz, Tx, Ty, Tz, limiter activations, max limiter restriction},
 * accumulated in FP64 in ascending cell order.  density/velocity4 are written
 * only when capture_fields != 0 (velocity4 holds ux,uy,uz,0 per cell). */
void bfo_step_f32(const bfo_params* p, const float* fin, float* fout,
                  const uint8_t* solid_prev, const uint8_t* solid_cur,
                  const float* wall4, float* density, float* velocity4,
                  double* loads);
void bfo_step_f64(const bfo_params* p, const double* fin, double* fout,
                  const uint8_t* solid_prev, const uint8_t* solid_cur,
                  const float* wall4, double* density, double* velocity4,
                  double* loads);
/* The same FP64 step with an FP64 wall field (test reference gradients). */
void bfo_step_f64w(const bfo_params* p, const double* fin, double* fout,
                   const uint8_t* solid_prev, const uint8_t* solid_cur,
                   const double* wall4, double* density, double* velocity4,
                   double* loads);

/* Link fractions read by BFO_WALL_INTERPOLATED_LINKS steps: [19][N], entry
 * q * N + x for fluid x whose source x - c_q is solid (NULL: none). */
void bfo_set_link_table(const float* links);
/* an FP64 table instead (takes precedence while set; NULL clears) */
void bfo_set_link_table_f64(const double* links);

/* Equilibrium fill: solid cells at their wall velocity, fluid at far_u. */
void bfo_init_equilibrium_f32(const bfo_params* p, float* f, const uint8_t* solid, const float* wall4);
void bfo_init_equilibrium_f64(const bfo_params* p, double* f, const uint8_t* solid, const float* wall4);
/* Shear wave u_x = amplitude * sin(2 pi y / ny), rho = 1. */
void bfo_init_shear_wave_f32(const bfo_params* p, float amplitude, float* f);
void bfo_init_shear_wave_f64(const bfo_params* p, double amplitude, double* f);
/* Macroscopic fields. */
void bfo_moments_f32(const bfo_params* p, const float* f, float* density, float* velocity3);
void bfo_moments_f64(const bfo_params* p, const double* f, double* density, double* velocity3);

/* Link fraction from implicit (signed-distance) values, BirdFlow.metal:2065-2071. */
float bfo_link_fraction_f32(float fluid_distance, float solid_distance);
double bfo_link_fraction_f64(double fluid_distance, double solid_distance);
/* Bouzidi two-branch population, InterpolatedBounceBackReference.swift:28-42. */
double bfo_bouzidi_population_f64(double q, double reflected, double farther,
                                  double previous, double correction);

/* ---- geometry (FP32) ---- */
/* Frame interval for a query time: returns first frame index and blend. */
int bfo_frame_interval(const float* times, int frame_count, float t, float* blend, float* duration);
/* Per-vertex prepare from frame-major xyz points. */
void bfo_prepare_vertices(const float* points, const float* times, int frame_count,
                          int vertex_count, int first, float query_time,
                          const float translation[3], float velocity_scale,
                          float* position3, float* velocity3);
/* Raster: keys must be pre-filled with all-ones.  key_bits = 32 reproduces the
 * reference (20-bit bin << 12 | 12-bit triangle, triangles >= 4096 skipped);
 * key_bits = 64 is the widened key (bin << 32 | triangle). */
void bfo_raster(const bfo_params* p, const float* position3, const uint32_t* tri,
                int triangle_count, float half_thickness, int key_bits, void* keys);
/* Resolve keys into part mask and wall (ux,uy,uz,signed distance/dx). Cells
 * whose key is empty are left untouched: callers pre-clear mask=0 and
 * wall=(0,0,0,16). */
void bfo_resolve(const bfo_params* p, const float* position3, const float* velocity3,
                 const uint32_t* tri, const uint8_t* tri_part, int triangle_count,
                 float half_thickness, int key_bits, const void* keys,
                 uint8_t* mask, float* wall4);
void bfo_clear_geometry(const bfo_params* p, uint8_t* mask, float* wall4);

/* Exact link fractions of a meshed shell (the ray-root wall rule): for every
 * fluid cell whose pull source along q (face treatment as the step's: wrapped
 * across a periodic face, clamped at an outlet, none beyond a far field) is
 * solid, the fraction q of the link from the fluid node toward the solid node
 * at which the segment first comes within half_thickness of the mesh (the
 * union of every triangle's shell: the offset surface the reference's
 * link-ray-root audit bisects, MetalIndexedBirdSurfaceValidation.swift:
 * 10317-10438), clamped to [1e-4, 1]; links[q * N + cell], 0 elsewhere.  The
 * first entry into each triangle's shell (the triangle swept by a ball of
 * radius half_thickness) is the smallest root over its two offset faces
 * (inside the triangle), three edge cylinders (within the edge) and three
 * vertex spheres; the root is the minimum over all triangles.  A link whose
 * segment enters no shell (a solid source only by the nearest-key
 * quantization) takes the signed-distance ratio of BirdFlow.metal:2044-2076
 * from wall4.  FP32, no contraction. */
void bfo_ray_links(const bfo_params* p, const float* position3, const uint32_t* tri, int triangle_count,
                   float half_thickness, const uint8_t* mask, const float* wall4, float* links);
/* The first entry t in [0, 1] of the segment F + t D into the shell of radius
 * h around triangle ABC: 0 when F lies in it, 2 when the segment does not
 * enter it. */
float bfo_ray_shell(const float F[3], const float D[3], const float A[3], const float B[3], const float C[3],
                    float h);
/* bfo_ray_links that also reports, per link with an entry (t <= 1), the
 * first triangle attaining the least entry (0xFFFFFFFF elsewhere, and where
 * the signed-distance ratio applies). */
void bfo_ray_links_winners(const bfo_params* p, const float* position3, const uint32_t* tri, int triangle_count,
                           float half_thickness, const uint8_t* mask, const float* wall4, float* links,
                           uint32_t* winners);
/* bfo_ray_shell in FP64 (bfref_ray64.inc), and over count rays ([count][3]
 * arrays). */
double bfo_ray_shell_f64(const double F[3], const double D[3], const double A[3], const double B[3],
                         const double C[3], double h);
void bfo_ray_shell_f64_n(int count, const double* F, const double* D, const double* A, const double* B,
                         const double* C, double h, double* out);

/* Translating sphere geometry at time t (geometry center = c0 + v_geom * t). */
void bfo_translating_sphere(const bfo_params* p, const float center_radius[4],
                            const float geometry_velocity[3], const float wall_velocity[4],
                            float t, uint8_t* mask, float* wall4);

/* Link census: per-direction count of (fluid cell, q) whose pull source is
 * solid in solid_cur, and the fresh cover / uncover counts. */
void bfo_link_census(const bfo_params* p, const uint8_t* solid_prev, const uint8_t* solid_cur,
                     uint64_t link_counts[19], uint64_t* covered, uint64_t* uncovered);

/* ---- rigid body (FP64) ---- */
typedef struct bfo_body {
    double position[3];
    double orientation[4];        /* x, y, z, w */
    double linear_velocity[3];
    double angular_velocity_body[3];
} bfo_body;

void bfo_body_integrate(bfo_body* b, const double force[3], const double torque_world[3],
                        double mass, const double inertia[3], const double gravity[3],
                        double dt, int substeps);

/* Runtime safety monitor (BirdFlow.metal:6629-6707) in FP64.  metrics:
 * [0] max Mach, [1] min clearance, [2] Mach and [3] clearance at the first
 * event; event[0..1] = step low/high words, event[2] = flags (1 Mach,
 * 2 clearance, 4 non-finite state), kept from the first event on.
 * Returns this call's flags. */
uint32_t bfo_body_safety(const bfo_body* b, const double extent[3], double rotation_radius,
                         double wing_speed, double mach_limit, const double domain_min[3],
                         const double domain_max[3], const double far_velocity_world[3],
                         double velocity_to_lattice, uint64_t step, double metrics[4], uint32_t event[3]);

/* Mesh vertices carried by the body pose (FP32, the order the device
 * uses): p = x_b + R(q) p_local, v = s (v_b + R(w_b x p_local) + R v_local),
 * with p_local / v_local the body-frame frame interpolation. */
void bfo_place_vertices(const bfo_body* b, int vertex_count, float velocity_scale,
                        const float* local_position3, const float* local_velocity3,
                        float* position3, float* velocity3);

/* ---- analytic signed-distance bird (FP32) ---- */
/* GPUBirdParameters geometry and analytic kinematics (GPUData.swift:82-140):
 * lengths in metres, angles in radians, frequency in Hz. */
typedef struct bfo_bird {
    float body_radii[3];
    float wing_span, wing_root_chord, wing_tip_chord, wing_thickness, wing_sweep;
    float wing_root_offset[3];
    float tail_length, tail_half_width, tail_thickness;
    float frequency, stroke_amplitude, stroke_bias, pitch_mean, pitch_amplitude, pitch_phase;
} bfo_bird;

/* GPUPreparedBirdGeometry (GPUData.swift:325-340): 14 float4. */
typedef struct bfo_bird_frame {
    float body_position[4], orientation[4], linear_velocity[4], omega_world_radius[4];
    float left_root[4], left_chord[4], left_span[4], left_normal[4], left_angular[4];
    float right_root[4], right_chord[4], right_span[4], right_normal[4], right_angular[4];
} bfo_bird_frame;

/* sin / cos of the geometry (quadrant reduction + Taylor polynomials). */
/* Articulated mesh with Fourier joint kinematics (this build's on-device
 * kinematics; the reference drives its bird's wings by stroke, deviation,
 * pitch and tip-twist angles about a hinge, BirdFlow.metal:611-814).  Same
 * layout as bf_kinematics in include/birdflow_hip.h. */
typedef struct {
    float side;
    float root[3];
    float stroke[4][2];
    float deviation[4][2];
    float pitch[4][2];
    float twist[4][2];
} bfo_joint;
typedef struct {
    float period;
    float amplitude_scale;
    uint32_t harmonics;
    uint32_t joint_count;
    float center[3];
    float reserved;
    bfo_joint joints[8];
} bfo_kinematics;
/* Positions and velocities (velocity_scale x d pos / dt) of an articulated
 * mesh at time t (the formula documented with bf_mesh_upload_articulated). */
void bfo_articulated_vertices(const float* rest, const uint8_t* vertex_joint, const float* span_fraction,
                              int vertex_count, const bfo_kinematics* k, float t, float velocity_scale,
                              float* position3, float* velocity3);

void bfo_sincos(float x, float* s, float* c);
void bfo_sincos_n(const float* x, int n, float* s, float* c);
/* prepareBirdGeometry (BirdFlow.metal:686-814) at `time` for a body pose;
 * keyframe_count > 0: measured keyframes (17 floats each: phase, left angles
 * stroke / deviation / pitch / tip twist, left rates, right angles, right
 * rates) instead of the analytic sinusoid. */
void bfo_bird_prepare(const bfo_bird* b, const float body_position[3], const float orientation[4],
                      const float linear_velocity[3], const float angular_velocity_body[3], float time, float dx,
                      const float* keyframes, int keyframe_count, bfo_bird_frame* out);
/* buildBirdGeometry (BirdFlow.metal:816-938) over the grid of p (origin, dx):
 * part mask 1 body, 2 left wing, 3 right wing, 4 tail; wall = (velocity *
 * velocity_to_lattice, 0).  solid_prev NULL: no previous geometry. */
void bfo_bird_build(const bfo_params* p, const bfo_bird* b, const bfo_bird_frame* f, const uint8_t* solid_prev,
                    float velocity_to_lattice, uint8_t* mask, float* wall4);

/* Rigid-wing inertial properties of the prescribed-wing dynamics
 * (PrescribedWingDynamics, BirdModel.swift:98-128, packed as
 * GPUData.swift:150-176): mass (kg); centre of mass from the hinge along the
 * wing's chord, span and normal axes (m); principal inertia about those axes
 * through the centre of mass (kg m^2). */
typedef struct bfo_wing_dynamics {
    float mass;
    float center_of_mass[3];
    float inertia[3];
} bfo_wing_dynamics;

/* updateWingInertialReaction (BirdFlow.metal:6493-6572) with
 * prescribedWingMomentum (:6461-6490): each wing's linear and angular
 * momentum about the body position from the prepared frame; reaction[16] =
 * left force, left torque, right force, right torque (float4 rows, w = 0) =
 * -(momentum - previous) / dt, or 0 when initialize; previous[16] (the same
 * rows: left linear, left angular, right linear, right angular) becomes this
 * frame's momentum.  FP32, no contraction. */
void bfo_wing_reaction(const bfo_bird_frame* f, const bfo_wing_dynamics* left, const bfo_wing_dynamics* right,
                       float dt, int initialize, float previous[16], float reaction[16]);

/* ---- prescribed beta-planform flapping wing (FP32) ---- */
/* GPUFlappingWingParameters (GPUData.swift:383-388): root position and mean
 * chord; span (radius), thickness, beta exponent (shape - 1) and chord
 * normalization; cycle length, stroke half amplitude, stroke acceleration
 * and pitch durations (cycle fractions); pitch low / high angles, maximum
 * stroke rate (per cycle), reference speed.  World units and time units as
 * the grid's (the reference runs it in cells and steps). */
typedef struct bfo_wing {
    float root[3], chord_mean;
    float radius, thickness, beta_exponent, beta_normalization;
    float cycle, stroke_amplitude, stroke_duration, pitch_duration;
    float pitch_low, pitch_high, stroke_rate_max, reference_speed;
} bfo_wing;
/* GPUPreparedFlappingWing: root (w: mean chord), chord, span, normal,
 * angular velocity, state (stroke, pitch, their rates per time unit). */
typedef struct bfo_wing_frame {
    float root[4], chord[4], span[4], normal[4], angular[4], state[4];
} bfo_wing_frame;

float bfo_wing_pow(float x, float y);
/* preparePrescribedFlappingWing (BirdFlow.metal:1040-1082) at `time`. */
void bfo_wing_prepare(const bfo_wing* w, float time, bfo_wing_frame* out);
/* prescribedWingLinkFraction (BirdFlow.metal:1146-1174): bisection from the
 * fluid to the solid point, fraction measured from the fluid side. */
float bfo_wing_link_fraction(const bfo_wing* w, const bfo_wing_frame* f, float dx, const float fluid[3],
                             const float solid[3]);
/* buildPrescribedFlappingWing (BirdFlow.metal:1178-1245): mask 1 on the
 * wing, wall = (omega x r * velocity_to_lattice, -1 on the wing, +1 near
 * it); (0, 0, 0, radius) beyond the guard sphere.  solid_prev NULL: none. */
void bfo_wing_build(const bfo_params* p, const bfo_wing* w, const bfo_wing_frame* f, const uint8_t* solid_prev,
                    float velocity_to_lattice, uint8_t* mask, float* wall4);
/* The link table of the step (BirdFlow.metal:1271-1296), fluid side:
 * links[q * N + x] for fluid x whose source x - c_q is solid; 0 elsewhere. */
void bfo_wing_links(const bfo_params* p, const bfo_wing* w, const bfo_wing_frame* f, const uint8_t* mask,
                    float* links);

#ifdef __cplusplus
}
#endif
#endif
