// Host-side launcher interface between the C ABI (bf_api.hip) and the
// kernel translation units (bf_fluid.hip, bf_geometry.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bf_device.h"

namespace bf {

struct FluidBuffers {
    const void* fin;
    void* fout;
    const uint8_t* solid_prev;
    const uint8_t* solid_cur;
    const float4* wall;
    const uint8_t* static_flags;
    const uint8_t* dyn_prev;
    const uint8_t* dyn_cur;
    uint32_t* general_list;
    uint32_t* general_count;
    double* partials;
    double* block_out;
    double* last_load;
    double* history;
    float* density;
    float4* velocity;
    uint32_t* activation;   // [count, max restriction bits] of this step's fast tiles
    const double* body;     // body state (position first): torque reference point
};

struct MeshParams {
    Grid g;
    float origin[3];
    float dx;
    float half_thickness;
    float translation[3];
    float velocity_scale;
    float query_time;
    uint32_t first;            // frame interval chosen on the host
    uint32_t frame_count;
    uint32_t vertex_count;
    uint32_t triangle_count;
    uint32_t raster_count;     // triangles the raster walks: MeshBuffers::raster_list, or 0..raster_count-1
    int periodic[3];
    float cull_cells;          // key only cells at most this far (cells) from a triangle; +inf = whole AABB
    int move_with_body;        // frames are body-frame; place them with the device body pose
    uint32_t raster_blocks;    // raster grid cap (grid-stride over triangle batches)
    uint32_t raster_batch;     // triangles per raster wave batch, 1..64
    uint32_t raster_split;     // wavefronts per batch (column chunks split among them), 1..64
    uint32_t tile_pass_blocks; // tile-pass grid (grid-stride over listed tiles)
};

// Analytic signed-distance bird (BirdFlow.metal:486-533, 611-938): shape
// and analytic wing kinematics as GPUBirdParameters holds them (metres,
// radians, Hz), same field order as bf_bird_geometry.
struct BirdShape {
    float body_radii[3];
    float wing_span, wing_root_chord, wing_tip_chord, wing_thickness, wing_sweep;
    float wing_root_offset[3];
    float tail_length, tail_half_width, tail_thickness;
    float frequency, stroke_amplitude, stroke_bias, pitch_mean, pitch_amplitude, pitch_phase;
};
constexpr int BIRD_FRAME_ROWS = 14;   // GPUPreparedBirdGeometry: 14 float4

struct BirdParams {
    Grid g;
    float origin[3];
    float dx;
    float time;                // geometry time of the step (s)
    float velocity_scale;      // metres / second -> lattice
    int periodic[3];
    int bc[3][2];              // face types (BC_*), [axis][low, high]: the wing's link sources
    uint32_t tile_pass_blocks;
};

// Prescribed-wing dynamics of the analytic bird (GPUData.swift:150-176):
// per wing (0 left, 1 right) mass, centre of mass from the hinge along chord /
// span / normal, principal inertia about those axes; the step's time (s).
struct WingDynamics {
    float mass[2];
    float center_of_mass[2][3];
    float inertia[2][3];
    float dt;
    int initialize;            // the first frame: momentum only, reaction 0
};

struct BirdBuffers {
    const double* body;               // device body state (FP64)
    const float* keyframes;           // measured wing keyframes (17 floats each), or null
    uint32_t keyframe_count;          // 0: the analytic sinusoid
    float4* frame;                    // prepared frame of this step (written)
    const uint8_t* mask_prev;         // previous step's mask, or null: no previous geometry
    const uint8_t* dyn_prev;          // tiles near the previous step's solid cells, or null
    uint8_t* tiles_new;               // tiles written this step (written here)
    const uint8_t* tiles_mask_old;    // tiles whose `mask` entries an earlier step wrote
    const uint8_t* tiles_wall_old;    // tiles whose `wall` entries an earlier step wrote
    uint32_t* list;                   // scratch: tiles visited by the tile pass
    uint32_t* list_count;
    uint8_t* mask;
    float4* wall;
    uint8_t* dyn;
    // optional hinge reaction of the prescribed wings: the previous frame's
    // wing momenta (16 floats, updated) and this step's reaction (16 floats)
    float* wing_momentum;
    float* wing_reaction;
};

// Prescribed beta-planform flapping wing (BirdFlow.metal:940-1297), as
// GPUFlappingWingParameters holds it; same field order as bf_flapping_wing.
struct WingShape {
    float root[3], chord_mean;
    float radius, thickness, beta_exponent, beta_normalization;
    float cycle, stroke_amplitude, stroke_duration, pitch_duration;
    float pitch_low, pitch_high, stroke_rate_max, reference_speed;
};
constexpr int WING_FRAME_ROWS = 6;   // GPUPreparedFlappingWing: 6 float4

struct WingBuffers {
    float4* frame;                    // prepared frame of this step (written)
    const uint8_t* mask_prev;         // previous step's mask, or null: no previous geometry
    const uint8_t* dyn_prev;
    uint8_t* tiles_new;
    const uint8_t* tiles_mask_old;
    const uint8_t* tiles_wall_old;
    uint32_t* list;
    uint32_t* list_count;
    uint8_t* mask;
    float4* wall;
    uint8_t* dyn;
    float* links;                     // [tile][19][256] link fractions of fluid cells with a solid source
};

// Fourier joint kinematics of an articulated mesh (layout of bf_kinematics).
struct KinJoint {
    float side;
    float root[3];
    float stroke[4][2];
    float deviation[4][2];
    float pitch[4][2];
    float twist[4][2];
};
struct Kinematics {
    float period;
    float amplitude_scale;
    uint32_t harmonics;
    uint32_t joint_count;
    float center[3];
    float reserved;
    KinJoint joints[8];
};

struct MeshBuffers {
    const double* body;                 // body state, read when MeshParams::move_with_body
    const float* points;
    const float* times;
    // articulated mesh (instead of frames): rest positions [V][3], joint of
    // each vertex, twist weight, kinematics (device copy)
    const float* rest;
    const uint8_t* vertex_joint;
    const float* span_fraction;
    const Kinematics* kinematics;
    const uint32_t* tri;
    const uint8_t* tri_part;
    float4* pos;
    float4* vel;
    unsigned long long* keys;
    uint8_t* tiles_new;                 // tiles keyed this step (written here)
    const uint8_t* tiles_mask_old;      // tiles whose `mask` entries an earlier step wrote
    const uint8_t* tiles_wall_old;      // tiles whose `wall` entries an earlier step wrote
    uint32_t* list;                     // scratch: tiles visited by the tile pass
    uint32_t* list_count;
    uint8_t* mask;
    float4* wall;
    uint8_t* dyn;
    uint32_t* tri_out;                  // optional: nearest triangle of each resolved cell
    // Optional: the raster walks only the listed triangles (ascending ids),
    // and the keys of the others, constant from step to step, are seeded
    // from a cache of (cell, key) pairs before it runs.
    const uint32_t* raster_list;
    const uint32_t* seed_cells;
    const unsigned long long* seed_keys;
    uint32_t seed_count;
};

// Exact link fractions against the meshed shell (bf_rays.hip): the table
// [tile][19][256] of the step, read by the step as a link table.
struct RayParams {
    Grid g;
    float origin[3];
    float dx;
    float half_thickness;
    int bc[3][2];
    uint32_t triangle_count;
    uint32_t triangle_blocks;  // triangle pass grid (one triangle per single-wavefront block, grid-stride)
    uint32_t split;            // wavefronts per triangle (its cells split among them), 1..64
    uint32_t tile_pass_blocks; // init / finalize grid (grid-stride over listed tiles)
};

struct RayBuffers {
    const float4* pos;       // prepared vertices of the step
    const uint32_t* tri;
    const uint8_t* mask;     // the step's part mask
    const float4* wall;      // the step's wall field (signed distance in .w)
    const uint8_t* dyn;      // tiles whose cells pull from a solid cell
    uint32_t* list;          // scratch: listed tiles
    uint32_t* list_count;
    uint32_t* dirs;          // per cell: link directions pulling from it (zero between steps)
    float* links;            // the step's table
    // optional (the adjoint): per table entry, the triangle whose shell
    // gives the entry (NO_WINNER where none lies in the link); same layout
    uint32_t* winners;
    // optional static cache (launch_ray_static): the static triangles'
    // entries per (source cell, q), merged each step, so that only the
    // moving triangles are walked
    const uint4* static_entries;
    uint32_t static_count;
    const uint32_t* dynamic_list;
    uint32_t dynamic_count;
};
constexpr uint32_t NO_WINNER = 0xFFFFFFFFu;
// Builds the static cache: scratch_table ([19][N] u32, filled with the bits
// of 2.0f by the caller) takes the static triangles' entries per source
// cell and direction, then its entries are listed in out (at most capacity;
// *count receives their number, which may exceed capacity).
hipError_t launch_ray_static(const float4* pos, const uint32_t* tri, const uint32_t* static_list,
                             uint32_t static_count, uint32_t* scratch_table, uint4* out, uint32_t capacity,
                             uint32_t* count, const RayParams& p, hipStream_t s);

// Adjoint of one step (bf_adjoint.hip).  lam_in = dJ/df_s, lam_out =
// dJ/df_{s-1}; f_prev = f_{s-1}; masks and wall of step s-1 / s; nu:
// per listed tile 19 x 256 floats; dW: [K][N] wall tangents of step s.
struct AdjointBuffers {
    const void* f_prev;           // the step's input populations, stored form of the storage
    const float* lam_in;
    float* lam_out;
    const uint8_t* solid_prev;
    const uint8_t* solid_cur;
    const float4* wall;
    const uint8_t* static_flags;
    const uint8_t* dyn_prev;
    const uint8_t* dyn_cur;
    const uint32_t* general_list;
    const uint32_t* general_count;
    float* nu;
    const float4* dW;
    double* grad_partials;        // [listed tile][8]
    double* grad_total;           // [8], accumulated over steps
    const double* body;
    float* outlet_extra;          // [x face 0 / 1][19][ny * nz]: clamped pulls at outlet x faces
    // exact-root wall rule (StepParams::wall_rule is then
    // WALL_INTERPOLATED_LINKS): the step's link table, the triangle giving
    // each entry (NO_WINNER: the signed-distance ratio), the vertex
    // positions of the step and their tangents [K][vertices], the triangles
    // and the shell radius
    const float* links;
    const uint32_t* winners;
    const float4* vpos;
    const float4* dvpos;
    const uint32_t* tri;
    uint32_t vertex_count;
    float half_thickness;
};
struct AdjointParams {
    float cF[3], cT[3];           // load weights of the step: J_s = cF . F + cT . T
    int K;                        // parameters (<= 8)
};
hipError_t launch_adjoint_step(int coll, int storage, const AdjointBuffers& b, const StepParams& p,
                               const AdjointParams& w, int general_blocks, hipStream_t s);
// Kinematic parameters of an articulated mesh (layout of
// bf_kinematic_parameter): coefficient [harmonic][component] of series
// `channel` (stroke, deviation, pitch, twist) of every joint in joint_mask.
struct KinParam {
    uint32_t joint_mask, channel, harmonic, component;
};
struct KinParams {
    KinParam p[8];
};
// Vertex tangents of step time m.query_time (from the frame tangents
// dpoints, or for an articulated mesh (mb.kinematics) from kp), placed by
// the body pose when m.move_with_body, then the wall-field tangents dW of
// the listed tiles' keyed cells.
hipError_t launch_wall_tangent(const uint32_t* list, const uint32_t* count, int blocks, const uint32_t* tri_of_cell,
                               const uint32_t* tri, const float4* pos, const float4* vel, const MeshBuffers& mb,
                               const float* dpoints, const KinParams& kp, float4* dpos, float4* dvel, float4* dW,
                               const MeshParams& m, int K, hipStream_t s);

// Rigid body (FP64): integration after a step's loads and the safety monitor.
struct BodyParams {
    double mass;
    double inertia[3];
    double gravity[3];
    double external_force[3];
    double external_torque[3];
    double dt;                       // world time per fluid step
    int substeps;
    int monitor;
    double extent[3];                // body-frame half extents for the clearance check
    double rotation_radius;
    double wing_speed;
    double mach_limit;
    double domain_min[3], domain_max[3];
    double far_velocity_world[3];
    double velocity_to_lattice;
};
// wing_reaction (optional): the step's prescribed-wing hinge reaction, 16
// floats (left force, left torque, right force, right torque; xyz + 0),
// added to the fluid load.
hipError_t launch_body_step(double* state, const double* load, double* history, uint32_t history_slot,
                            double* metrics, uint32_t* event, uint64_t step, const BodyParams& bp, hipStream_t s,
                            const float* wing_reaction = nullptr);

hipError_t launch_fluid_fast(int arith, int storage, int coll, const FluidBuffers& b, const StepParams& p,
                             hipStream_t s);
hipError_t launch_fluid_general(int arith, int storage, int coll, const FluidBuffers& b, const StepParams& p,
                                int general_blocks, hipStream_t s);
// Per translation unit: fused arithmetic for every storage, strict
// arithmetic for FP32 storage.
hipError_t launch_fluid_fast_fused(int storage, int coll, const FluidBuffers& b, const StepParams& p,
                                   hipStream_t s);
hipError_t launch_fluid_general_fused(int storage, int coll, const FluidBuffers& b, const StepParams& p,
                                      int general_blocks, hipStream_t s);
hipError_t launch_fluid_fast_strict(int coll, const FluidBuffers& b, const StepParams& p, hipStream_t s);
hipError_t launch_fluid_general_strict(int coll, const FluidBuffers& b, const StepParams& p, int general_blocks,
                                       hipStream_t s);
hipError_t launch_load_reduction(const FluidBuffers& b, const StepParams& p, uint32_t history_slot, hipStream_t s);
hipError_t launch_build_general_list(const uint8_t* static_flags, const uint8_t* dyn_prev, const uint8_t* dyn_cur,
                                     uint32_t tiles, uint32_t* list, uint32_t* count, hipStream_t s);
hipError_t launch_init_equilibrium(int storage, void* f, const uint8_t* solid, const float4* wall,
                                   const float* far_u, const Grid& g, hipStream_t s);
hipError_t launch_init_shear_wave(int storage, void* f, float amplitude, const Grid& g, hipStream_t s);
hipError_t launch_extract_fields(int storage, const void* f, float* rho, float* u, const Grid& g, hipStream_t s);
// host layout (direction-major FP32) <-> device population buffer
// Stored values (4 or 2 bytes by storage) between host direction-major and
// device row-interleaved order, unchanged.
hipError_t launch_permute_populations(int storage, bool to_device, const void* src, void* dst, size_t count,
                                      const Grid& g, hipStream_t s);
// The population minimum (bf_fluid.hip): comparison value (a non-finite
// value compares as -infinity), the value itself, its direction-major index
// and whether it is non-finite.
struct PopulationMinimum {
    float comparison;
    float value;
    unsigned long long index;
    uint32_t non_finite;
    uint32_t pad;
};
constexpr uint32_t POPULATION_MINIMUM_BLOCKS = 4096;
hipError_t launch_population_minimum(int storage, const void* f, const Grid& g, PopulationMinimum* partials,
                                     uint32_t blocks, PopulationMinimum* out, hipStream_t s);
hipError_t launch_poke_population(int storage, void* f, const Grid& g, unsigned long long index, float value,
                                  hipStream_t s);
hipError_t launch_convert_populations(int storage, bool pack, const void* src, void* dst, size_t count,
                                      const Grid& g, hipStream_t s);
hipError_t launch_mesh_geometry(const MeshBuffers& b, const MeshParams& m, hipStream_t s);
hipError_t launch_ray_links(const RayBuffers& b, const RayParams& p, hipStream_t s);
// Keys of the listed (static) triangles alone: vertex prepare, then the
// raster of b.raster_list into b.keys (tiles into b.tiles_new), then the
// number of keyed cells into *count.
hipError_t launch_mesh_static_keys(const MeshBuffers& b, const MeshParams& m, unsigned long long* count,
                                   hipStream_t s);
// Moves every key of `keys` (n cells) into (cells, out) pairs, at most
// `capacity`, and empties it; *counter ends at the number of keys.
hipError_t launch_collect_keys(unsigned long long* keys, uint32_t n, uint32_t* cells, unsigned long long* out,
                               uint32_t capacity, unsigned long long* counter, hipStream_t s);
hipError_t launch_mesh_prepare(const MeshBuffers& b, const MeshParams& m, hipStream_t s);

// Momentum ledger (bf_ledger.hip).  sources: the completed step's change,
// open-boundary, sponge and persistent-link sums (16 doubles) and counts (4);
// otherwise the mass and momentum of f_out over the fluid of solid_cur (4
// doubles).  partial_v / partial_n hold blocks x 16 / blocks x 4 entries.
struct LedgerBuffers {
    const void* f_in;
    const void* f_out;
    const uint8_t* solid_prev;
    const uint8_t* solid_cur;
    const float4* wall;
    uint32_t blocks;
    double* partial_v;
    unsigned long long* partial_n;
    double* out_v;
    unsigned long long* out_n;
};
hipError_t launch_ledger(int storage, const LedgerBuffers& b, const StepParams& p, bool sources, hipStream_t s);
// Block count of the ledger kernels (grid-stride over cells): at most this.
constexpr uint32_t LEDGER_SCRATCH_BLOCKS = 4096;
// The no-allocation ledger's file-scope device scratch (bf_ledger.hip):
// setup looks up its address on `device` (the current device) once;
// buffers fills b's partials and outputs from it for `blocks` blocks
// (hipErrorInvalidValue beyond LEDGER_SCRATCH_BLOCKS, hipErrorNotInitialized
// before setup on that device).
hipError_t ledger_scratch_setup(int device);
hipError_t ledger_scratch_buffers(int device, uint32_t blocks, LedgerBuffers& b);
hipError_t launch_bird_geometry(const BirdShape& shape, const BirdBuffers& b, const BirdParams& p, hipStream_t s,
                                const WingDynamics* wing_dynamics = nullptr);
hipError_t launch_wing_geometry(const WingShape& shape, const WingBuffers& b, const BirdParams& p, hipStream_t s);
hipError_t launch_sphere_geometry(uint8_t* mask, float4* wall, uint8_t* dyn, const Grid& g, float4 cr, float3 gv,
                                  float4 wv, float t, const int periodic[3], hipStream_t s);
hipError_t launch_mark_tiles(const uint8_t* mask, uint8_t* dyn, const Grid& g, const int periodic[3], bool clear,
                             hipStream_t s);
hipError_t launch_fill_u8(uint8_t* p, uint8_t v, uint32_t n, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, uint32_t n, hipStream_t s);
hipError_t launch_fill_keys(unsigned long long* p, uint32_t n, hipStream_t s);
hipError_t launch_fill_wall(float4* p, uint32_t n, hipStream_t s);
hipError_t launch_link_census(const uint8_t* prev, const uint8_t* cur, const Grid& g, const int bc[3][2],
                              unsigned long long* counts, hipStream_t s);

}  // namespace bf
