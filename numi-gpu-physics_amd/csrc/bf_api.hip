// C ABI of libbirdflow_hip.so (declared in include/birdflow_hip.h).
//
// Owns the device state of one simulation and sequences each step the way
// BirdFlowSimulation.encodeBatch does (Sources/BirdFlowMetal/
// BirdFlowSimulation.swift:1359-1449): geometry of step n+1 into the next
// mask, fluid step current -> next lattice with the previous / current
// masks, load reduction, swap.
//
// Two HIP streams form a software pipeline, per step s:
//   stream A  fluid_fast(s)                              (HBM-bound, wall-free tiles)
//   stream B  geometry(s+1) + tile list(s+1)  [overlaps fast(s)],
//             then fluid_general(s)                       [after fast(s)]
//             and the load reduction of step s            [overlaps fast(s+1)]
// fast(s) waits for event R[(s-1)&1] (general(s-1) done, which follows
// geometry(s) on B); general(s) waits for F[s&1] (fast(s) done).  Geometry of
// step s+1 only writes buffers that steps s-1 and s-2 last read: masks and
// solid-proximity tile flags rotate over 3 buffers, the wall field and the
// general-tile list over 2, the touched-tile flags over 4.  Every entry point
// other than bf_step drains both streams first.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/birdflow_hip.h"
#include "bf_launch.h"

using namespace bf;

static_assert(BF_STORAGE_FP32 == STORE_FP32 && BF_STORAGE_FP16_OFFSET == STORE_FP16_OFFSET &&
                  BF_STORAGE_BF16_OFFSET == STORE_BF16_OFFSET, "storage ids");
static_assert(sizeof(bf_kinematics) == sizeof(Kinematics) && sizeof(bf_joint) == sizeof(KinJoint) &&
                  offsetof(bf_kinematics, joints) == offsetof(Kinematics, joints) && BF_MAX_JOINTS == 8 &&
                  BF_MAX_HARMONICS == 4,
              "kinematics layouts differ");
static_assert(sizeof(bf_kinematic_parameter) == sizeof(KinParam), "kinematic parameter layouts differ");
static_assert(BF_COLLISION_RR3 == COLL_RR3 && BF_WALL_INTERPOLATED_HALF == WALL_INTERPOLATED_HALF, "operator ids");

namespace {
constexpr uint32_t HISTORY_CAPACITY = 1u << 16;
constexpr uint32_t PROFILE_CAPACITY = 4096;
constexpr int PROFILE_EVENTS = 6;   // A: fast start/end; B: geometry start/end, general start/end

struct DeviceAlloc {
    std::vector<void*> ptrs;
    uint64_t bytes = 0;
    void* take(hipError_t& err, size_t n) {
        void* p = nullptr;
        if (err != hipSuccess) return nullptr;
        err = hipMalloc(&p, n == 0 ? 1 : n);
        if (err == hipSuccess) {
            ptrs.push_back(p);
            bytes += n;
        }
        return p;
    }
    void release() {
        for (void* p : ptrs) (void)hipFree(p);
        ptrs.clear();
        bytes = 0;
    }
};

// Arithmetic of the collision kernels for a configuration.
inline int arithmetic_of(const bf_config& cfg) {
    if (cfg.arithmetic == BF_ARITHMETIC_STRICT) return ARITH_STRICT;
    if (cfg.arithmetic == BF_ARITHMETIC_FUSED) return ARITH_FUSED;
    return cfg.storage == BF_STORAGE_FP32 ? ARITH_STRICT : ARITH_FUSED;
}

inline int mod3(int64_t s) { return (int)(((s % 3) + 3) % 3); }
inline int mod4(int64_t s) { return (int)(((s % 4) + 4) % 4); }
inline int par2(int64_t s) { return (int)(s & 1); }
}  // namespace

struct bf_context {
    bf_config cfg{};
    Grid g{};
    hipStream_t sa = nullptr;             // fast fluid
    hipStream_t sb = nullptr;             // general fluid, reductions, geometry, everything else
    hipEvent_t ev_fast[2] = {nullptr, nullptr};
    hipEvent_t ev_ready[2] = {nullptr, nullptr};
    hipEvent_t ev_general[2] = {nullptr, nullptr};   // general(s) done (stream A)
    int device = 0;
    DeviceAlloc alloc;
    DeviceAlloc mesh_alloc;
    size_t pop_elem = 4;
    void* f[2] = {nullptr, nullptr};
    int cur = 0;                          // f[cur] holds the current lattice
    uint8_t* mask[3] = {nullptr, nullptr, nullptr};
    float4* wall[2] = {nullptr, nullptr};
    uint8_t* dyn[3] = {nullptr, nullptr, nullptr};
    uint8_t* static_flags = nullptr;
    uint8_t* tflags[4] = {nullptr, nullptr, nullptr, nullptr};   // tiles keyed at step s -> tflags[s % 4]
    uint8_t* no_tiles = nullptr;          // all-zero tile flags
    uint32_t* general_list[2] = {nullptr, nullptr};
    uint32_t* geo_list = nullptr;
    uint32_t* counters = nullptr;         // [0],[1] general-list counts by step parity, [2] geometry tile list
    double* partials = nullptr;
    double* block_out = nullptr;
    double* last_load = nullptr;
    double* history = nullptr;
    unsigned long long* census = nullptr;
    float* density = nullptr;
    float4* velocity = nullptr;
    float* staging = nullptr;             // host-layout FP32 buffer for population transfers
    // geometry
    int source = BF_GEOMETRY_STATIC;
    bool static_prev_differs = false;
    float time_start = 0.0f;              // time of step 0
    int64_t geometry_ready = 0;           // highest step whose geometry is enqueued
    // mesh
    float* m_points = nullptr;
    float* m_times = nullptr;
    // articulated mesh (instead of frames): rest positions, joints, twist
    // weights and the kinematics on the device
    bool m_articulated = false;
    float* m_rest = nullptr;
    uint8_t* m_joint = nullptr;
    float* m_span = nullptr;
    Kinematics* m_kin = nullptr;
    uint32_t m_joint_count = 0, m_harmonics = 0;
    uint32_t* m_tri = nullptr;
    uint8_t* m_part = nullptr;
    float4* m_pos = nullptr;
    float4* m_vel = nullptr;
    unsigned long long* keys = nullptr;
    std::vector<float> host_times;
    uint32_t m_frames = 0, m_vertices = 0, m_triangles = 0;
    float m_half = 0.0f;
    float m_translation[3] = {0, 0, 0};
    // static triangles (the same vertex positions in every frame): their keys
    // within the stepping band are cached once as (cell, key) pairs and seeded
    // each step, and the raster walks the moving triangles only
    uint32_t* m_dyn_list = nullptr;        // ascending ids of the moving triangles
    uint32_t* m_static_list = nullptr;     // ascending ids of the static ones
    uint32_t m_dyn_count = 0, m_static_count = 0;
    bool static_keys_enabled = true;       // BIRDFLOW_STATIC_KEYS=0 rasters every triangle every step
    bool static_ready = false;
    DeviceAlloc static_alloc;
    uint32_t* static_cells = nullptr;
    unsigned long long* static_keys = nullptr;
    uint32_t static_key_count = 0;
    // momentum ledger: block partials and results (allocated on first use);
    // the last step's inputs are intact while no population writer ran
    DeviceAlloc ledger_alloc;
    double* ledger_pv = nullptr;
    unsigned long long* ledger_pn = nullptr;
    double* ledger_out = nullptr;
    unsigned long long* ledger_out_n = nullptr;
    uint32_t ledger_blocks = 0;
    const uint8_t* ledger_prev = nullptr;   // previous-step mask of the last step
    uint64_t f_writes = 0;                  // population writes other than steps
    uint64_t ledger_epoch = ~0ull;          // f_writes when the last step ran
    int64_t ledger_step = -1;               // and its index
    // prescribed flapping wing: shape, prepared frames, link-fraction table
    WingShape wing{};
    float4* wing_frame = nullptr;         // [2][WING_FRAME_ROWS]
    DeviceAlloc wing_alloc;
    float* wing_links[2] = {nullptr, nullptr};   // [step parity][tiles][19][256]
    // exact link fractions of a mesh (BF_WALL_INTERPOLATED_RAY): the table
    // of each step parity and the scratch list of the tiles it covers
    DeviceAlloc ray_alloc;
    float* ray_links[2] = {nullptr, nullptr};    // [step parity][tiles][19][256]
    uint32_t* ray_list = nullptr;
    uint32_t* ray_count = nullptr;
    uint32_t* ray_dirs = nullptr;                // per cell: link directions pulling from it
    // the static triangles' entries per (source cell, q), built once per
    // mesh when static triangles exist (BIRDFLOW_RAY_STATIC=0 walks every
    // triangle every step)
    DeviceAlloc ray_static_alloc;
    uint4* ray_static = nullptr;
    uint32_t ray_static_count = 0;
    bool ray_static_ready = false;
    bool ray_static_enabled = true;
    // analytic bird: shape and the prepared frame of each step parity
    BirdShape bird{};
    float4* bird_frame = nullptr;         // [2][BIRD_FRAME_ROWS]
    DeviceAlloc keyframe_alloc;           // measured wing keyframes, 17 floats each
    float* bird_keyframes = nullptr;
    uint32_t bird_keyframe_count = 0;
    // prescribed-wing dynamics of the analytic bird: hinge reaction added to
    // the body's load in free flight
    bool wing_dynamics = false;
    WingDynamics wing_dyn{};
    // fluid health: the non-finite population guard of the fluid kernels
    // ([0] flags, [1] first step) and the population-minimum reduction
    unsigned long long* health = nullptr;
    bool health_reported = false;
    PopulationMinimum* popmin = nullptr;  // [POPULATION_MINIMUM_BLOCKS + 1]
    float* wing_momentum = nullptr;       // [16] previous frame's wing momenta
    float* wing_reaction = nullptr;       // [2][16] reaction of step s in slot s & 1
    // sphere
    float4 sphere_cr{};
    float3 sphere_gv{};
    float4 sphere_wv{};
    // rigid body: FP64 device state (position[3], orientation[4] x y z w,
    // linear velocity[3], body angular velocity[3]), per-step history,
    // safety record
    double* body = nullptr;
    double* body_history = nullptr;
    double* safety_metrics = nullptr;     // max Mach, min clearance, Mach / clearance at the first event
    uint32_t* safety_event = nullptr;     // step low, step high, flags
    BodyParams body_params{};
    bool free_flight = false;
    bool move_mesh = false;
    uint32_t safety_reported = 0;
    uint64_t body_history_count = 0;
    // adjoint: parameter tangents of the mesh frames and sweep buffers
    struct {
        DeviceAlloc alloc;                // sweep buffers (sized by the checkpoint interval)
        DeviceAlloc tangent_alloc;
        uint32_t K = 0;
        float* dpoints = nullptr;         // [K][frames][vertices][3]
        bool kinematic = false;           // K kinematic parameters of an articulated mesh instead
        KinParams kin{};
        uint32_t interval = 0, segments = 0;
        uint64_t verts = 0;               // vertex capacity of dpos, dvel, gpos, gvel
        uint8_t* checkpoints = nullptr;   // [segments][19N], stored form (pop_elem bytes each)
        uint8_t* seg_f = nullptr;         // [interval][19N]: f_a .. f_{b-1}, stored form
        // geometry of steps a .. b of a segment, two sets: the reverse sweep
        // of one segment (stream B) reads one while the geometry of the next
        // segment is built into the other (stream A)
        uint8_t* seg_mask[2] = {nullptr, nullptr};    // [interval + 1][N]
        float4* seg_wall[2] = {nullptr, nullptr};
        uint32_t* seg_tri[2] = {nullptr, nullptr};
        uint8_t* seg_dyn[2] = {nullptr, nullptr};     // [interval + 1][tiles]
        uint8_t* slot_tiles[2] = {nullptr, nullptr};  // [2][interval + 1][tiles]: touched tiles per build
        int gen[2] = {0, 0};                          // which half of slot_tiles the next build writes
        uint32_t* slot_list[2] = {nullptr, nullptr};  // [interval + 1][tiles]: general list of steps (k-1, k)
        uint32_t* slot_count[2] = {nullptr, nullptr}; // [interval + 1]
        float4* gpos = nullptr;                       // vertex positions / velocities of that build
        float4* gvel = nullptr;
        hipEvent_t ev_geo[2] = {nullptr, nullptr};    // set built (stream A)
        hipEvent_t ev_done[2] = {nullptr, nullptr};   // set released (stream B)
        float* lam[2] = {nullptr, nullptr};
        int lam_cur = 0;
        bool have_sensitivity = false;
        float* nu = nullptr;              // [tiles][19][256]
        float4* dW[2] = {nullptr, nullptr};           // [8][N], by step parity
        hipStream_t sc = nullptr;                     // wall tangents (third stream, default priority)
        hipEvent_t ev_tan[2] = {nullptr, nullptr};    // dW[j] written (stream C)
        hipEvent_t ev_used[2] = {nullptr, nullptr};   // dW[j] read (stream B)
        float* outlet_extra = nullptr;    // [2][19][ny nz]
        // by step parity (stream C writes step s while stream B reads step
        // s + 1): vertex positions / velocities of the step and their
        // tangents [8][vertices]
        float4* vpos[2] = {nullptr, nullptr};
        float4* vvel[2] = {nullptr, nullptr};
        float4* dpos[2] = {nullptr, nullptr};
        float4* dvel[2] = {nullptr, nullptr};
        // exact-root wall rule: the recompute's vertices (stream B); per step
        // parity the reverse step's link table and entry triangles, and
        // stream C's own scratch of the ray pass
        float4* rpos = nullptr;
        float4* rvel = nullptr;
        float* rtab[2] = {nullptr, nullptr};
        uint32_t* rwin[2] = {nullptr, nullptr};
        uint32_t* rdirs = nullptr;
        uint32_t* rlist = nullptr;
        uint32_t* rcount = nullptr;
        double* grad_partials = nullptr;  // [tiles][8]
        double* grad_total = nullptr;     // [8]
    } adj;
    // stepping
    int64_t step = 0;
    uint64_t history_count = 0;
    bool invalidated = false;
    std::string error;
    // profiling
    bool profiling = false;
    int profiling_level = 0;              // 1: fast kernel only; 2: also general and geometry
    std::vector<hipEvent_t> events;
    uint32_t profiled = 0;
    int general_blocks = 1024;
    // replayed part masks (finite-difference checks with the link set fixed)
    DeviceAlloc frozen_alloc;
    uint8_t* frozen = nullptr;
    uint32_t frozen_count = 0;
    int64_t frozen_first = 0;
    // geometry grids: bounded so that the geometry of step s+1 leaves the
    // fluid kernel of step s most of the machine
    uint32_t raster_blocks = 1u << 30;
    uint32_t raster_batch = 0;            // triangles per raster wave; 0: from the mesh size
    int fast_priority = 0, general_priority = 0;   // s_setprio of the fluid kernels' wavefronts
    bool pair_cells = true;                        // 2-byte storage: fast tiles two cells per lane
    bool geometry_after_general = true;            // stream B starts the next geometry after general(s)
    uint32_t raster_batch_cap = 32;                // largest automatic raster batch
    uint32_t compute_units = 1;
    uint32_t tile_pass_blocks = 512;
    uint32_t ray_blocks_per_cu = 64;        // triangle pass: single-wavefront blocks per CU
    uint32_t ray_tile_blocks_per_cu = 8;    // exact-root init / finalize: 256-lane blocks per CU
};

namespace {

bf_status fail(bf_context* c, bf_status st, const std::string& msg) {
    if (c) c->error = msg;
    return st;
}

bf_status hip_fail(bf_context* c, hipError_t e, const char* what, bool invalidate = true) {
    if (c) {
        c->error = std::string(what) + ": " + hipGetErrorString(e);
        if (invalidate) c->invalidated = true;
    }
    return BF_ERROR_LAUNCH_FAILED;
}

bool periodic_axis(const bf_config& c, int a) {
    return c.boundary[a][0] == BF_BC_PERIODIC && c.boundary[a][1] == BF_BC_PERIODIC;
}

hipError_t drain(bf_context* c) {
    hipError_t e = hipStreamSynchronize(c->sa);
    hipError_t e2 = hipStreamSynchronize(c->sb);
    hipError_t e3 = c->adj.sc ? hipStreamSynchronize(c->adj.sc) : hipSuccess;
    return e != hipSuccess ? e : (e2 != hipSuccess ? e2 : e3);
}

#define BF_ENTER(c)                                                                                \
    do {                                                                                           \
        if (!(c)) return BF_ERROR_INVALID_ARGUMENT;                                                \
        if ((c)->invalidated) return BF_ERROR_STATE_INVALIDATED;                                   \
        (void)hipSetDevice((c)->device);                                                           \
        hipError_t drain_err_ = drain(c);                                                          \
        if (drain_err_ != hipSuccess) return hip_fail((c), drain_err_, "earlier device work");     \
    } while (0)

StepParams make_step_params(const bf_context* c) {
    StepParams p{};
    p.g = c->g;
    p.fast_priority = c->fast_priority;
    p.general_priority = c->general_priority;
    p.pair_cells = c->pair_cells ? 1 : 0;
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < 2; ++s) p.bc[a][s] = (int)c->cfg.boundary[a][s];
    // the exact-root rule reads its per-step table like the wing's
    p.wall_rule = c->cfg.wall_rule == BF_WALL_INTERPOLATED_RAY ? WALL_INTERPOLATED_LINKS : (int)c->cfg.wall_rule;
    p.force_mode = (int)c->cfg.force_mode;
    p.load_component = (int)c->cfg.load_component;
    p.selected_part = (int)c->cfg.selected_part;
    p.accumulate_loads = c->cfg.accumulate_loads ? 1 : 0;
    p.capture_fields = 0;
    p.omega_plus = c->cfg.omega_plus;
    p.omega_minus = c->cfg.omega_minus;
    p.sponge_strength = c->cfg.sponge_strength;
    p.sponge_width = c->cfg.sponge_width;
    for (int i = 0; i < 3; ++i) {
        p.far_u[i] = c->cfg.far_velocity[i];
        p.origin[i] = c->cfg.origin[i];
    }
    p.far_rho = c->cfg.far_density;
    p.dx = c->cfg.cell_size;
    p.force_scale = c->cfg.force_to_physical;
    for (int q = 0; q < 19; ++q) p.far_eq[q] = equilibrium<float>(q, p.far_rho, p.far_u[0], p.far_u[1], p.far_u[2]);
    return p;
}

// Tiles that always take the general path: partial tiles at the +x / +y
// edges (the fast kernel streams whole tiles; the sponge band streams there
// too).
std::vector<uint8_t> static_tile_flags(const bf_config& cfg, const Grid& g) {
    (void)cfg;
    std::vector<uint8_t> flags(g.tiles, 0);
    for (int z = 0; z < g.nz; ++z)
        for (int ty = 0; ty < g.tiles_y; ++ty)
            for (int tx = 0; tx < g.tiles_x; ++tx) {
                const int x0 = tx * TILE_X, y0 = ty * TILE_Y;
                const bool general = x0 + TILE_X > g.nx || y0 + TILE_Y > g.ny;
                flags[((size_t)z * g.tiles_y + ty) * g.tiles_x + tx] = general ? 1 : 0;
            }
    return flags;
}

// Geometry views of step s.
struct StepGeometry {
    const uint8_t* mask_cur;
    const uint8_t* mask_prev;
    const float4* wall;
    const uint8_t* dyn_cur;
    const uint8_t* dyn_prev;
};

StepGeometry geometry_of(const bf_context* c, int64_t s) {
    if (c->source == BF_GEOMETRY_STATIC) {
        (void)s;
        const bool differs = c->static_prev_differs;   // only the first step after bf_set_geometry
        return {c->mask[0], differs ? c->mask[1] : c->mask[0], c->wall[0], c->dyn[0], differs ? c->dyn[1] : c->dyn[0]};
    }
    return {c->mask[mod3(s)], c->mask[mod3(s - 1)], c->wall[par2(s)], c->dyn[mod3(s)], c->dyn[mod3(s - 1)]};
}

// Raster launch shape for `count` triangles: enough batches for two waves
// per CU, at most raster_batch_cap triangles each (one gather round trip
// serves many triangles of a large mesh, while a small mesh of large
// triangles spreads one triangle per wave); a mesh of fewer triangles than
// that splits each triangle's box among several wavefronts.
void set_raster_shape(const bf_context* c, MeshParams& m, uint32_t count) {
    m.raster_count = count;
    m.raster_batch = c->raster_batch ? c->raster_batch
                                     : std::min(c->raster_batch_cap, std::max(1u, count / (2u * c->compute_units)));
    const uint32_t batches = (count + m.raster_batch - 1) / std::max(1u, m.raster_batch);
    m.raster_split = std::min(64u, std::max(1u, (2u * c->compute_units) / std::max(1u, batches)));
}

MeshParams mesh_params(const bf_context* c, float t, bool exact) {
    MeshParams m{};
    m.g = c->g;
    for (int i = 0; i < 3; ++i) {
        m.origin[i] = c->cfg.origin[i];
        m.translation[i] = c->m_translation[i];
        m.periodic[i] = periodic_axis(c->cfg, i) ? 1 : 0;
    }
    m.dx = c->cfg.cell_size;
    m.half_thickness = c->m_half;
    m.velocity_scale = c->cfg.velocity_to_lattice;
    m.query_time = t;
    m.frame_count = c->m_frames;
    m.vertex_count = c->m_vertices;
    m.triangle_count = c->m_triangles;
    m.raster_count = c->m_triangles;
    // Stepping keys only the band that the fluid step reads: solid cells
    // (distance <= h), fluid cells one link away from them (<= h + sqrt(2)),
    // and cells uncovered this step (<= h + one step's wall travel).
    m.cull_cells = exact ? INFINITY : c->m_half / c->cfg.cell_size + 1.5f;
    m.move_with_body = c->move_mesh ? 1 : 0;
    m.raster_blocks = c->raster_blocks;
    set_raster_shape(c, m, c->m_triangles);
    m.tile_pass_blocks = c->tile_pass_blocks;
    if (c->m_articulated) return m;   // no frames: the kinematics give every time
    // frame interval, MeasuredBirdSurfaceSequence.swift:174-204
    const std::vector<float>& ts = c->host_times;
    const int last = (int)c->m_frames - 1;
    uint32_t first;
    if (t <= ts[0]) first = 0;
    else if (t >= ts[last]) first = (uint32_t)(last - 1);
    else {
        int lo = 0, hi = last;
        while (lo + 1 < hi) {
            const int mid = lo + (hi - lo) / 2;
            if (t < ts[mid]) hi = mid;
            else lo = mid;
        }
        first = (uint32_t)lo;
    }
    m.first = first;
    return m;
}

float time_of(const bf_context* c, int64_t s) {
    return c->time_start + (float)((double)s * c->cfg.time_step);
}

// The mesh arrays of either upload form.
MeshBuffers mesh_buffers(const bf_context* c) {
    MeshBuffers b{};
    b.body = c->body;
    b.points = c->m_points;
    b.times = c->m_times;
    b.rest = c->m_rest;
    b.vertex_joint = c->m_joint;
    b.span_fraction = c->m_span;
    b.kinematics = c->m_articulated ? c->m_kin : nullptr;
    b.tri = c->m_tri;
    b.tri_part = c->m_part;
    b.pos = c->m_pos;
    b.vel = c->m_vel;
    return b;
}

// Enqueue on stream B the geometry of step s (dynamic sources).  fresh: the
// mask and wall buffers of step s hold no entries from earlier steps (they
// were cleared), so no stale tiles need clearing.
// Replaces the part mask in slot `slot` (step s) by replayed mask k and
// rebuilds its solid-proximity flags; marks every tile as written so that
// later incremental steps clear the whole slot.
hipError_t replay_mask(bf_context* c, int64_t s, uint32_t k, int slot) {
    int periodic[3];
    for (int a = 0; a < 3; ++a) periodic[a] = periodic_axis(c->cfg, a) ? 1 : 0;
    hipError_t e = hipMemcpyAsync(c->mask[slot], c->frozen + (uint64_t)k * c->g.n, c->g.n, hipMemcpyDeviceToDevice,
                                  c->sb);
    if (e == hipSuccess) e = launch_mark_tiles(c->mask[slot], c->dyn[slot], c->g, periodic, true, c->sb);
    if (e == hipSuccess) e = launch_fill_u8(c->tflags[mod4(s)], 1, c->g.tiles, c->sb);
    return e;
}

RayParams ray_params(const bf_context* c) {
    RayParams r{};
    r.g = c->g;
    for (int a = 0; a < 3; ++a) {
        r.origin[a] = c->cfg.origin[a];
        r.bc[a][0] = (int)c->cfg.boundary[a][0];
        r.bc[a][1] = (int)c->cfg.boundary[a][1];
    }
    r.dx = c->cfg.cell_size;
    r.half_thickness = c->m_half;
    r.triangle_count = c->m_triangles;
    // one triangle per single-wavefront block (the pass waits on loads of
    // the link directions: many wavefronts per CU hide that); a mesh of fewer
    // triangles than two wavefronts per CU splits each triangle's cells among
    // several
    const uint32_t waves = 2u * c->compute_units;
    r.triangle_blocks = std::max(1u, std::min(c->m_triangles, c->ray_blocks_per_cu * c->compute_units));
    r.split = std::min(64u, std::max(1u, waves / std::max(1u, c->m_triangles)));
    r.tile_pass_blocks = c->ray_tile_blocks_per_cu * c->compute_units;
    return r;
}

bool use_static_keys(const bf_context* c) {
    return c->static_keys_enabled && c->m_static_count > 0 && !c->move_mesh;
}

// The static cache for a plain ray pass (the adjoint's pass with winners
// walks every triangle).
void set_ray_static(const bf_context* c, RayBuffers& rb) {
    if (!c->ray_static_ready || c->ray_static_count == 0) return;
    rb.static_entries = c->ray_static;
    rb.static_count = c->ray_static_count;
    rb.dynamic_list = c->m_dyn_list;
    rb.dynamic_count = c->m_dyn_count;
}

// The cached keys of the static triangles within the stepping band: the
// static triangles rastered alone into the (empty) key grid, then collected
// into (cell, key) pairs, which empties the grid again.  Once per mesh
// upload; synchronizes stream B.
bf_status ensure_static_keys(bf_context* c) {
    if (c->static_ready) return BF_OK;
    hipStream_t st = c->sb;
    DeviceAlloc& A = c->static_alloc;
    A.release();
    c->static_cells = nullptr;
    c->static_keys = nullptr;
    hipError_t e = hipSuccess;
    uint8_t* tiles = (uint8_t*)A.take(e, c->g.tiles);
    unsigned long long* counter = (unsigned long long*)A.take(e, 8);
    if (e != hipSuccess) {
        A.release();
        return fail(c, BF_ERROR_ALLOCATION_FAILED, std::string("static key cache: ") + hipGetErrorString(e));
    }
    MeshParams m = mesh_params(c, time_of(c, c->step), false);
    set_raster_shape(c, m, c->m_static_count);
    MeshBuffers b = mesh_buffers(c);
    b.keys = c->keys;
    b.tiles_new = tiles;
    b.raster_list = c->m_static_list;
    e = launch_mesh_static_keys(b, m, counter, st);
    unsigned long long n = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&n, counter, 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(c, e, "static key raster");
    if (n > c->g.n) return fail(c, BF_ERROR_LAUNCH_FAILED, "static key count exceeds the grid");
    c->static_cells = (uint32_t*)A.take(e, 4 * std::max<unsigned long long>(n, 1));
    c->static_keys = (unsigned long long*)A.take(e, 8 * std::max<unsigned long long>(n, 1));
    if (e != hipSuccess) {
        A.release();
        return fail(c, BF_ERROR_ALLOCATION_FAILED, std::string("static key cache: ") + hipGetErrorString(e));
    }
    e = launch_collect_keys(c->keys, c->g.n, c->static_cells, c->static_keys, (uint32_t)n, counter, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(c, e, "static key collection");
    c->static_key_count = (uint32_t)n;
    c->static_ready = true;
    return BF_OK;
}

// The exact-root rule's static cache: entries of the static triangles'
// shells for every direction of every source cell near them, from the
// prepared vertices in m_pos (static vertices do not move).  Needs a
// scratch table of 76 B per cell while it builds; without the room, or
// without static triangles, the pass walks every triangle.
bf_status ensure_ray_static(bf_context* c) {
    if (c->ray_static_ready) return BF_OK;
    c->ray_static_ready = true;
    c->ray_static_alloc.release();
    c->ray_static = nullptr;
    c->ray_static_count = 0;
    if (!c->ray_static_enabled || !use_static_keys(c)) return BF_OK;
    const uint64_t scratch_bytes = 19ull * c->g.n * 4;
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    if (scratch_bytes + (64ull << 20) > (uint64_t)free_b) return BF_OK;
    hipStream_t st = c->sb;
    DeviceAlloc scratch;                                  // released before every return below
    hipError_t e = hipSuccess;
    uint32_t* table = (uint32_t*)scratch.take(e, scratch_bytes);
    uint32_t* counter = (uint32_t*)scratch.take(e, 256);
    if (e != hipSuccess) {                                // no room: walk every triangle
        scratch.release();
        return BF_OK;
    }
    const RayParams rp = ray_params(c);
    uint32_t n = 0;
    // 19 N exceeds 32 bits at 1024 x 512 x 512: fill in chunks
    for (uint64_t off = 0, total = 19ull * c->g.n; off < total && e == hipSuccess; off += 1ull << 30)
        e = launch_fill_u32(table + off, 0x40000000u, (uint32_t)std::min<uint64_t>(1ull << 30, total - off), st);
    if (e == hipSuccess)
        e = launch_ray_static(c->m_pos, c->m_tri, c->m_static_list, c->m_static_count, table, nullptr, 0, counter, rp,
                              st);
    if (e == hipSuccess) e = hipMemcpyAsync(&n, counter, 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && n > 0) {
        c->ray_static = (uint4*)c->ray_static_alloc.take(e, 16ull * n);
        if (e == hipSuccess)
            e = launch_ray_static(c->m_pos, c->m_tri, c->m_static_list, c->m_static_count, table, c->ray_static, n,
                                  counter, rp, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e == hipErrorOutOfMemory) {                   // no room for the list: walk every triangle
            (void)hipGetLastError();
            c->ray_static_alloc.release();
            c->ray_static = nullptr;
            scratch.release();
            return BF_OK;
        }
    }
    scratch.release();
    if (e != hipSuccess) {
        c->ray_static_alloc.release();
        c->ray_static = nullptr;
        return hip_fail(c, e, "exact-root static cache");
    }
    c->ray_static_count = n;
    return BF_OK;
}

bf_status enqueue_geometry(bf_context* c, int64_t s, bool exact, bool fresh) {
    hipError_t e = hipSuccess;
    if (c->source == BF_GEOMETRY_MESH) {
        MeshBuffers b = mesh_buffers(c);
        b.keys = c->keys;
        b.tiles_new = c->tflags[mod4(s)];
        b.tiles_mask_old = fresh ? c->no_tiles : c->tflags[mod4(s - 3)];   // mask[s%3] last written at s-3
        b.tiles_wall_old = fresh ? c->no_tiles : c->tflags[mod4(s - 2)];   // wall[s&1] last written at s-2
        b.list = c->geo_list;
        b.list_count = &c->counters[2];
        b.mask = c->mask[mod3(s)];
        b.wall = c->wall[par2(s)];
        b.dyn = c->dyn[mod3(s)];
        MeshParams m = mesh_params(c, time_of(c, s), exact);
        if (!exact && use_static_keys(c)) {
            if (bf_status r = ensure_static_keys(c)) return r;
            set_raster_shape(c, m, c->m_dyn_count);
            b.raster_list = c->m_dyn_list;
            b.seed_cells = c->static_cells;
            b.seed_keys = c->static_keys;
            b.seed_count = c->static_key_count;
        }
        e = launch_mesh_geometry(b, m, c->sb);
        if (e == hipSuccess && c->frozen_count && s >= c->frozen_first && s < c->frozen_first + c->frozen_count)
            e = replay_mask(c, s, (uint32_t)(s - c->frozen_first), mod3(s));
        if (e == hipSuccess && c->cfg.wall_rule == BF_WALL_INTERPOLATED_RAY) {
            if (bf_status r = ensure_ray_static(c)) return r;
            RayBuffers rb{};
            rb.pos = c->m_pos;
            rb.tri = c->m_tri;
            rb.mask = b.mask;
            rb.wall = b.wall;
            rb.dyn = b.dyn;
            rb.list = c->ray_list;
            rb.list_count = c->ray_count;
            rb.dirs = c->ray_dirs;
            rb.links = c->ray_links[par2(s)];
            set_ray_static(c, rb);
            e = launch_ray_links(rb, ray_params(c), c->sb);
        }
    } else if (c->source == BF_GEOMETRY_BIRD) {
        // fresh: built from cleared buffers with no previous geometry
        // (the reference's initialization build)
        BirdBuffers b{};
        b.body = c->body;
        b.keyframes = c->bird_keyframes;
        b.keyframe_count = c->bird_keyframe_count;
        b.frame = c->bird_frame + BIRD_FRAME_ROWS * par2(s);
        b.mask_prev = fresh ? nullptr : c->mask[mod3(s - 1)];
        b.dyn_prev = fresh ? nullptr : c->dyn[mod3(s - 1)];
        b.tiles_new = c->tflags[mod4(s)];
        b.tiles_mask_old = fresh ? c->no_tiles : c->tflags[mod4(s - 3)];
        b.tiles_wall_old = fresh ? c->no_tiles : c->tflags[mod4(s - 2)];
        b.list = c->geo_list;
        b.list_count = &c->counters[2];
        b.mask = c->mask[mod3(s)];
        b.wall = c->wall[par2(s)];
        b.dyn = c->dyn[mod3(s)];
        BirdParams bp{};
        bp.g = c->g;
        for (int a = 0; a < 3; ++a) {
            bp.origin[a] = c->cfg.origin[a];
            bp.periodic[a] = periodic_axis(c->cfg, a) ? 1 : 0;
        }
        bp.dx = c->cfg.cell_size;
        bp.time = time_of(c, s);
        bp.velocity_scale = c->cfg.velocity_to_lattice;
        bp.tile_pass_blocks = 4u * c->tile_pass_blocks;
        WingDynamics wd = c->wing_dyn;
        if (c->wing_dynamics) {
            b.wing_momentum = c->wing_momentum;
            b.wing_reaction = c->wing_reaction + 16 * par2(s);
            wd.dt = c->cfg.time_step;
            wd.initialize = fresh ? 1 : 0;
        }
        e = launch_bird_geometry(c->bird, b, bp, c->sb, c->wing_dynamics ? &wd : nullptr);
    } else if (c->source == BF_GEOMETRY_WING) {
        WingBuffers b{};
        b.frame = c->wing_frame + WING_FRAME_ROWS * par2(s);
        b.mask_prev = fresh ? nullptr : c->mask[mod3(s - 1)];
        b.dyn_prev = fresh ? nullptr : c->dyn[mod3(s - 1)];
        b.tiles_new = c->tflags[mod4(s)];
        b.tiles_mask_old = fresh ? c->no_tiles : c->tflags[mod4(s - 3)];
        b.tiles_wall_old = fresh ? c->no_tiles : c->tflags[mod4(s - 2)];
        b.list = c->geo_list;
        b.list_count = &c->counters[2];
        b.mask = c->mask[mod3(s)];
        b.wall = c->wall[par2(s)];
        b.dyn = c->dyn[mod3(s)];
        b.links = c->wing_links[par2(s)];
        BirdParams bp{};
        bp.g = c->g;
        for (int a = 0; a < 3; ++a) {
            bp.origin[a] = c->cfg.origin[a];
            bp.periodic[a] = periodic_axis(c->cfg, a) ? 1 : 0;
            bp.bc[a][0] = (int)c->cfg.boundary[a][0];
            bp.bc[a][1] = (int)c->cfg.boundary[a][1];
        }
        bp.dx = c->cfg.cell_size;
        bp.time = time_of(c, s);
        bp.velocity_scale = c->cfg.velocity_to_lattice;
        bp.tile_pass_blocks = 4u * c->tile_pass_blocks;
        e = launch_wing_geometry(c->wing, b, bp, c->sb);
    } else if (c->source == BF_GEOMETRY_SPHERE) {
        int periodic[3];
        for (int a = 0; a < 3; ++a) periodic[a] = periodic_axis(c->cfg, a) ? 1 : 0;
        e = launch_sphere_geometry(c->mask[mod3(s)], c->wall[par2(s)], c->dyn[mod3(s)], c->g, c->sphere_cr,
                                   c->sphere_gv, c->sphere_wv, time_of(c, s), periodic, c->sb);
    }
    if (e != hipSuccess) return hip_fail(c, e, "geometry launch");
    return BF_OK;
}

// List of the tiles that take the general path in step s (stream B).
bf_status enqueue_tile_list(bf_context* c, int64_t s) {
    const StepGeometry sg = geometry_of(c, s);
    hipError_t e = launch_fill_u32(&c->counters[par2(s)], 0u, 1u, c->sb);
    if (e == hipSuccess)
        e = launch_build_general_list(c->static_flags, sg.dyn_prev, sg.dyn_cur, c->g.tiles, c->general_list[par2(s)],
                                      &c->counters[par2(s)], c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "tile list launch");
    return BF_OK;
}

FluidBuffers fluid_buffers(const bf_context* c, int64_t s) {
    const StepGeometry sg = geometry_of(c, s);
    FluidBuffers b{};
    b.fin = c->f[c->cur];
    b.fout = c->f[1 - c->cur];
    b.solid_cur = sg.mask_cur;
    b.solid_prev = sg.mask_prev;
    b.wall = sg.wall;
    b.dyn_cur = sg.dyn_cur;
    b.dyn_prev = sg.dyn_prev;
    b.static_flags = c->static_flags;
    b.general_list = c->general_list[par2(s)];
    b.general_count = &c->counters[par2(s)];
    b.partials = c->partials + 8ull * c->g.tiles * (uint64_t)par2(s);
    b.block_out = c->block_out;
    b.last_load = c->last_load;
    b.history = c->history;
    b.density = c->density;
    b.velocity = c->velocity;
    b.activation = &c->counters[4 + 2 * par2(s)];
    b.body = c->body;
    return b;
}

bf_status check_host_array(bf_context* c, const void* p, const char* name) {
    if (!p) return fail(c, BF_ERROR_INVALID_ARGUMENT, std::string(name) + " is NULL");
    return BF_OK;
}

}  // namespace

__global__ void copy_probe(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

extern "C" {

uint32_t bf_abi_version(void) { return BF_ABI_VERSION; }

bf_status bf_copy_bandwidth(uint64_t bytes, uint32_t iterations, double* gbs) {
    if (!gbs || bytes < 16 || iterations == 0) return BF_ERROR_INVALID_ARGUMENT;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return BF_ERROR_DEVICE_UNAVAILABLE;
    const size_t n = (size_t)(bytes / 16);
    if ((n + 255) / 256 > 0x7fffffffull) return BF_ERROR_INVALID_ARGUMENT;
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 16) != hipSuccess) return BF_ERROR_ALLOCATION_FAILED;
    if (hipMalloc(&b, n * 16) != hipSuccess) {
        (void)hipFree(a);
        return BF_ERROR_ALLOCATION_FAILED;
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipMemset(a, 0, n * 16);
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(copy_probe, dim3(blocks), dim3(256), 0, 0, (const float4*)a, (float4*)b, n);
    (void)hipEventRecord(e0, 0);
    for (uint32_t i = 0; i < iterations; ++i)
        hipLaunchKernelGGL(copy_probe, dim3(blocks), dim3(256), 0, 0, (const float4*)a, (float4*)b, n);
    (void)hipEventRecord(e1, 0);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return BF_ERROR_LAUNCH_FAILED;
    *gbs = 2.0 * (double)n * 16.0 * iterations / (ms * 1e-3) / 1e9;
    return BF_OK;
}

const char* bf_status_string(bf_status s) {
    switch (s) {
        case BF_OK: return "ok";
        case BF_ERROR_DEVICE_UNAVAILABLE: return "no gfx950 device is available";
        case BF_ERROR_INVALID_CONFIGURATION: return "invalid configuration";
        case BF_ERROR_ALLOCATION_FAILED: return "device allocation failed";
        case BF_ERROR_BUFFER_EXCEEDS_DEVICE_LIMIT: return "a buffer exceeds the device limit";
        case BF_ERROR_WORKING_SET_EXCEEDS_MEMORY: return "the working set exceeds device memory";
        case BF_ERROR_INVALID_ADVANCE_REQUEST: return "invalid advance request";
        case BF_ERROR_INVALID_ARGUMENT: return "invalid argument";
        case BF_ERROR_LAUNCH_FAILED: return "a device launch failed";
        case BF_ERROR_RUNTIME_SAFETY_VIOLATION: return "runtime safety violation";
        case BF_ERROR_STATE_INVALIDATED: return "the simulation state was invalidated by an earlier failure";
        case BF_ERROR_NOT_SUPPORTED: return "not supported";
    }
    return "unknown status";
}

const char* bf_last_error(const bf_context* c) { return c ? c->error.c_str() : "null context"; }

bf_status bf_device_info(char* name, size_t cap, uint64_t* total) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return BF_ERROR_DEVICE_UNAVAILABLE;
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return BF_ERROR_DEVICE_UNAVAILABLE;
    if (name && cap) std::snprintf(name, cap, "%s (%s)", prop.name, prop.gcnArchName);
    if (total) *total = (uint64_t)prop.totalGlobalMem;
    return BF_OK;
}

uint64_t bf_allocated_bytes(const bf_context* c) {
    if (!c) return 0;
    return c->alloc.bytes + c->mesh_alloc.bytes + c->static_alloc.bytes + c->ledger_alloc.bytes + c->wing_alloc.bytes +
           c->ray_alloc.bytes + c->ray_static_alloc.bytes + c->keyframe_alloc.bytes + c->adj.alloc.bytes +
           c->adj.tangent_alloc.bytes + c->frozen_alloc.bytes;
}

bf_status bf_create(const bf_config* cfg, bf_context** out) {
    if (!cfg || !out) return BF_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    const uint64_t n64 = (uint64_t)cfg->nx * cfg->ny * cfg->nz;
    if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1 || n64 >= (1ull << 32)) return BF_ERROR_INVALID_CONFIGURATION;
    if (cfg->storage > BF_STORAGE_BF16_OFFSET || cfg->collision > BF_COLLISION_RR3 ||
        cfg->wall_rule > BF_WALL_INTERPOLATED_RAY || cfg->force_mode > 6 || cfg->load_component > 2 ||
        cfg->arithmetic > BF_ARITHMETIC_FUSED)
        return BF_ERROR_INVALID_CONFIGURATION;
    if (cfg->arithmetic == BF_ARITHMETIC_STRICT && cfg->storage != BF_STORAGE_FP32)
        return BF_ERROR_INVALID_CONFIGURATION;
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < 2; ++s)
            if (cfg->boundary[a][s] > BF_BC_OUTLET) return BF_ERROR_INVALID_CONFIGURATION;
    for (int a = 0; a < 3; ++a)
        if ((cfg->boundary[a][0] == BF_BC_PERIODIC) != (cfg->boundary[a][1] == BF_BC_PERIODIC))
            return BF_ERROR_INVALID_CONFIGURATION;
    if (!(cfg->omega_plus > 0.0f && cfg->omega_plus < 2.0f && cfg->omega_minus > 0.0f && cfg->omega_minus < 2.0f) ||
        !(cfg->cell_size > 0.0f) || !(cfg->far_density > 0.0f) || !(cfg->sponge_strength >= 0.0f) ||
        !(cfg->sponge_strength <= 1.0f))
        return BF_ERROR_INVALID_CONFIGURATION;

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return BF_ERROR_DEVICE_UNAVAILABLE;
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return BF_ERROR_DEVICE_UNAVAILABLE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return BF_ERROR_DEVICE_UNAVAILABLE;

    bf_context* c = new (std::nothrow) bf_context();
    if (!c) return BF_ERROR_ALLOCATION_FAILED;
    c->cfg = *cfg;
    c->device = dev;
    Grid& g = c->g;
    g.nx = (int)cfg->nx;
    g.ny = (int)cfg->ny;
    g.nz = (int)cfg->nz;
    g.tiles_x = (g.nx + TILE_X - 1) / TILE_X;
    g.tiles_y = (g.ny + TILE_Y - 1) / TILE_Y;
    g.n = (uint32_t)n64;
    g.tiles = (uint32_t)((uint64_t)g.tiles_x * g.tiles_y * g.nz);
    c->pop_elem = cfg->storage == BF_STORAGE_FP32 ? 4 : 2;
    c->general_blocks = std::max(1, prop.multiProcessorCount * 4);
    c->compute_units = (uint32_t)std::max(1, prop.multiProcessorCount);
    // one tile-pass workgroup per CU: the fastest of a sweep (CU x 1/2/4/8)
    // on the 512x256x256 bird, where the passes overlap the fluid kernel;
    // the raster launches every triangle batch at once
    c->tile_pass_blocks = (uint32_t)std::max(1, prop.multiProcessorCount);
    // tuning overrides (workgroup counts)
    if (const char* v = std::getenv("BIRDFLOW_RASTER_BLOCKS")) c->raster_blocks = (uint32_t)std::max(1, std::atoi(v));
    // Scheduling of the next step's geometry against the fluid kernels (sweeps
    // on the 512x256x256 bird, 40k triangles):
    //  * FP32: the fast kernel runs at issue priority 2 over the raster, which
    //    starts after the general kernel and still finishes within the fast
    //    kernel (1.03 -> 1.01 ms per step);
    //  * FP16 / BF16: the fast kernel is too short for a deprioritized raster;
    //    the raster starts at once, beside the general kernel (priority 3),
    //    in batches of 16 triangles (0.664 -> 0.640 ms per step).
    const bool fp32 = cfg->storage == BF_STORAGE_FP32;
    c->fast_priority = fp32 ? 2 : 0;
    c->general_priority = 3;
    c->geometry_after_general = fp32;
    c->raster_batch_cap = fp32 ? 32u : 16u;
    if (const char* v = std::getenv("BIRDFLOW_FAST_PRIORITY")) c->fast_priority = std::min(3, std::max(0, std::atoi(v)));
    // FP16 / BF16: fast tiles two cells per lane (0 selects one cell per lane)
    if (const char* v = std::getenv("BIRDFLOW_PAIR_CELLS")) c->pair_cells = std::atoi(v) != 0;
    // meshes: cached keys of static triangles (0 rasters every triangle every step)
    if (const char* v = std::getenv("BIRDFLOW_STATIC_KEYS")) c->static_keys_enabled = std::atoi(v) != 0;
    if (const char* v = std::getenv("BIRDFLOW_RAY_STATIC")) c->ray_static_enabled = std::atoi(v) != 0;
    if (const char* v = std::getenv("BIRDFLOW_GENERAL_PRIORITY"))
        c->general_priority = std::min(3, std::max(0, std::atoi(v)));
    if (const char* v = std::getenv("BIRDFLOW_GEOMETRY_AFTER_GENERAL")) c->geometry_after_general = std::atoi(v) != 0;
    if (const char* v = std::getenv("BIRDFLOW_RASTER_BATCH"))
        c->raster_batch = (uint32_t)std::min(64, std::max(1, std::atoi(v)));
    if (const char* v = std::getenv("BIRDFLOW_RAY_BLOCKS_PER_CU"))
        c->ray_blocks_per_cu = (uint32_t)std::min(256, std::max(1, std::atoi(v)));
    if (const char* v = std::getenv("BIRDFLOW_RAY_TILE_BLOCKS_PER_CU"))
        c->ray_tile_blocks_per_cu = (uint32_t)std::min(64, std::max(1, std::atoi(v)));
    if (const char* v = std::getenv("BIRDFLOW_TILE_PASS_BLOCKS"))
        c->tile_pass_blocks = (uint32_t)std::max(1, std::atoi(v));

    // Working-set plan (MetalBackend.validateAllocationPlan, :77-106).
    const uint64_t n = g.n;
    const uint64_t pop_bytes = 19ull * n * c->pop_elem;
    const uint64_t plan = 2 * pop_bytes + 3 * n + 32 * n + 3ull * g.tiles + 4ull * g.tiles + 128ull * g.tiles +
                          64ull * HISTORY_CAPACITY + 128ull * HISTORY_CAPACITY + 4 * n + 16 * n;
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    if (plan > (uint64_t)free_b) {
        delete c;
        return BF_ERROR_WORKING_SET_EXCEEDS_MEMORY;
    }
    // The no-allocation ledger's device scratch: its address on this device
    // (looked up once per device, with the device current).
    if (ledger_scratch_setup(dev) != hipSuccess) {
        delete c;
        return BF_ERROR_DEVICE_UNAVAILABLE;
    }
    // Both streams at the default priority.
    if (hipStreamCreateWithFlags(&c->sa, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->sb, hipStreamNonBlocking) != hipSuccess) {
        bf_destroy(c);
        return BF_ERROR_DEVICE_UNAVAILABLE;
    }
    for (int i = 0; i < 2; ++i) {
        if (hipEventCreateWithFlags(&c->ev_fast[i], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_ready[i], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_general[i], hipEventDisableTiming) != hipSuccess) {
            bf_destroy(c);
            return BF_ERROR_DEVICE_UNAVAILABLE;
        }
    }
    hipError_t e = hipSuccess;
    DeviceAlloc& A = c->alloc;
    c->f[0] = A.take(e, pop_bytes);
    c->f[1] = A.take(e, pop_bytes);
    for (int i = 0; i < 3; ++i) c->mask[i] = (uint8_t*)A.take(e, n);
    for (int i = 0; i < 2; ++i) c->wall[i] = (float4*)A.take(e, 16 * n);
    for (int i = 0; i < 3; ++i) c->dyn[i] = (uint8_t*)A.take(e, g.tiles);
    c->static_flags = (uint8_t*)A.take(e, g.tiles);
    for (int i = 0; i < 4; ++i) c->tflags[i] = (uint8_t*)A.take(e, g.tiles);
    c->no_tiles = (uint8_t*)A.take(e, g.tiles);
    for (int i = 0; i < 2; ++i) c->general_list[i] = (uint32_t*)A.take(e, 4ull * g.tiles);
    c->geo_list = (uint32_t*)A.take(e, 4ull * g.tiles);
    c->bird_frame = (float4*)A.take(e, 2ull * BIRD_FRAME_ROWS * sizeof(float4));
    c->wing_momentum = (float*)A.take(e, 16 * sizeof(float));
    c->health = (unsigned long long*)A.take(e, 2 * sizeof(unsigned long long));
    c->popmin = (PopulationMinimum*)A.take(e, (POPULATION_MINIMUM_BLOCKS + 1) * sizeof(PopulationMinimum));
    c->wing_reaction = (float*)A.take(e, 32 * sizeof(float));
    c->counters = (uint32_t*)A.take(e, 64);
    c->partials = (double*)A.take(e, 128ull * g.tiles);   // two steps' per-tile load partials
    c->block_out = (double*)A.take(e, 64ull * ((g.tiles + 1023) / 1024 + 1));
    c->last_load = (double*)A.take(e, 64);
    c->history = (double*)A.take(e, 64ull * HISTORY_CAPACITY);
    c->census = (unsigned long long*)A.take(e, 8 * 24);
    c->density = (float*)A.take(e, 4 * n);
    c->velocity = (float4*)A.take(e, 16 * n);
    c->body = (double*)A.take(e, 16 * sizeof(double));
    c->body_history = (double*)A.take(e, 16 * sizeof(double) * HISTORY_CAPACITY);
    c->safety_metrics = (double*)A.take(e, 4 * sizeof(double));
    c->safety_event = (uint32_t*)A.take(e, 4 * sizeof(uint32_t));
    if (e != hipSuccess) {
        bf_destroy(c);
        return BF_ERROR_ALLOCATION_FAILED;
    }
    std::vector<uint8_t> sf = static_tile_flags(*cfg, g);
    hipStream_t s = c->sb;
    e = hipMemcpyAsync(c->static_flags, sf.data(), sf.size(), hipMemcpyHostToDevice, s);
    for (int i = 0; i < 3 && e == hipSuccess; ++i) e = launch_fill_u8(c->mask[i], 0, g.n, s);
    for (int i = 0; i < 2 && e == hipSuccess; ++i) e = launch_fill_wall(c->wall[i], g.n, s);
    for (int i = 0; i < 3 && e == hipSuccess; ++i) e = launch_fill_u8(c->dyn[i], 0, g.tiles, s);
    for (int i = 0; i < 4 && e == hipSuccess; ++i) e = launch_fill_u8(c->tflags[i], 0, g.tiles, s);
    if (e == hipSuccess) e = launch_fill_u8(c->no_tiles, 0, g.tiles, s);
    if (e == hipSuccess) e = launch_fill_u32(c->counters, 0u, 16u, s);
    if (e == hipSuccess) e = launch_fill_u8((uint8_t*)c->last_load, 0, 64, s);
    if (e == hipSuccess) e = launch_fill_u8((uint8_t*)c->partials, 0, (uint32_t)(128ull * g.tiles), s);
    if (e == hipSuccess) e = launch_fill_u8((uint8_t*)c->wing_momentum, 0, 16 * sizeof(float), s);
    if (e == hipSuccess) e = launch_fill_u8((uint8_t*)c->health, 0, 8, s);
    if (e == hipSuccess) e = launch_fill_u8((uint8_t*)c->health + 8, 0xFF, 8, s);
    if (e == hipSuccess) e = launch_fill_u8((uint8_t*)c->wing_reaction, 0, 32 * sizeof(float), s);
    // body at rest at the torque reference point, identity orientation
    const double body0[16] = {cfg->body_position[0], cfg->body_position[1], cfg->body_position[2], 0.0, 0.0, 0.0,
                              1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const double safety0[4] = {0.0, DBL_MAX, 0.0, 0.0};
    const uint32_t event0[4] = {0u, 0u, 0u, 0u};
    if (e == hipSuccess) e = hipMemcpyAsync(c->body, body0, sizeof body0, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(c->safety_metrics, safety0, sizeof safety0, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(c->safety_event, event0, sizeof event0, hipMemcpyHostToDevice, s);
    // populations defined from the start: the far-field equilibrium of the
    // empty geometry (initializePopulations before any body is placed)
    if (e == hipSuccess)
        e = launch_init_equilibrium((int)cfg->storage, c->f[c->cur], c->mask[0], c->wall[0], cfg->far_velocity, g, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        bf_destroy(c);
        return BF_ERROR_LAUNCH_FAILED;
    }
    *out = c;
    return BF_OK;
}

void bf_destroy(bf_context* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->sa) (void)hipStreamSynchronize(c->sa);
    if (c->sb) (void)hipStreamSynchronize(c->sb);
    if (c->adj.sc) (void)hipStreamSynchronize(c->adj.sc);
    for (hipEvent_t ev : c->events) (void)hipEventDestroy(ev);
    for (int i = 0; i < 2; ++i) {
        if (c->ev_fast[i]) (void)hipEventDestroy(c->ev_fast[i]);
        if (c->ev_ready[i]) (void)hipEventDestroy(c->ev_ready[i]);
        if (c->ev_general[i]) (void)hipEventDestroy(c->ev_general[i]);
        if (c->adj.ev_geo[i]) (void)hipEventDestroy(c->adj.ev_geo[i]);
        if (c->adj.ev_done[i]) (void)hipEventDestroy(c->adj.ev_done[i]);
        if (c->adj.ev_tan[i]) (void)hipEventDestroy(c->adj.ev_tan[i]);
        if (c->adj.ev_used[i]) (void)hipEventDestroy(c->adj.ev_used[i]);
    }
    c->mesh_alloc.release();
    c->static_alloc.release();
    c->ledger_alloc.release();
    c->keyframe_alloc.release();
    c->wing_alloc.release();
    c->ray_alloc.release();
    c->ray_static_alloc.release();
    c->adj.alloc.release();
    c->adj.tangent_alloc.release();
    c->frozen_alloc.release();
    c->alloc.release();
    if (c->sa) (void)hipStreamDestroy(c->sa);
    if (c->sb) (void)hipStreamDestroy(c->sb);
    if (c->adj.sc) (void)hipStreamDestroy(c->adj.sc);
    delete c;
}

uint64_t bf_step_index(const bf_context* c) { return c ? (uint64_t)c->step : 0; }

static bf_status rebuild_mesh_geometry(bf_context* c);
static bf_status rebuild_sphere_geometry(bf_context* c);
static bf_status rebuild_bird_geometry(bf_context* c);
static bf_status rebuild_wing_geometry(bf_context* c);

bf_status bf_read_restart_state(bf_context* c, uint64_t* step, float* mesh_time_origin) {
    BF_ENTER(c);
    if (!step || !mesh_time_origin) return fail(c, BF_ERROR_INVALID_ARGUMENT, "output pointers must not be NULL");
    *step = (uint64_t)c->step;
    *mesh_time_origin = c->time_start;
    return BF_OK;
}

bf_status bf_set_restart_state(bf_context* c, uint64_t step, float mesh_time_origin, const double last_load[8],
                               const uint8_t* solid_mask) {
    BF_ENTER(c);
    ++c->f_writes;
    if (step > (uint64_t)(INT64_MAX / 4) || !std::isfinite(mesh_time_origin))
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "restart step or mesh time origin out of range");
    if (last_load) {
        hipError_t e = hipMemcpy(c->last_load, last_load, 64, hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(c, e, "restart load");
    }
    c->step = (int64_t)step;
    c->time_start = mesh_time_origin;
    c->history_count = 0;
    c->body_history_count = 0;
    c->geometry_ready = c->step;
    // the clock moved: the current step's geometry is built again
    bf_status r = BF_OK;
    if (c->source == BF_GEOMETRY_MESH) r = rebuild_mesh_geometry(c);
    if (c->source == BF_GEOMETRY_SPHERE) r = rebuild_sphere_geometry(c);
    if (c->source == BF_GEOMETRY_BIRD) r = rebuild_bird_geometry(c);
    if (c->source == BF_GEOMETRY_WING) r = rebuild_wing_geometry(c);
    if (r != BF_OK || !solid_mask) return r;
    // The saved mask is the previous-step mask of the next step.  It equals
    // the rebuilt one unless the geometry followed a free-flying body whose
    // pose moved on after it was built; then the saved mask replaces it, its
    // tiles are marked, and the whole slot is cleared when it is reused.
    const size_t n = c->g.n;
    std::vector<uint8_t> cur(n);
    hipError_t e = hipMemcpy(cur.data(), geometry_of(c, c->step).mask_cur, n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(c, e, "restart mask");
    if (std::memcmp(cur.data(), solid_mask, n) == 0) return BF_OK;
    if (c->source == BF_GEOMETRY_STATIC)
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "restart solid mask differs from the static geometry");
    int periodic[3];
    for (int a = 0; a < 3; ++a) periodic[a] = periodic_axis(c->cfg, a) ? 1 : 0;
    const int k = mod3(c->step);
    e = hipMemcpy(c->mask[k], solid_mask, n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_mark_tiles(c->mask[k], c->dyn[k], c->g, periodic, false, c->sb);
    if (e == hipSuccess) e = launch_fill_u8(c->tflags[mod4(c->step)], 1, c->g.tiles, c->sb);
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "restart mask");
    return BF_OK;
}

// A new population state clears the non-finite population record.
static hipError_t reset_health(bf_context* c) {
    c->health_reported = false;
    hipError_t e = launch_fill_u8((uint8_t*)c->health, 0, 8, c->sb);
    if (e == hipSuccess) e = launch_fill_u8((uint8_t*)c->health + 8, 0xFF, 8, c->sb);
    return e;
}

bf_status bf_init_equilibrium(bf_context* c) {
    BF_ENTER(c);
    ++c->f_writes;
    if (hipError_t e0 = reset_health(c)) return hip_fail(c, e0, "reset health");
    const StepGeometry sg = geometry_of(c, c->step);
    hipError_t e = launch_init_equilibrium((int)c->cfg.storage, c->f[c->cur], sg.mask_cur, sg.wall,
                                           c->cfg.far_velocity, c->g, c->sb);
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "initialize populations");
    return BF_OK;
}

bf_status bf_init_shear_wave(bf_context* c, float amplitude) {
    BF_ENTER(c);
    ++c->f_writes;
    if (hipError_t e0 = reset_health(c)) return hip_fail(c, e0, "reset health");
    hipError_t e = launch_init_shear_wave((int)c->cfg.storage, c->f[c->cur], amplitude, c->g, c->sb);
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "initialize shear wave");
    return BF_OK;
}

// Host-layout FP32 staging for population transfers (every storage: the
// device layout is not the host's direction-major one).
static bf_status ensure_staging(bf_context* c) {
    if (c->staging) return BF_OK;
    hipError_t e = hipSuccess;
    c->staging = (float*)c->alloc.take(e, 19ull * c->g.n * 4);
    if (e != hipSuccess || !c->staging) {
        c->staging = nullptr;
        return fail(c, BF_ERROR_ALLOCATION_FAILED, "population staging buffer");
    }
    return BF_OK;
}

bf_status bf_write_populations(bf_context* c, const float* host, uint64_t count) {
    BF_ENTER(c);
    ++c->f_writes;
    if (hipError_t e0 = reset_health(c)) return hip_fail(c, e0, "reset health");
    if (check_host_array(c, host, "populations")) return BF_ERROR_INVALID_ARGUMENT;
    if (count != 19ull * c->g.n) return fail(c, BF_ERROR_INVALID_ARGUMENT, "population count must be 19 * N");
    // through the staging buffer: the device layout differs from the host's
    // direction-major one for every storage
    if (bf_status st = ensure_staging(c)) return st;
    hipError_t e = hipMemcpyAsync(c->staging, host, count * 4, hipMemcpyHostToDevice, c->sb);
    if (e == hipSuccess)
        e = launch_convert_populations((int)c->cfg.storage, true, c->staging, c->f[c->cur], count, c->g, c->sb);
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "write populations");
    return BF_OK;
}

bf_status bf_read_populations(bf_context* c, float* host, uint64_t count) {
    BF_ENTER(c);
    if (check_host_array(c, host, "populations")) return BF_ERROR_INVALID_ARGUMENT;
    if (count != 19ull * c->g.n) return fail(c, BF_ERROR_INVALID_ARGUMENT, "population count must be 19 * N");
    if (bf_status st = ensure_staging(c)) return st;
    hipError_t e = launch_convert_populations((int)c->cfg.storage, false, c->f[c->cur], c->staging, count, c->g, c->sb);
    if (e == hipSuccess) e = hipMemcpyAsync(host, c->staging, count * 4, hipMemcpyDeviceToHost, c->sb);
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "read populations");
    return BF_OK;
}

static uint64_t stored_bytes(const bf_context* c) {
    return 19ull * c->g.n * (c->cfg.storage == BF_STORAGE_FP32 ? 4u : 2u);
}

bf_status bf_write_population_storage(bf_context* c, const void* host, uint64_t bytes) {
    BF_ENTER(c);
    ++c->f_writes;
    if (hipError_t e0 = reset_health(c)) return hip_fail(c, e0, "reset health");
    if (check_host_array(c, host, "population storage")) return BF_ERROR_INVALID_ARGUMENT;
    if (bytes != stored_bytes(c))
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "population storage size must be 19 * N * element size");
    if (bf_status st = ensure_staging(c)) return st;
    hipError_t e = hipMemcpyAsync(c->staging, host, bytes, hipMemcpyHostToDevice, c->sb);
    if (e == hipSuccess)
        e = launch_permute_populations((int)c->cfg.storage, true, c->staging, c->f[c->cur], 19ull * c->g.n, c->g,
                                       c->sb);
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "write population storage");
    return BF_OK;
}

bf_status bf_read_population_storage(bf_context* c, void* host, uint64_t bytes) {
    BF_ENTER(c);
    if (check_host_array(c, host, "population storage")) return BF_ERROR_INVALID_ARGUMENT;
    if (bytes != stored_bytes(c))
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "population storage size must be 19 * N * element size");
    if (bf_status st = ensure_staging(c)) return st;
    hipError_t e = launch_permute_populations((int)c->cfg.storage, false, c->f[c->cur], c->staging, 19ull * c->g.n,
                                              c->g, c->sb);
    if (e == hipSuccess) e = hipMemcpyAsync(host, c->staging, bytes, hipMemcpyDeviceToHost, c->sb);
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "read population storage");
    return BF_OK;
}

bf_status bf_read_fields(bf_context* c, float* density, float* velocity3) {
    BF_ENTER(c);
    if (check_host_array(c, density, "density") || check_host_array(c, velocity3, "velocity"))
        return BF_ERROR_INVALID_ARGUMENT;
    float* dv = (float*)c->velocity;   // 16 B per cell; holds 3 floats per cell here
    hipError_t e = launch_extract_fields((int)c->cfg.storage, c->f[c->cur], c->density, dv, c->g, c->sb);
    if (e == hipSuccess) e = hipMemcpyAsync(density, c->density, 4ull * c->g.n, hipMemcpyDeviceToHost, c->sb);
    if (e == hipSuccess) e = hipMemcpyAsync(velocity3, dv, 12ull * c->g.n, hipMemcpyDeviceToHost, c->sb);
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "read fields");
    return BF_OK;
}

bf_status bf_set_geometry(bf_context* c, const uint8_t* mask_prev, const uint8_t* mask_cur, const float* wall4) {
    BF_ENTER(c);
    if (check_host_array(c, mask_cur, "mask_cur") || check_host_array(c, wall4, "wall"))
        return BF_ERROR_INVALID_ARGUMENT;
    const size_t n = c->g.n;
    c->source = BF_GEOMETRY_STATIC;
    c->static_prev_differs = mask_prev && std::memcmp(mask_prev, mask_cur, n) != 0;
    hipStream_t s = c->sb;
    hipError_t e = hipMemcpyAsync(c->mask[0], mask_cur, n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && c->static_prev_differs) e = hipMemcpyAsync(c->mask[1], mask_prev, n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(c->wall[0], wall4, 16 * n, hipMemcpyHostToDevice, s);
    int periodic[3];
    for (int a = 0; a < 3; ++a) periodic[a] = periodic_axis(c->cfg, a) ? 1 : 0;
    if (e == hipSuccess) e = launch_mark_tiles(c->mask[0], c->dyn[0], c->g, periodic, true, s);
    if (e == hipSuccess && c->static_prev_differs) e = launch_mark_tiles(c->mask[1], c->dyn[1], c->g, periodic, true, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(c, e, "set geometry");
    c->geometry_ready = c->step;
    return BF_OK;
}

bf_status bf_read_geometry(bf_context* c, uint8_t* mask, float* wall4) {
    BF_ENTER(c);
    const StepGeometry sg = geometry_of(c, c->step);
    const size_t n = c->g.n;
    hipError_t e = hipSuccess;
    if (mask) e = hipMemcpyAsync(mask, sg.mask_cur, n, hipMemcpyDeviceToHost, c->sb);
    if (e == hipSuccess && wall4) e = hipMemcpyAsync(wall4, sg.wall, 16 * n, hipMemcpyDeviceToHost, c->sb);
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "read geometry");
    return BF_OK;
}

// Geometry of the current step from cleared buffers (masks, walls, flags,
// keys), with the reference's full key band; the previous step's mask starts
// equal to it (its touched-tile flags are copied so that the step that reuses
// that buffer clears it).
static bf_status rebuild_mesh_geometry(bf_context* c) {
    hipStream_t s = c->sb;
    hipError_t e = launch_fill_keys(c->keys, c->g.n, s);
    for (int i = 0; i < 3 && e == hipSuccess; ++i) e = launch_fill_u8(c->mask[i], 0, c->g.n, s);
    for (int i = 0; i < 2 && e == hipSuccess; ++i) e = launch_fill_wall(c->wall[i], c->g.n, s);
    for (int i = 0; i < 3 && e == hipSuccess; ++i) e = launch_fill_u8(c->dyn[i], 0, c->g.tiles, s);
    for (int i = 0; i < 4 && e == hipSuccess; ++i) e = launch_fill_u8(c->tflags[i], 0, c->g.tiles, s);
    if (e != hipSuccess) return hip_fail(c, e, "mesh geometry clear");
    const int64_t st = c->step;
    if (bf_status r = enqueue_geometry(c, st, true, true)) return r;
    const int k = mod3(st), kp = mod3(st - 1);
    e = hipMemcpyAsync(c->mask[kp], c->mask[k], c->g.n, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(c->dyn[kp], c->dyn[k], c->g.tiles, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->tflags[mod4(st - 1)], c->tflags[mod4(st)], c->g.tiles, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(c, e, "initial mesh geometry");
    c->geometry_ready = st;
    return BF_OK;
}

// Host arrays of one mesh upload: recorded frames, or an articulated mesh.
struct MeshSource {
    const float* points = nullptr;        // [frames][V][3]
    const float* frame_times = nullptr;
    uint32_t frame_count = 0;
    const float* rest = nullptr;          // [V][3]
    const uint8_t* vertex_joint = nullptr;
    const float* span_fraction = nullptr;
    const bf_kinematics* kinematics = nullptr;
};

// Triangle checks shared by both upload forms; `pose` gives vertex
// positions [V][3] of one pose.
static bf_status check_triangles(bf_context* c, const float* pose, uint32_t vertex_count, const uint32_t* triangles,
                                 const uint8_t* triangle_parts, uint32_t triangle_count, const char* which) {
    for (uint32_t t = 0; t < triangle_count; ++t) {
        const uint32_t* v = triangles + 3ull * t;
        if (v[0] >= vertex_count || v[1] >= vertex_count || v[2] >= vertex_count)
            return fail(c, BF_ERROR_INVALID_ARGUMENT, "triangle index out of range");
        if (triangle_parts[t] == 0) return fail(c, BF_ERROR_INVALID_ARGUMENT, "triangle part ids must be 1..255");
        // A zero-area triangle makes the closest-point branches divide 0 by 0.
        const float* a = pose + 3ull * v[0];
        const float* b = pose + 3ull * v[1];
        const float* d = pose + 3ull * v[2];
        const double ux = (double)b[0] - a[0], uy = (double)b[1] - a[1], uz = (double)b[2] - a[2];
        const double wx = (double)d[0] - a[0], wy = (double)d[1] - a[1], wz = (double)d[2] - a[2];
        const double nx = uy * wz - uz * wy, ny = uz * wx - ux * wz, nz = ux * wy - uy * wx;
        if (!(nx * nx + ny * ny + nz * nz > 0.0))
            return fail(c, BF_ERROR_INVALID_ARGUMENT, std::string("degenerate (zero-area) triangle in ") + which);
    }
    return BF_OK;
}

// Installs a validated mesh: buffers, static / moving triangle lists (a
// triangle is static when its three vertices are), the exact-root tables
// when that wall rule is configured, and the geometry at time_start.
static bf_status install_mesh(bf_context* c, const MeshSource& src, uint32_t vertex_count, const uint32_t* triangles,
                              const uint8_t* triangle_parts, uint32_t triangle_count, float half_thickness,
                              const float translation[3], float time_start,
                              const std::vector<uint8_t>& vertex_static) {
    if (c->cfg.wall_rule == BF_WALL_INTERPOLATED_RAY && !c->ray_links[0]) {
        // the exact-root tables of both step parities (every tile's 19 x 256
        // entries; the passes write the tiles next to solid cells), the tile
        // list and the per-cell link directions (zero between steps)
        hipError_t e = hipSuccess;
        const size_t entries = (size_t)c->g.tiles * 19 * TILE_CELLS;
        for (int k = 0; k < 2; ++k) c->ray_links[k] = (float*)c->ray_alloc.take(e, entries * sizeof(float));
        c->ray_list = (uint32_t*)c->ray_alloc.take(e, 4ull * c->g.tiles);
        c->ray_count = (uint32_t*)c->ray_alloc.take(e, 256);
        c->ray_dirs = (uint32_t*)c->ray_alloc.take(e, 4ull * c->g.n);
        if (e == hipSuccess) e = launch_fill_u32(c->ray_dirs, 0u, c->g.n, c->sb);
        if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
        if (e != hipSuccess) {
            c->ray_alloc.release();
            c->ray_links[0] = c->ray_links[1] = nullptr;
            c->ray_list = c->ray_count = c->ray_dirs = nullptr;
            return fail(c, BF_ERROR_ALLOCATION_FAILED, std::string("exact-root link tables: ") + hipGetErrorString(e));
        }
    }
    std::vector<uint32_t> dyn_list, static_list;
    for (uint32_t t = 0; t < triangle_count; ++t) {
        const uint32_t* v = triangles + 3ull * t;
        if (vertex_static[v[0]] && vertex_static[v[1]] && vertex_static[v[2]]) static_list.push_back(t);
        else dyn_list.push_back(t);
    }

    c->mesh_alloc.release();
    c->static_alloc.release();
    c->static_ready = false;
    c->static_cells = nullptr;
    c->static_keys = nullptr;
    c->static_key_count = 0;
    c->ray_static_alloc.release();
    c->ray_static = nullptr;
    c->ray_static_count = 0;
    c->ray_static_ready = false;
    c->adj.tangent_alloc.release();   // tangents and parameters belong to the previous mesh
    c->adj.K = 0;
    c->adj.dpoints = nullptr;
    c->adj.kinematic = false;
    c->m_points = c->m_times = c->m_rest = c->m_span = nullptr;
    c->m_joint = nullptr;
    c->m_kin = nullptr;
    DeviceAlloc& A = c->mesh_alloc;
    hipError_t e = hipSuccess;
    const bool articulated = src.kinematics != nullptr;
    const uint64_t pts = 3ull * src.frame_count * vertex_count;
    if (articulated) {
        c->m_rest = (float*)A.take(e, 12ull * vertex_count);
        c->m_joint = (uint8_t*)A.take(e, vertex_count);
        c->m_span = (float*)A.take(e, 4ull * vertex_count);
        c->m_kin = (Kinematics*)A.take(e, sizeof(Kinematics));
    } else {
        c->m_points = (float*)A.take(e, 4 * pts);
        c->m_times = (float*)A.take(e, 4ull * src.frame_count);
    }
    c->m_tri = (uint32_t*)A.take(e, 12ull * triangle_count);
    c->m_part = (uint8_t*)A.take(e, triangle_count);
    c->m_pos = (float4*)A.take(e, 16ull * vertex_count);
    c->m_vel = (float4*)A.take(e, 16ull * vertex_count);
    c->keys = (unsigned long long*)A.take(e, 8ull * c->g.n);
    c->m_dyn_list = (uint32_t*)A.take(e, 4ull * std::max<size_t>(dyn_list.size(), 1));
    c->m_static_list = (uint32_t*)A.take(e, 4ull * std::max<size_t>(static_list.size(), 1));
    if (e != hipSuccess) {
        A.release();
        c->source = BF_GEOMETRY_STATIC;
        return fail(c, BF_ERROR_ALLOCATION_FAILED, std::string("mesh buffers: ") + hipGetErrorString(e));
    }
    hipStream_t s = c->sb;
    if (articulated) {
        e = hipMemcpyAsync(c->m_rest, src.rest, 12ull * vertex_count, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(c->m_joint, src.vertex_joint, vertex_count, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(c->m_span, src.span_fraction, 4ull * vertex_count, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(c->m_kin, src.kinematics, sizeof(Kinematics), hipMemcpyHostToDevice, s);
    } else {
        e = hipMemcpyAsync(c->m_points, src.points, 4 * pts, hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(c->m_times, src.frame_times, 4ull * src.frame_count, hipMemcpyHostToDevice, s);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(c->m_tri, triangles, 12ull * triangle_count, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(c->m_part, triangle_parts, triangle_count, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && !dyn_list.empty())
        e = hipMemcpyAsync(c->m_dyn_list, dyn_list.data(), 4 * dyn_list.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && !static_list.empty())
        e = hipMemcpyAsync(c->m_static_list, static_list.data(), 4 * static_list.size(), hipMemcpyHostToDevice, s);
    // the host lists are read by the copies: finish them before they go
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(c, e, "mesh upload");
    c->m_dyn_count = (uint32_t)dyn_list.size();
    c->m_static_count = (uint32_t)static_list.size();
    if (articulated) c->host_times.clear();
    else c->host_times.assign(src.frame_times, src.frame_times + src.frame_count);
    c->m_articulated = articulated;
    c->m_joint_count = articulated ? src.kinematics->joint_count : 0;
    c->m_harmonics = articulated ? src.kinematics->harmonics : 0;
    c->m_frames = src.frame_count;
    c->m_vertices = vertex_count;
    c->m_triangles = triangle_count;
    c->m_half = half_thickness;
    for (int i = 0; i < 3; ++i) c->m_translation[i] = translation ? translation[i] : 0.0f;
    c->source = BF_GEOMETRY_MESH;
    c->static_prev_differs = false;
    c->time_start = time_start - (float)((double)c->step * c->cfg.time_step);
    return rebuild_mesh_geometry(c);
}

bf_status bf_mesh_upload(bf_context* c, const float* points, const float* frame_times, uint32_t frame_count,
                         uint32_t vertex_count, const uint32_t* triangles, const uint8_t* triangle_parts,
                         uint32_t triangle_count, float half_thickness, const float translation[3], float time_start) {
    BF_ENTER(c);
    if (!points || !frame_times || !triangles || !triangle_parts)
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "mesh arrays must not be NULL");
    if (frame_count < 2 || vertex_count == 0 || triangle_count == 0)
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "mesh needs >= 2 frames, >= 1 vertex and >= 1 triangle");
    for (uint32_t i = 0; i + 1 < frame_count; ++i)
        if (!(frame_times[i] < frame_times[i + 1]) || !std::isfinite(frame_times[i]) ||
            !std::isfinite(frame_times[i + 1]))
            return fail(c, BF_ERROR_INVALID_ARGUMENT, "frame times must be finite and strictly increasing");
    for (uint64_t i = 0; i < 3ull * frame_count * vertex_count; ++i)
        if (!std::isfinite(points[i])) return fail(c, BF_ERROR_INVALID_ARGUMENT, "vertex positions must be finite");
    if (bf_status r = check_triangles(c, points, vertex_count, triangles, triangle_parts, triangle_count,
                                      "the first frame"))
        return r;
    if (!(half_thickness > 0.0f)) return fail(c, BF_ERROR_INVALID_ARGUMENT, "half thickness must be positive");
    // Static triangles: every vertex at the same position in every frame
    // (their keys are the same every step unless a moving body carries the
    // mesh).
    std::vector<uint8_t> vertex_static(vertex_count, 1);
    for (uint32_t f = 1; f < frame_count; ++f)
        for (uint32_t v = 0; v < vertex_count; ++v) {
            const float* p0 = points + 3ull * v;
            const float* pf = points + 3ull * ((uint64_t)f * vertex_count + v);
            if (std::memcmp(p0, pf, 12) != 0) vertex_static[v] = 0;
        }
    MeshSource src;
    src.points = points;
    src.frame_times = frame_times;
    src.frame_count = frame_count;
    return install_mesh(c, src, vertex_count, triangles, triangle_parts, triangle_count, half_thickness, translation,
                        time_start, vertex_static);
}

bf_status bf_mesh_upload_articulated(bf_context* c, const float* rest, const uint8_t* vertex_joint,
                                     const float* span_fraction, uint32_t vertex_count, const uint32_t* triangles,
                                     const uint8_t* triangle_parts, uint32_t triangle_count, float half_thickness,
                                     const bf_kinematics* kinematics, float time_start) {
    BF_ENTER(c);
    if (!rest || !vertex_joint || !span_fraction || !triangles || !triangle_parts || !kinematics)
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "mesh arrays and kinematics must not be NULL");
    if (vertex_count == 0 || triangle_count == 0)
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "mesh needs >= 1 vertex and >= 1 triangle");
    const bf_kinematics& k = *kinematics;
    if (!(k.period > 0.0f) || !std::isfinite(k.period) || !std::isfinite(k.amplitude_scale))
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "the kinematics period must be positive and the scale finite");
    if (k.harmonics < 1 || k.harmonics > BF_MAX_HARMONICS || k.joint_count > BF_MAX_JOINTS)
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "1..4 harmonics and at most 8 joints");
    for (int i = 0; i < 3; ++i)
        if (!std::isfinite(k.center[i])) return fail(c, BF_ERROR_INVALID_ARGUMENT, "the centre must be finite");
    for (uint32_t j = 0; j < k.joint_count; ++j) {
        const bf_joint& J = k.joints[j];
        if (J.side != 1.0f && J.side != -1.0f) return fail(c, BF_ERROR_INVALID_ARGUMENT, "a joint's side is +1 or -1");
        const float* f = &J.root[0];
        for (size_t i = 0; i < 3 + 4 * 2 * BF_MAX_HARMONICS; ++i)
            if (!std::isfinite(f[i])) return fail(c, BF_ERROR_INVALID_ARGUMENT, "joint parameters must be finite");
    }
    // the rest pose with each joint's side applied (the degenerate-triangle
    // check), and the static vertices: those of joint 0
    std::vector<float> pose(3ull * vertex_count);
    std::vector<uint8_t> vertex_static(vertex_count, 0);
    for (uint32_t v = 0; v < vertex_count; ++v) {
        const uint32_t j = vertex_joint[v];
        if (j > k.joint_count) return fail(c, BF_ERROR_INVALID_ARGUMENT, "vertex joint out of range");
        for (int i = 0; i < 3; ++i)
            if (!std::isfinite(rest[3ull * v + i]))
                return fail(c, BF_ERROR_INVALID_ARGUMENT, "rest positions must be finite");
        if (!std::isfinite(span_fraction[v])) return fail(c, BF_ERROR_INVALID_ARGUMENT, "span fractions must be finite");
        const float side = j == 0 ? 1.0f : k.joints[j - 1].side;
        pose[3ull * v] = rest[3ull * v];
        pose[3ull * v + 1] = side * rest[3ull * v + 1];
        pose[3ull * v + 2] = rest[3ull * v + 2];
        vertex_static[v] = j == 0 ? 1 : 0;
    }
    if (bf_status r = check_triangles(c, pose.data(), vertex_count, triangles, triangle_parts, triangle_count,
                                      "the rest pose"))
        return r;
    if (!(half_thickness > 0.0f)) return fail(c, BF_ERROR_INVALID_ARGUMENT, "half thickness must be positive");
    MeshSource src;
    src.rest = rest;
    src.vertex_joint = vertex_joint;
    src.span_fraction = span_fraction;
    src.kinematics = kinematics;
    return install_mesh(c, src, vertex_count, triangles, triangle_parts, triangle_count, half_thickness, nullptr,
                        time_start, vertex_static);
}

bf_status bf_mesh_vertices_at(bf_context* c, float time, float* position3, float* velocity3, uint32_t vertex_count) {
    BF_ENTER(c);
    if (c->source != BF_GEOMETRY_MESH) return fail(c, BF_ERROR_INVALID_ARGUMENT, "no mesh uploaded");
    if (!position3 || !velocity3) return fail(c, BF_ERROR_INVALID_ARGUMENT, "outputs must not be NULL");
    if (vertex_count != c->m_vertices) return fail(c, BF_ERROR_INVALID_ARGUMENT, "vertex count differs from the mesh");
    // the prepare pass alone, on the geometry stream after any prefetched
    // geometry (which has consumed the prepared vertices by then; every
    // geometry build prepares them again)
    MeshBuffers b = mesh_buffers(c);
    const MeshParams m = mesh_params(c, time, true);
    hipError_t e = launch_mesh_prepare(b, m, c->sb);
    std::vector<float4> p(vertex_count), v(vertex_count);
    if (e == hipSuccess) e = hipMemcpyAsync(p.data(), c->m_pos, 16ull * vertex_count, hipMemcpyDeviceToHost, c->sb);
    if (e == hipSuccess) e = hipMemcpyAsync(v.data(), c->m_vel, 16ull * vertex_count, hipMemcpyDeviceToHost, c->sb);
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "mesh vertices");
    for (uint32_t i = 0; i < vertex_count; ++i) {
        position3[3ull * i] = p[i].x;
        position3[3ull * i + 1] = p[i].y;
        position3[3ull * i + 2] = p[i].z;
        velocity3[3ull * i] = v[i].x;
        velocity3[3ull * i + 1] = v[i].y;
        velocity3[3ull * i + 2] = v[i].z;
    }
    return BF_OK;
}

bf_status bf_mesh_build(bf_context* c, float t) {
    BF_ENTER(c);
    if (c->source != BF_GEOMETRY_MESH) return fail(c, BF_ERROR_INVALID_ARGUMENT, "no mesh uploaded");
    // Rebuild the current step's geometry at time t with the full key band
    // (geometry audit).  A prefetched next-step geometry stays valid: it
    // neither reads nor writes the current step's buffers.
    const int64_t st = c->step;
    hipError_t e0 = launch_fill_u8(c->mask[mod3(st)], 0, c->g.n, c->sb);
    if (e0 == hipSuccess) e0 = launch_fill_wall(c->wall[par2(st)], c->g.n, c->sb);
    if (e0 != hipSuccess) return hip_fail(c, e0, "mesh build");
    const float saved = c->time_start;
    c->time_start = t - (float)((double)st * c->cfg.time_step);
    bf_status r = enqueue_geometry(c, st, true, true);
    c->time_start = saved;
    if (r) return r;
    hipError_t e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "mesh build");
    return BF_OK;
}

// Sphere geometry of the current step; the previous step's mask starts
// equal to it.
static bf_status rebuild_sphere_geometry(bf_context* c) {
    const int64_t st = c->step;
    if (bf_status r = enqueue_geometry(c, st, true, false)) return r;
    const int k = mod3(st), kp = mod3(st - 1);
    hipError_t e = hipMemcpyAsync(c->mask[kp], c->mask[k], c->g.n, hipMemcpyDeviceToDevice, c->sb);
    if (e == hipSuccess) e = hipMemcpyAsync(c->dyn[kp], c->dyn[k], c->g.tiles, hipMemcpyDeviceToDevice, c->sb);
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "sphere geometry");
    c->geometry_ready = st;
    return BF_OK;
}

bf_status bf_set_translating_sphere(bf_context* c, const float cr[4], const float gv[3], const float wv[4]) {
    BF_ENTER(c);
    if (!cr || !gv || !wv) return fail(c, BF_ERROR_INVALID_ARGUMENT, "sphere parameters must not be NULL");
    c->sphere_cr = make_float4(cr[0], cr[1], cr[2], cr[3]);
    c->sphere_gv = make_float3(gv[0], gv[1], gv[2]);
    c->sphere_wv = make_float4(wv[0], wv[1], wv[2], wv[3]);
    c->source = BF_GEOMETRY_SPHERE;
    c->static_prev_differs = false;
    c->time_start = 0.0f;
    return rebuild_sphere_geometry(c);
}

// Bird geometry of the current step from cleared buffers (walls zero, as
// the reference's far cells), with no previous geometry; the previous
// step's mask starts equal to it.
static bf_status rebuild_bird_geometry(bf_context* c) {
    hipStream_t s = c->sb;
    hipError_t e = hipSuccess;
    for (int i = 0; i < 3 && e == hipSuccess; ++i) e = launch_fill_u8(c->mask[i], 0, c->g.n, s);
    for (int i = 0; i < 2 && e == hipSuccess; ++i) e = launch_fill_u32((uint32_t*)c->wall[i], 0u, 4u * c->g.n, s);
    for (int i = 0; i < 3 && e == hipSuccess; ++i) e = launch_fill_u8(c->dyn[i], 0, c->g.tiles, s);
    for (int i = 0; i < 4 && e == hipSuccess; ++i) e = launch_fill_u8(c->tflags[i], 0, c->g.tiles, s);
    if (e != hipSuccess) return hip_fail(c, e, "bird geometry clear");
    const int64_t st = c->step;
    if (bf_status r = enqueue_geometry(c, st, true, true)) return r;
    const int k = mod3(st), kp = mod3(st - 1);
    e = hipMemcpyAsync(c->mask[kp], c->mask[k], c->g.n, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(c->dyn[kp], c->dyn[k], c->g.tiles, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->tflags[mod4(st - 1)], c->tflags[mod4(st)], c->g.tiles, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(c, e, "initial bird geometry");
    c->geometry_ready = st;
    return BF_OK;
}

bf_status bf_set_bird(bf_context* c, const bf_bird_geometry* bird) {
    BF_ENTER(c);
    if (!bird) return fail(c, BF_ERROR_INVALID_ARGUMENT, "bird geometry must not be NULL");
    static_assert(sizeof(bf_bird_geometry) == sizeof(BirdShape), "bird layouts differ");
    const float* v = (const float*)bird;
    for (size_t i = 0; i < sizeof(bf_bird_geometry) / sizeof(float); ++i)
        if (!std::isfinite(v[i])) return fail(c, BF_ERROR_INVALID_ARGUMENT, "bird parameters must be finite");
    const bf_bird_geometry& b = *bird;
    if (!(b.body_radii[0] > 0.0f && b.body_radii[1] > 0.0f && b.body_radii[2] > 0.0f))
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "body radii must be positive");
    if (!(b.wing_span > 0.0f && b.wing_root_chord > 0.0f && b.wing_tip_chord > 0.0f && b.wing_thickness > 0.0f))
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "wing span, chords and thickness must be positive");
    if (!(b.tail_length >= 0.0f && b.tail_half_width >= 0.0f && b.tail_thickness >= 0.0f && b.frequency >= 0.0f))
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "tail sizes and flapping frequency must not be negative");
    if (c->bird_keyframe_count > 0 && !(b.frequency > 0.0f))
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "measured keyframes need a positive flapping frequency");
    std::memcpy(&c->bird, bird, sizeof(BirdShape));
    c->source = BF_GEOMETRY_BIRD;
    c->static_prev_differs = false;
    c->time_start = 0.0f;
    return rebuild_bird_geometry(c);
}

// Wing geometry of the current step from cleared buffers, no previous
// geometry; the previous step's mask starts equal to it.
static bf_status rebuild_wing_geometry(bf_context* c) {
    hipStream_t s = c->sb;
    hipError_t e = hipSuccess;
    for (int i = 0; i < 3 && e == hipSuccess; ++i) e = launch_fill_u8(c->mask[i], 0, c->g.n, s);
    for (int i = 0; i < 2 && e == hipSuccess; ++i) e = launch_fill_u32((uint32_t*)c->wall[i], 0u, 4u * c->g.n, s);
    for (int i = 0; i < 3 && e == hipSuccess; ++i) e = launch_fill_u8(c->dyn[i], 0, c->g.tiles, s);
    for (int i = 0; i < 4 && e == hipSuccess; ++i) e = launch_fill_u8(c->tflags[i], 0, c->g.tiles, s);
    if (e != hipSuccess) return hip_fail(c, e, "wing geometry clear");
    const int64_t st = c->step;
    if (bf_status r = enqueue_geometry(c, st, true, true)) return r;
    const int k = mod3(st), kp = mod3(st - 1);
    e = hipMemcpyAsync(c->mask[kp], c->mask[k], c->g.n, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(c->dyn[kp], c->dyn[k], c->g.tiles, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->tflags[mod4(st - 1)], c->tflags[mod4(st)], c->g.tiles, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(c, e, "initial wing geometry");
    c->geometry_ready = st;
    return BF_OK;
}

bf_status bf_set_flapping_wing(bf_context* c, const bf_flapping_wing* wing) {
    BF_ENTER(c);
    if (!wing) return fail(c, BF_ERROR_INVALID_ARGUMENT, "wing parameters must not be NULL");
    static_assert(sizeof(bf_flapping_wing) == sizeof(WingShape), "wing layouts differ");
    const float* v = (const float*)wing;
    for (size_t i = 0; i < sizeof(bf_flapping_wing) / sizeof(float); ++i)
        if (!std::isfinite(v[i])) return fail(c, BF_ERROR_INVALID_ARGUMENT, "wing parameters must be finite");
    const bf_flapping_wing& w = *wing;
    if (!(w.radius > 0.0f && w.chord_mean > 0.0f && w.thickness >= 0.0f && w.cycle > 0.0f &&
          w.beta_normalization > 0.0f))
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "wing span, mean chord, cycle and normalization must be positive");
    if (!(w.stroke_duration > 0.0f && w.stroke_duration <= 0.5f && w.pitch_duration > 0.0f &&
          w.pitch_duration <= 0.5f))
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "stroke and pitch durations must lie in (0, 1/2] of a cycle");
    if (!c->wing_links[0]) {
        hipError_t e = hipSuccess;
        const size_t entries = (size_t)c->g.tiles * 19 * TILE_CELLS;
        if (entries > 0xFFFFFF00ull)
            return fail(c, BF_ERROR_NOT_SUPPORTED, "the wing link table holds at most 2^32 - 256 entries");
        c->wing_frame = (float4*)c->wing_alloc.take(e, 2ull * WING_FRAME_ROWS * sizeof(float4));
        for (int k = 0; k < 2; ++k) c->wing_links[k] = (float*)c->wing_alloc.take(e, entries * sizeof(float));
        for (int k = 0; k < 2 && e == hipSuccess; ++k)
            e = launch_fill_u32((uint32_t*)c->wing_links[k], 0u, (uint32_t)entries, c->sb);
        if (e != hipSuccess) {
            c->wing_alloc.release();
            c->wing_frame = nullptr;
            c->wing_links[0] = c->wing_links[1] = nullptr;
            return fail(c, BF_ERROR_ALLOCATION_FAILED, std::string("wing link table: ") + hipGetErrorString(e));
        }
    }
    std::memcpy(&c->wing, wing, sizeof(WingShape));
    c->source = BF_GEOMETRY_WING;
    c->static_prev_differs = false;
    c->time_start = 0.0f;
    return rebuild_wing_geometry(c);
}

bf_status bf_read_wing_frame(bf_context* c, float frame[24]) {
    BF_ENTER(c);
    if (!frame) return fail(c, BF_ERROR_INVALID_ARGUMENT, "frame must not be NULL");
    if (c->source != BF_GEOMETRY_WING) return fail(c, BF_ERROR_INVALID_ARGUMENT, "no flapping wing set");
    hipError_t e = hipMemcpy(frame, c->wing_frame + WING_FRAME_ROWS * par2(c->step),
                             WING_FRAME_ROWS * sizeof(float4), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(c, e, "read wing frame");
    return BF_OK;
}

bf_status bf_read_wing_links(bf_context* c, float* links, uint64_t count) {
    BF_ENTER(c);
    if (!links) return fail(c, BF_ERROR_INVALID_ARGUMENT, "links must not be NULL");
    if (c->source != BF_GEOMETRY_WING) return fail(c, BF_ERROR_INVALID_ARGUMENT, "no flapping wing set");
    if (count != 19ull * c->g.n) return fail(c, BF_ERROR_INVALID_ARGUMENT, "link count must be 19 * N");
    std::vector<float> t((size_t)c->g.tiles * 19 * TILE_CELLS);
    hipError_t e = hipMemcpy(t.data(), c->wing_links[par2(c->step)], t.size() * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(c, e, "read wing links");
    const Grid& g = c->g;
    for (int z = 0; z < g.nz; ++z)
        for (int y = 0; y < g.ny; ++y)
            for (int x = 0; x < g.nx; ++x) {
                const size_t tile = ((size_t)z * g.tiles_y + (size_t)(y / TILE_Y)) * g.tiles_x + (size_t)(x / TILE_X);
                const size_t lane = (size_t)(x % TILE_X) + (size_t)TILE_X * (size_t)(y % TILE_Y);
                const size_t gid = (size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * (size_t)z);
                for (int q = 0; q < 19; ++q) links[(size_t)q * g.n + gid] = t[(tile * 19 + q) * TILE_CELLS + lane];
            }
    return BF_OK;
}

bf_status bf_read_ray_links(bf_context* c, float* links, uint64_t count) {
    BF_ENTER(c);
    if (!links) return fail(c, BF_ERROR_INVALID_ARGUMENT, "links must not be NULL");
    if (count != 19ull * c->g.n) return fail(c, BF_ERROR_INVALID_ARGUMENT, "links needs 19 * N floats");
    if (c->cfg.wall_rule != BF_WALL_INTERPOLATED_RAY || c->source != BF_GEOMETRY_MESH || !c->ray_links[0])
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "no exact-root table: the wall rule is not BF_WALL_INTERPOLATED_RAY "
                                                  "or no mesh is uploaded");
    hipError_t e = hipDeviceSynchronize();
    std::vector<float> t((size_t)c->g.tiles * 19 * TILE_CELLS);
    std::vector<uint8_t> dyn(c->g.tiles);
    if (e == hipSuccess)
        e = hipMemcpy(t.data(), c->ray_links[par2(c->step)], t.size() * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(dyn.data(), c->dyn[mod3(c->step)], dyn.size(), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(c, e, "read exact-root links");
    const Grid& g = c->g;
    std::memset(links, 0, sizeof(float) * 19ull * g.n);
    for (uint32_t tile = 0; tile < g.tiles; ++tile) {
        if (!dyn[tile]) continue;   // the table's other tiles hold no entry of this step
        int x0, y0, z;
        tile_origin(g, tile, x0, y0, z);
        for (uint32_t lane = 0; lane < TILE_CELLS; ++lane) {
            const int x = x0 + (int)(lane & (TILE_X - 1)), y = y0 + (int)(lane / TILE_X);
            if (x >= g.nx || y >= g.ny) continue;
            const uint64_t gid = (uint64_t)x + (uint64_t)g.nx * ((uint64_t)y + (uint64_t)g.ny * (uint64_t)z);
            for (int q = 0; q < 19; ++q) links[(uint64_t)q * g.n + gid] = t[((size_t)tile * 19 + q) * TILE_CELLS + lane];
        }
    }
    return BF_OK;
}

bf_status bf_set_bird_keyframes(bf_context* c, const float* keyframes, uint32_t count) {
    BF_ENTER(c);
    if (count == 0) {
        c->keyframe_alloc.release();
        c->bird_keyframes = nullptr;
        c->bird_keyframe_count = 0;
    } else {
        if (!keyframes) return fail(c, BF_ERROR_INVALID_ARGUMENT, "keyframes must not be NULL");
        if (count < 4 || count > 4096) return fail(c, BF_ERROR_INVALID_ARGUMENT, "keyframes must number 4...4096");
        if (c->source == BF_GEOMETRY_BIRD && !(c->bird.frequency > 0.0f))
            return fail(c, BF_ERROR_INVALID_ARGUMENT, "measured keyframes need a positive flapping frequency");
        if (keyframes[0] != 0.0f) return fail(c, BF_ERROR_INVALID_ARGUMENT, "the first keyframe phase must be zero");
        float previous = -1.0f;
        for (uint32_t k = 0; k < count; ++k) {
            const float* f = keyframes + 17ull * k;
            if (!(std::isfinite(f[0]) && f[0] >= 0.0f && f[0] < 1.0f && f[0] > previous))
                return fail(c, BF_ERROR_INVALID_ARGUMENT,
                            "keyframe phases must be finite, strictly increasing and in [0, 1)");
            previous = f[0];
            for (int i = 1; i < 17; ++i)
                if (!std::isfinite(f[i])) return fail(c, BF_ERROR_INVALID_ARGUMENT, "keyframe values must be finite");
        }
        c->keyframe_alloc.release();
        hipError_t e = hipSuccess;
        float* d = (float*)c->keyframe_alloc.take(e, 17ull * 4 * count);
        if (e == hipSuccess) e = hipMemcpy(d, keyframes, 17ull * 4 * count, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            c->keyframe_alloc.release();
            c->bird_keyframes = nullptr;
            c->bird_keyframe_count = 0;
            return hip_fail(c, e, "keyframe upload");
        }
        c->bird_keyframes = d;
        c->bird_keyframe_count = count;
    }
    // the current step's geometry follows the new kinematics
    return c->source == BF_GEOMETRY_BIRD ? rebuild_bird_geometry(c) : BF_OK;
}

bf_status bf_set_bird_wing_dynamics(bf_context* c, const bf_wing_dynamics* left, const bf_wing_dynamics* right) {
    BF_ENTER(c);
    if ((left == nullptr) != (right == nullptr))
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "give both wings' dynamics, or neither");
    if (!left) {
        c->wing_dynamics = false;
        return BF_OK;
    }
    const bf_wing_dynamics* w[2] = {left, right};
    for (int k = 0; k < 2; ++k) {
        bool ok = std::isfinite(w[k]->mass) && w[k]->mass >= 0.0f;
        for (int i = 0; i < 3; ++i)
            ok = ok && std::isfinite(w[k]->center_of_mass[i]) && std::isfinite(w[k]->inertia[i]) &&
                 w[k]->inertia[i] >= 0.0f;
        if (!ok)
            return fail(c, BF_ERROR_INVALID_ARGUMENT,
                        "wing mass and inertia must be finite and non-negative, the centre of mass finite");
        c->wing_dyn.mass[k] = w[k]->mass;
        for (int i = 0; i < 3; ++i) {
            c->wing_dyn.center_of_mass[k][i] = w[k]->center_of_mass[i];
            c->wing_dyn.inertia[k][i] = w[k]->inertia[i];
        }
    }
    c->wing_dynamics = true;
    // the current frame's momenta start the history (no reaction on it)
    return c->source == BF_GEOMETRY_BIRD ? rebuild_bird_geometry(c) : BF_OK;
}

bf_status bf_read_wing_reaction(bf_context* c, float reaction[16]) {
    BF_ENTER(c);
    if (!reaction) return fail(c, BF_ERROR_INVALID_ARGUMENT, "reaction must not be NULL");
    if (!c->wing_dynamics || c->source != BF_GEOMETRY_BIRD)
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "no analytic bird with wing dynamics set");
    hipError_t e = hipMemcpy(reaction, c->wing_reaction + 16 * par2(c->step), 16 * sizeof(float),
                             hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(c, e, "read wing reaction");
    return BF_OK;
}

bf_status bf_read_bird_frame(bf_context* c, float frame[56]) {
    BF_ENTER(c);
    if (!frame) return fail(c, BF_ERROR_INVALID_ARGUMENT, "frame must not be NULL");
    if (c->source != BF_GEOMETRY_BIRD) return fail(c, BF_ERROR_INVALID_ARGUMENT, "no bird geometry set");
    hipError_t e = hipMemcpy(frame, c->bird_frame + BIRD_FRAME_ROWS * par2(c->step),
                             BIRD_FRAME_ROWS * sizeof(float4), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(c, e, "read bird frame");
    return BF_OK;
}

static uint32_t ledger_blocks(const bf_context* c) {
    return std::min<uint32_t>(LEDGER_SCRATCH_BLOCKS, (c->g.n + 255u) / 256u);
}

static bf_status ensure_ledger(bf_context* c) {
    if (c->ledger_pv) return BF_OK;
    hipError_t e = hipSuccess;
    const uint32_t blocks = ledger_blocks(c);
    DeviceAlloc& A = c->ledger_alloc;
    double* pv = (double*)A.take(e, 16ull * 8 * blocks);
    unsigned long long* pn = (unsigned long long*)A.take(e, 4ull * 8 * blocks);
    double* out = (double*)A.take(e, 16 * 8);
    unsigned long long* out_n = (unsigned long long*)A.take(e, 4 * 8);
    if (e != hipSuccess) {
        A.release();
        return fail(c, BF_ERROR_ALLOCATION_FAILED, std::string("ledger buffers: ") + hipGetErrorString(e));
    }
    c->ledger_pv = pv;
    c->ledger_pn = pn;
    c->ledger_out = out;
    c->ledger_out_n = out_n;
    c->ledger_blocks = blocks;
    return BF_OK;
}

bf_status bf_ledger_fluid(bf_context* c, double out[4]) {
    BF_ENTER(c);
    if (!out) return fail(c, BF_ERROR_INVALID_ARGUMENT, "output must not be NULL");
    if (bf_status r = ensure_ledger(c)) return r;
    LedgerBuffers b{};
    b.f_out = c->f[c->cur];
    b.solid_cur = geometry_of(c, c->step).mask_cur;
    b.blocks = c->ledger_blocks;
    b.partial_v = c->ledger_pv;
    b.partial_n = c->ledger_pn;
    b.out_v = c->ledger_out;
    b.out_n = c->ledger_out_n;
    hipError_t e = launch_ledger((int)c->cfg.storage, b, make_step_params(c), false, c->sb);
    if (e == hipSuccess) e = hipMemcpyAsync(out, c->ledger_out, 4 * 8, hipMemcpyDeviceToHost, c->sb);
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "fluid momentum");
    return BF_OK;
}

// The step-source sums of bf_ledger_step_sources into `sums` / `counts`
// through the partial and output buffers of `b`.
static bf_status ledger_step_sources_into(bf_context* c, LedgerBuffers b, double sums[16], uint64_t counts[4]) {
    const StepGeometry sg = geometry_of(c, c->step);
    b.f_in = c->f[1 - c->cur];
    b.f_out = c->f[c->cur];
    b.solid_prev = c->ledger_prev;
    b.solid_cur = sg.mask_cur;
    b.wall = sg.wall;
    hipError_t e = launch_ledger((int)c->cfg.storage, b, make_step_params(c), true, c->sb);
    if (e == hipSuccess) e = hipMemcpyAsync(sums, b.out_v, 16 * 8, hipMemcpyDeviceToHost, c->sb);
    if (e == hipSuccess) e = hipMemcpyAsync(counts, b.out_n, 4 * 8, hipMemcpyDeviceToHost, c->sb);
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "step momentum sources");
    return BF_OK;
}

static bf_status ledger_sources_ready(bf_context* c, const double* sums, const uint64_t* counts) {
    if (!sums || !counts) return fail(c, BF_ERROR_INVALID_ARGUMENT, "outputs must not be NULL");
    if (c->ledger_step != c->step || c->ledger_epoch != c->f_writes || c->step < 1)
        return fail(c, BF_ERROR_INVALID_ARGUMENT,
                    "the ledger needs the inputs of the last step: call it right after bf_step");
    return BF_OK;
}

bf_status bf_ledger_step_sources(bf_context* c, double sums[16], uint64_t counts[4]) {
    BF_ENTER(c);
    if (bf_status r = ledger_sources_ready(c, sums, counts)) return r;
    if (bf_status r = ensure_ledger(c)) return r;
    LedgerBuffers b{};
    b.blocks = c->ledger_blocks;
    b.partial_v = c->ledger_pv;
    b.partial_n = c->ledger_pn;
    b.out_v = c->ledger_out;
    b.out_n = c->ledger_out_n;
    return ledger_step_sources_into(c, b, sums, counts);
}

bf_status bf_ledger_step_sources_noalloc(bf_context* c, double sums[16], uint64_t counts[4]) {
    BF_ENTER(c);
    if (bf_status r = ledger_sources_ready(c, sums, counts)) return r;
    const uint32_t blocks = ledger_blocks(c);
    if (blocks > LEDGER_SCRATCH_BLOCKS)
        return fail(c, BF_ERROR_INVALID_ARGUMENT,
                    "the ledger needs " + std::to_string(blocks) + " block partials; the no-allocation scratch holds " +
                        std::to_string(LEDGER_SCRATCH_BLOCKS));
    LedgerBuffers b{};
    hipError_t e = ledger_scratch_buffers(c->device, blocks, b);
    if (e != hipSuccess) return hip_fail(c, e, "no-allocation ledger scratch", false);
    return ledger_step_sources_into(c, b, sums, counts);
}

bf_status bf_step(bf_context* c, uint32_t steps) {
    if (!c) return BF_ERROR_INVALID_ARGUMENT;
    if (c->invalidated) return BF_ERROR_STATE_INVALIDATED;
    (void)hipSetDevice(c->device);
    if (steps == 0) return BF_OK;
    if (c->cfg.wall_rule == BF_WALL_INTERPOLATED_LINKS && c->source != BF_GEOMETRY_WING)
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "the link-table wall rule needs the prescribed flapping wing");
    if (c->cfg.wall_rule == BF_WALL_INTERPOLATED_RAY && (c->source != BF_GEOMETRY_MESH || !c->ray_links[0]))
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "the exact-root wall rule needs an uploaded mesh");
    const int storage = (int)c->cfg.storage, coll = (int)c->cfg.collision, arith = arithmetic_of(c->cfg);
    hipError_t e = hipSuccess;
    // Prologue: geometry (unless already built) and tile list of the first
    // step of this call; event R of the previous step then releases fast(s).
    {
        const int64_t s = c->step + 1;
        if (c->source != BF_GEOMETRY_STATIC && c->geometry_ready < s) {
            if (bf_status r = enqueue_geometry(c, s, false, false)) return r;
        }
        if (bf_status r = enqueue_tile_list(c, s)) return r;
        e = hipEventRecord(c->ev_ready[par2(s)], c->sb);
        if (e != hipSuccess) return hip_fail(c, e, "event record");
        c->geometry_ready = s;
    }
    for (uint32_t i = 0; i < steps; ++i) {
        const int64_t s = c->step + 1;
        const bool prof = c->profiling && c->profiled < PROFILE_CAPACITY;
        const bool prof_all = prof && c->profiling_level >= 2;
        hipEvent_t* pe = nullptr;
        if (prof) {
            while (c->events.size() < (size_t)(c->profiled + 1) * PROFILE_EVENTS) {
                hipEvent_t ev;
                if (hipEventCreate(&ev) != hipSuccess) return fail(c, BF_ERROR_ALLOCATION_FAILED, "profile event");
                c->events.push_back(ev);
            }
            pe = &c->events[(size_t)c->profiled * PROFILE_EVENTS];
        }
        StepParams p = make_step_params(c);
        // the link table of step s; geometry(s + 1) fills the other slot meanwhile
        p.links = (c->cfg.wall_rule == BF_WALL_INTERPOLATED_RAY ? c->ray_links : c->wing_links)[par2(s)];
        p.health = c->health;
        p.step_index = (unsigned long long)s;
        const FluidBuffers fb = fluid_buffers(c, s);
        c->ledger_prev = fb.solid_prev;
        if (c->source == BF_GEOMETRY_STATIC) c->static_prev_differs = false;
        // stream A: the wall tiles of step s, then the wall-free tiles, once
        // the tile list of step s (event L[s], stream B) is built.  Both read
        // f_in, which the previous step finished writing on this stream.
        e = hipStreamWaitEvent(c->sa, c->ev_ready[par2(s)], 0);
        if (prof_all && e == hipSuccess) e = hipEventRecord(pe[4], c->sa);
        if (e == hipSuccess) e = launch_fluid_general(arith, storage, coll, fb, p, c->general_blocks, c->sa);
        if (prof_all && e == hipSuccess) e = hipEventRecord(pe[5], c->sa);
        if (e == hipSuccess) e = hipEventRecord(c->ev_general[par2(s)], c->sa);
        if (prof && e == hipSuccess) e = hipEventRecord(pe[0], c->sa);
        if (e == hipSuccess) e = launch_fluid_fast(arith, storage, coll, fb, p, c->sa);
        if (prof && e == hipSuccess) e = hipEventRecord(pe[1], c->sa);
        if (e == hipSuccess) e = hipEventRecord(c->ev_fast[par2(s)], c->sa);
        if (e != hipSuccess) return hip_fail(c, e, "fluid step launch");
        // stream B: geometry and tile list of step s+1 while step s streams
        // (they write only buffers that step s-1 last read; this stream
        // waited for step s-1 before its load reduction), then event L[s+1];
        // then the load reduction of step s (partials double-buffered by
        // step parity) and the body update.  general(s+1) reads the torque
        // reference point from the body state, so in free flight L[s+1]
        // follows the body update; a mesh riding on a free-flying body also
        // needs the pose after step s, so its geometry follows it too.
        const bool coupled = c->free_flight && ((c->move_mesh && c->source == BF_GEOMETRY_MESH) ||
                                                c->source == BF_GEOMETRY_BIRD);
        auto next_geometry = [&]() -> bf_status {
            if (prof_all) e = hipEventRecord(pe[2], c->sb);
            if (e == hipSuccess && c->source != BF_GEOMETRY_STATIC) {
                if (bf_status r = enqueue_geometry(c, s + 1, false, false)) return r;
            }
            if (bf_status r = enqueue_tile_list(c, s + 1)) return r;
            if (prof_all) e = hipEventRecord(pe[3], c->sb);
            return e == hipSuccess ? BF_OK : hip_fail(c, e, "geometry launch");
        };
        if (!coupled) {
            // after general(s): its latency-bound wavefronts would otherwise
            // share the machine with the raster; the geometry then overlaps
            // fast(s) only
            if (c->geometry_after_general) e = hipStreamWaitEvent(c->sb, c->ev_general[par2(s)], 0);
            if (e != hipSuccess) return hip_fail(c, e, "event wait");
            if (bf_status r = next_geometry()) return r;
        }
        if (!c->free_flight) e = hipEventRecord(c->ev_ready[par2(s + 1)], c->sb);
        // Coupled free flight puts the reduction, body update and next
        // geometry on the critical path.  With TRT the loads come from the
        // wall tiles alone (the fast kernel adds limiter activations only for
        // the other collisions), so they start when general(s) ends and
        // overlap fast(s): fast(s) reads none of the buffers they write
        // (geometry s+1 writes the dyn flags of s-2, which fast(s-1) read and
        // general(s) followed), and general(s) itself follows fast(s-1).
        // Uncoupled, the geometry of s+1 may precede general(s) on this
        // stream, so the reduction keeps waiting for fast(s).
        const bool early_loads = coupled && c->cfg.collision == BF_COLLISION_TRT;
        if (e == hipSuccess)
            e = hipStreamWaitEvent(c->sb, early_loads ? c->ev_general[par2(s)] : c->ev_fast[par2(s)], 0);
        if (e == hipSuccess) e = launch_load_reduction(fb, p, (uint32_t)(s % HISTORY_CAPACITY), c->sb);
        if (e == hipSuccess && c->free_flight)
            e = launch_body_step(c->body, c->last_load, c->body_history, (uint32_t)(s % HISTORY_CAPACITY),
                                 c->safety_metrics, c->safety_event, (uint64_t)s, c->body_params, c->sb,
                                 c->wing_dynamics && c->source == BF_GEOMETRY_BIRD ? c->wing_reaction + 16 * par2(s)
                                                                                   : nullptr);
        if (e != hipSuccess) return hip_fail(c, e, "fluid step launch");
        if (coupled) {
            if (bf_status r = next_geometry()) return r;
        }
        if (c->free_flight) {
            e = hipEventRecord(c->ev_ready[par2(s + 1)], c->sb);
            if (e != hipSuccess) return hip_fail(c, e, "event record");
        }
        c->step = s;
        c->cur = 1 - c->cur;
        ++c->history_count;
        if (c->free_flight) ++c->body_history_count;
        c->geometry_ready = s + 1;
        if (prof) ++c->profiled;
    }
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(c, e, "step");
    c->ledger_epoch = c->f_writes;
    c->ledger_step = c->step;
    return BF_OK;
}

bf_status bf_synchronize(bf_context* c) {
    if (!c) return BF_ERROR_INVALID_ARGUMENT;
    if (c->invalidated) return BF_ERROR_STATE_INVALIDATED;
    (void)hipSetDevice(c->device);
    hipError_t e = drain(c);
    if (e != hipSuccess) return hip_fail(c, e, "synchronize");
    if (!c->health_reported) {
        // a non-finite population is reported once, with the first step
        // that stored one
        unsigned long long h[2];
        e = hipMemcpy(h, c->health, sizeof h, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(c, e, "read fluid health");
        if (h[0] != 0ull) {
            c->health_reported = true;
            char msg[160];
            std::snprintf(msg, sizeof msg, "runtime safety violation: non-finite populations from step %llu",
                          (unsigned long long)h[1]);
            c->error = msg;
            return BF_ERROR_RUNTIME_SAFETY_VIOLATION;
        }
    }
    if (c->free_flight && c->body_params.monitor) {
        // a new runtime-safety event is reported once (the reference throws
        // runtimeSafetyViolation after the batch, BirdFlowSimulation.swift:735-741)
        uint32_t ev[4];
        e = hipMemcpy(ev, c->safety_event, sizeof ev, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(c, e, "read safety record");
        if (ev[2] != 0u && c->safety_reported == 0u) {
            c->safety_reported = ev[2];
            char msg[160];
            std::snprintf(msg, sizeof msg, "runtime safety violation at step %llu (flags %u: 1 Mach, 2 clearance, "
                          "4 non-finite)", (unsigned long long)ev[0] | ((unsigned long long)ev[1] << 32), ev[2]);
            c->error = msg;
            return BF_ERROR_RUNTIME_SAFETY_VIOLATION;
        }
    }
    return BF_OK;
}

bf_status bf_body_set(bf_context* c, const bf_body_state* state, const bf_body_params* params) {
    BF_ENTER(c);
    if (params) {
        if (!(params->mass > 0.0) || !(params->inertia[0] > 0.0) || !(params->inertia[1] > 0.0) ||
            !(params->inertia[2] > 0.0) || params->substeps < 1 || params->free_flight > 1u || params->move_mesh > 1u)
            return fail(c, BF_ERROR_INVALID_ARGUMENT, "body parameters: mass and inertia > 0, substeps >= 1");
        if (params->free_flight && !c->cfg.accumulate_loads)
            return fail(c, BF_ERROR_INVALID_ARGUMENT, "free flight integrates the fluid loads: accumulate_loads must be on");
        BodyParams& b = c->body_params;
        b.mass = params->mass;
        for (int i = 0; i < 3; ++i) {
            b.inertia[i] = params->inertia[i];
            b.gravity[i] = params->gravity[i];
            b.external_force[i] = params->external_force[i];
            b.external_torque[i] = params->external_torque[i];
            b.extent[i] = params->safety_extent[i];
            const double dx = c->cfg.cell_size;
            const double margin = ((double)c->cfg.sponge_width + 3.0) * dx;   // BirdFlow.metal:6650-6657
            const double n = i == 0 ? c->cfg.nx : (i == 1 ? c->cfg.ny : c->cfg.nz);
            b.domain_min[i] = (double)c->cfg.origin[i] + margin;
            b.domain_max[i] = (double)c->cfg.origin[i] + n * dx - margin;
            b.far_velocity_world[i] = (double)c->cfg.far_velocity[i] / (double)c->cfg.velocity_to_lattice;
        }
        b.dt = c->cfg.time_step;
        b.substeps = (int)params->substeps;
        b.monitor = params->mach_limit > 0.0 ? 1 : 0;
        b.rotation_radius = params->safety_radius;
        b.wing_speed = params->wing_speed;
        b.mach_limit = params->mach_limit;
        b.velocity_to_lattice = c->cfg.velocity_to_lattice;
        c->free_flight = params->free_flight != 0u;
        c->move_mesh = params->move_mesh != 0u;
    }
    if (state) {
        double v[16] = {0};
        for (int i = 0; i < 3; ++i) {
            v[i] = state->position[i];
            v[7 + i] = state->linear_velocity[i];
            v[10 + i] = state->angular_velocity_body[i];
        }
        for (int i = 0; i < 4; ++i) v[3 + i] = state->orientation[i];
        hipError_t e = hipMemcpy(c->body, v, sizeof v, hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(c, e, "body upload");
        // a mesh carried by the body is placed again for the next step
        if ((c->move_mesh && c->source == BF_GEOMETRY_MESH) || c->source == BF_GEOMETRY_BIRD)
            c->geometry_ready = c->step;
    }
    if (state || params) {
        // a new body (or new body parameters) starts a new safety record
        const double safety0[4] = {0.0, DBL_MAX, 0.0, 0.0};
        const uint32_t event0[4] = {0u, 0u, 0u, 0u};
        hipError_t e = hipMemcpy(c->safety_metrics, safety0, sizeof safety0, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(c->safety_event, event0, sizeof event0, hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(c, e, "safety reset");
        c->safety_reported = 0u;
        c->body_history_count = 0;
    }
    return BF_OK;
}

bf_status bf_body_get(bf_context* c, bf_body_state* state) {
    BF_ENTER(c);
    if (!state) return BF_ERROR_INVALID_ARGUMENT;
    double v[16];
    hipError_t e = hipMemcpy(v, c->body, sizeof v, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(c, e, "body read");
    for (int i = 0; i < 3; ++i) {
        state->position[i] = v[i];
        state->linear_velocity[i] = v[7 + i];
        state->angular_velocity_body[i] = v[10 + i];
    }
    for (int i = 0; i < 4; ++i) state->orientation[i] = v[3 + i];
    return BF_OK;
}

bf_status bf_read_body_history(bf_context* c, bf_body_state* out, uint32_t max_records, uint32_t* records) {
    BF_ENTER(c);
    if (!out || !records) return BF_ERROR_INVALID_ARGUMENT;
    const uint64_t avail = std::min<uint64_t>(c->body_history_count, HISTORY_CAPACITY);
    const uint64_t want = std::min<uint64_t>(avail, max_records);
    std::vector<double> all(16ull * HISTORY_CAPACITY);
    hipError_t e = hipMemcpy(all.data(), c->body_history, 128ull * HISTORY_CAPACITY, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(c, e, "read body history");
    for (uint64_t i = 0; i < want; ++i) {
        const uint64_t step = (uint64_t)c->step - want + 1 + i;
        const double* v = all.data() + 16 * (step % HISTORY_CAPACITY);
        for (int k = 0; k < 3; ++k) {
            out[i].position[k] = v[k];
            out[i].linear_velocity[k] = v[7 + k];
            out[i].angular_velocity_body[k] = v[10 + k];
        }
        for (int k = 0; k < 4; ++k) out[i].orientation[k] = v[3 + k];
    }
    *records = (uint32_t)want;
    return BF_OK;
}

bf_status bf_read_safety(bf_context* c, double metrics[4], uint64_t* event_step, uint32_t* flags) {
    BF_ENTER(c);
    if (!metrics || !event_step || !flags) return BF_ERROR_INVALID_ARGUMENT;
    uint32_t ev[4];
    hipError_t e = hipMemcpy(metrics, c->safety_metrics, 4 * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(ev, c->safety_event, sizeof ev, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(c, e, "read safety record");
    *event_step = (uint64_t)ev[0] | ((uint64_t)ev[1] << 32);
    *flags = ev[2];
    unsigned long long h[2];
    e = hipMemcpy(h, c->health, sizeof h, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(c, e, "read fluid health");
    if (h[0] != 0ull) {
        if (*flags == 0u) *event_step = (uint64_t)h[1];
        *flags |= 8u;
    }
    return BF_OK;
}

bf_status bf_population_minimum(bf_context* c, float* value, uint64_t* index, uint32_t* non_finite) {
    BF_ENTER(c);
    if (!value || !index || !non_finite) return fail(c, BF_ERROR_INVALID_ARGUMENT, "outputs must not be NULL");
    const uint64_t total = 19ull * c->g.n;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(POPULATION_MINIMUM_BLOCKS, (total + 255) / 256);
    hipError_t e = launch_population_minimum((int)c->cfg.storage, c->f[c->cur], c->g, c->popmin, blocks,
                                             c->popmin + POPULATION_MINIMUM_BLOCKS, c->sb);
    PopulationMinimum r{};
    if (e == hipSuccess)
        e = hipMemcpyAsync(&r, c->popmin + POPULATION_MINIMUM_BLOCKS, sizeof r, hipMemcpyDeviceToHost, c->sb);
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "population minimum");
    *value = r.value;
    *index = r.index;
    *non_finite = r.non_finite;
    return BF_OK;
}

bf_status bf_inject_population(bf_context* c, uint64_t index, float value) {
    BF_ENTER(c);
    if (index >= 19ull * c->g.n) return fail(c, BF_ERROR_INVALID_ARGUMENT, "population index beyond 19 * N");
    ++c->f_writes;
    hipError_t e = launch_poke_population((int)c->cfg.storage, c->f[c->cur], c->g, index, value, c->sb);
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "inject population");
    return BF_OK;
}

// ---------------------------------------------------------------------------
// Adjoint (reverse sweep with checkpoint / recompute).
// ---------------------------------------------------------------------------
namespace {

bf_status adjoint_supported(bf_context* c) {
    const bf_config& g = c->cfg;
    if (g.force_mode != 6) return fail(c, BF_ERROR_NOT_SUPPORTED, "the adjoint differentiates force mode 6");
    if (!g.accumulate_loads) return fail(c, BF_ERROR_INVALID_ARGUMENT, "the objective needs accumulate_loads");
    for (int a = 1; a < 3; ++a)
        for (int sd = 0; sd < 2; ++sd)
            if (g.boundary[a][sd] == BF_BC_OUTLET)
                return fail(c, BF_ERROR_NOT_SUPPORTED,
                            "the adjoint covers periodic and far-field faces, and outlets on the x faces");
    if ((g.boundary[0][0] == BF_BC_OUTLET || g.boundary[0][1] == BF_BC_OUTLET) && g.nx < 2)
        return fail(c, BF_ERROR_NOT_SUPPORTED, "an outlet face needs at least two cells along x");
    if (c->free_flight) return fail(c, BF_ERROR_NOT_SUPPORTED, "the adjoint holds the body fixed (no free flight)");
    if (c->source == BF_GEOMETRY_SPHERE || c->source == BF_GEOMETRY_BIRD || c->source == BF_GEOMETRY_WING ||
        g.wall_rule == BF_WALL_INTERPOLATED_LINKS)
        return fail(c, BF_ERROR_NOT_SUPPORTED, "the adjoint covers mesh and static geometry");
    if (g.wall_rule == BF_WALL_INTERPOLATED_RAY && (c->source != BF_GEOMETRY_MESH || !c->ray_links[0]))
        return fail(c, BF_ERROR_NOT_SUPPORTED, "the exact-root adjoint needs a mesh");
    if (c->source == BF_GEOMETRY_STATIC && c->static_prev_differs)
        return fail(c, BF_ERROR_NOT_SUPPORTED, "step once after bf_set_geometry with a different previous mask");
    return BF_OK;
}

bf_status adjoint_alloc(bf_context* c, uint32_t interval, uint32_t segments) {
    auto& a = c->adj;
    // the vertex buffers are sized by the mesh of the allocation: a larger
    // mesh uploaded since needs new ones
    if (a.interval == interval && a.segments >= segments && a.checkpoints && a.verts >= c->m_vertices) return BF_OK;
    a.alloc.release();
    // pop: one FP32 lattice (lambda); popS: one lattice in the storage's form
    const uint64_t n = c->g.n, tiles = c->g.tiles, pop = 19ull * n * 4, popS = 19ull * n * c->pop_elem;
    const uint64_t verts = std::max<uint32_t>(c->m_vertices, 1u);
    const bool ray = c->cfg.wall_rule == BF_WALL_INTERPOLATED_RAY;
    const uint64_t need = popS * segments + popS * interval + 2 * (uint64_t)(interval + 1) * (21 * n + 7 * tiles + 4) +
                          2 * pop + tiles * 19 * 256 * 4 + 2 * 8 * 16 * n + 4 * 8 * 16 * verts + 6 * 16 * verts +
                          tiles * 64 + 4 * tiles + 2ull * 19 * c->g.ny * c->g.nz * 4 + 1024 +
                          (ray ? 2 * 16 * verts + 4 * tiles * 19 * 256 * 4 + 4 * n + 4 * tiles + 256 : 0);
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    if (need > (uint64_t)free_b) {
        char msg[200];
        std::snprintf(msg, sizeof msg, "adjoint buffers need %.2f GB, %.2f GB free: lower the checkpoint interval",
                      need / 1e9, free_b / 1e9);
        return fail(c, BF_ERROR_WORKING_SET_EXCEEDS_MEMORY, msg);
    }
    hipError_t e = hipSuccess;
    DeviceAlloc& A = a.alloc;
    a.checkpoints = (uint8_t*)A.take(e, popS * segments);
    a.seg_f = (uint8_t*)A.take(e, popS * interval);
    for (int i = 0; i < 2; ++i) {
        a.seg_mask[i] = (uint8_t*)A.take(e, (interval + 1) * n);
        a.seg_wall[i] = (float4*)A.take(e, (interval + 1) * 16 * n);
        a.seg_tri[i] = (uint32_t*)A.take(e, (interval + 1) * 4 * n);
        a.seg_dyn[i] = (uint8_t*)A.take(e, (interval + 1) * tiles);
    }
    for (int i = 0; i < 2; ++i) {
        a.slot_tiles[i] = (uint8_t*)A.take(e, 2 * (interval + 1) * tiles);
        a.slot_list[i] = (uint32_t*)A.take(e, (interval + 1) * 4 * tiles);
        a.slot_count[i] = (uint32_t*)A.take(e, 4 * (interval + 1));
    }
    a.gpos = (float4*)A.take(e, 16 * verts);
    a.gvel = (float4*)A.take(e, 16 * verts);
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
        if (!a.ev_geo[i]) e = hipEventCreateWithFlags(&a.ev_geo[i], hipEventDisableTiming);
        if (e == hipSuccess && !a.ev_done[i]) e = hipEventCreateWithFlags(&a.ev_done[i], hipEventDisableTiming);
    }
    a.lam[0] = (float*)A.take(e, pop);
    a.lam[1] = (float*)A.take(e, pop);
    a.nu = (float*)A.take(e, tiles * 19 * 256 * 4);
    a.dW[0] = (float4*)A.take(e, 8 * 16 * n);
    a.dW[1] = (float4*)A.take(e, 8 * 16 * n);
    if (e == hipSuccess && !a.sc) e = hipStreamCreateWithFlags(&a.sc, hipStreamNonBlocking);
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
        if (!a.ev_tan[i]) e = hipEventCreateWithFlags(&a.ev_tan[i], hipEventDisableTiming);
        if (e == hipSuccess && !a.ev_used[i]) e = hipEventCreateWithFlags(&a.ev_used[i], hipEventDisableTiming);
    }
    for (int i = 0; i < 2; ++i) {
        a.vpos[i] = (float4*)A.take(e, 16 * verts);
        a.vvel[i] = (float4*)A.take(e, 16 * verts);
        a.dpos[i] = (float4*)A.take(e, 8 * 16 * verts);
        a.dvel[i] = (float4*)A.take(e, 8 * 16 * verts);
    }
    a.rpos = a.rvel = nullptr;
    a.rtab[0] = a.rtab[1] = nullptr;
    a.rwin[0] = a.rwin[1] = nullptr;
    a.rdirs = a.rlist = a.rcount = nullptr;
    if (ray) {
        a.rpos = (float4*)A.take(e, 16 * verts);
        a.rvel = (float4*)A.take(e, 16 * verts);
        for (int i = 0; i < 2; ++i) {
            a.rtab[i] = (float*)A.take(e, tiles * 19 * 256 * 4);
            a.rwin[i] = (uint32_t*)A.take(e, tiles * 19 * 256 * 4);
        }
        a.rdirs = (uint32_t*)A.take(e, 4 * n);
        a.rlist = (uint32_t*)A.take(e, 4 * tiles);
        a.rcount = (uint32_t*)A.take(e, 256);
        // the ray pass leaves its link directions cleared; they start so
        if (e == hipSuccess) e = launch_fill_u32(a.rdirs, 0u, (uint32_t)n, c->sb);
    }
    a.grad_partials = (double*)A.take(e, tiles * 64);
    a.grad_total = (double*)A.take(e, 64);
    // clamped pulls of the outlet x faces: [face][19][ny nz]
    a.outlet_extra = (float*)A.take(e, 2ull * 19 * c->g.ny * c->g.nz * 4);
    if (e != hipSuccess) {
        A.release();
        a.checkpoints = nullptr;
        return fail(c, BF_ERROR_ALLOCATION_FAILED, "adjoint buffers");
    }
    // geometry slots start cleared, with no touched tiles
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
        for (uint64_t k = 0; k <= interval && e == hipSuccess; ++k) {
            e = launch_fill_u8(a.seg_mask[i] + k * n, 0, (uint32_t)n, c->sb);
            if (e == hipSuccess) e = launch_fill_wall(a.seg_wall[i] + k * n, (uint32_t)n, c->sb);
            if (e == hipSuccess) e = launch_fill_u32(a.seg_tri[i] + k * n, 0xFFFFFFFFu, (uint32_t)n, c->sb);
        }
        if (e == hipSuccess) e = launch_fill_u8(a.slot_tiles[i], 0, (uint32_t)(2 * (interval + 1) * tiles), c->sb);
        a.gen[i] = 0;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) {
        A.release();
        a.checkpoints = nullptr;
        return hip_fail(c, e, "adjoint buffers");
    }
    a.interval = interval;
    a.segments = segments;
    a.verts = verts;
    return BF_OK;
}

struct SlotGeometry {
    const uint8_t* mask;
    const float4* wall;
    const uint8_t* dyn;
    const uint32_t* tri;
};

// The ray pass's buffers for one step's geometry slot.
RayBuffers ray_buffers(const bf_context* c, const SlotGeometry& gs, const float4* pos, uint32_t* list,
                       uint32_t* count, uint32_t* dirs, float* links, uint32_t* winners) {
    RayBuffers rb{};
    rb.pos = pos;
    rb.tri = c->m_tri;
    rb.mask = gs.mask;
    rb.wall = gs.wall;
    rb.dyn = gs.dyn;
    rb.list = list;
    rb.list_count = count;
    rb.dirs = dirs;
    rb.links = links;
    rb.winners = winners;
    if (!winners) set_ray_static(c, rb);
    return rb;
}

// Geometry of step s into segment slot k: the same raster / resolve as the
// forward pass (masks identical, walls identical wherever the step reads
// them), plus each keyed cell's nearest triangle.
SlotGeometry slot_geometry(const bf_context* c, uint32_t k, int set) {
    const auto& a = c->adj;
    if (c->source == BF_GEOMETRY_STATIC) return {c->mask[0], c->wall[0], c->dyn[0], nullptr};
    const uint64_t n = c->g.n;
    return {a.seg_mask[set] + k * n, a.seg_wall[set] + k * n, a.seg_dyn[set] + (uint64_t)k * c->g.tiles,
            a.seg_tri[set] + k * n};
}

// Geometry of step s into slot k of set `set`, on stream `st` (its own
// vertex buffers; keys, geometry list and touched-tile flags are used by no
// other work during a sweep).
bf_status adjoint_geometry(bf_context* c, int64_t s, uint32_t k, int set, hipStream_t st) {
    auto& a = c->adj;
    if (c->source == BF_GEOMETRY_STATIC) return BF_OK;
    const uint64_t n = c->g.n;
    uint8_t* mask = a.seg_mask[set] + k * n;
    float4* wall = a.seg_wall[set] + k * n;
    uint32_t* tri = a.seg_tri[set] + k * n;
    uint8_t* dyn = a.seg_dyn[set] + (uint64_t)k * c->g.tiles;
    // The slot holds the geometry of its previous build: the tile pass
    // clears the tiles that build touched (old flags) and resolves the new
    // ones, as in stepping.
    const uint64_t tiles = c->g.tiles, slots = a.interval + 1;
    uint8_t* touched_new = a.slot_tiles[set] + ((uint64_t)a.gen[set] * slots + k) * tiles;
    uint8_t* touched_old = a.slot_tiles[set] + ((uint64_t)(1 - a.gen[set]) * slots + k) * tiles;
    hipError_t e = hipSuccess;
    MeshBuffers b = mesh_buffers(c);   // frames or articulated arrays
    b.pos = a.gpos;
    b.vel = a.gvel;
    b.keys = c->keys;
    b.tiles_new = touched_new;
    b.tiles_mask_old = touched_old;
    b.tiles_wall_old = touched_old;
    b.list = c->geo_list;
    b.list_count = &c->counters[2];
    b.mask = mask;
    b.wall = wall;
    b.dyn = dyn;
    b.tri_out = tri;
    e = launch_mesh_geometry(b, mesh_params(c, time_of(c, s), false), st);
    if (e != hipSuccess) return hip_fail(c, e, "adjoint geometry");
    return BF_OK;
}

// Geometry of every step of segment `seg` into set seg & 1, on stream A.
bf_status segment_geometry(bf_context* c, int64_t s0, uint32_t steps, uint32_t interval, int64_t seg) {
    const int64_t sa_ = s0 + seg * interval;
    const int64_t sb_ = std::min<int64_t>(sa_ + interval, s0 + steps);
    const int set = (int)(seg & 1);
    auto& a = c->adj;
    a.gen[set] ^= 1;
    for (int64_t s = sa_; s <= sb_; ++s)
        if (bf_status r = adjoint_geometry(c, s, (uint32_t)(s - sa_), set, c->sa)) return r;
    // slots a short (last) segment leaves alone keep their contents, and so
    // their touched-tile flags
    const uint64_t tiles = c->g.tiles, slots = a.interval + 1;
    for (uint64_t k = (uint64_t)(sb_ - sa_) + 1; k < slots; ++k) {
        hipError_t e = hipMemcpyAsync(a.slot_tiles[set] + ((uint64_t)a.gen[set] * slots + k) * tiles,
                                      a.slot_tiles[set] + ((uint64_t)(1 - a.gen[set]) * slots + k) * tiles, tiles,
                                      hipMemcpyDeviceToDevice, c->sa);
        if (e != hipSuccess) return hip_fail(c, e, "adjoint geometry");
    }
    // the general list of each step (k - 1, k), used by its recompute and its
    // reverse step
    hipError_t e = launch_fill_u32(a.slot_count[set], 0u, (uint32_t)(sb_ - sa_ + 1), c->sa);
    for (int64_t s = sa_ + 1; s <= sb_ && e == hipSuccess; ++s) {
        const uint32_t k = (uint32_t)(s - sa_);
        const SlotGeometry gp = slot_geometry(c, k - 1, set), gc = slot_geometry(c, k, set);
        e = launch_build_general_list(c->static_flags, gp.dyn, gc.dyn, c->g.tiles,
                                      a.slot_list[set] + (uint64_t)k * c->g.tiles, a.slot_count[set] + k, c->sa);
    }
    if (e == hipSuccess) e = hipEventRecord(a.ev_geo[set], c->sa);
    if (e != hipSuccess) return hip_fail(c, e, "adjoint geometry");
    return BF_OK;
}



}  // namespace

bf_status bf_adjoint_set_tangents(bf_context* c, uint32_t parameter_count, const float* point_tangents) {
    BF_ENTER(c);
    if (parameter_count > 8) return fail(c, BF_ERROR_INVALID_ARGUMENT, "at most 8 parameters");
    if (parameter_count > 0 && !point_tangents) return fail(c, BF_ERROR_INVALID_ARGUMENT, "tangents are NULL");
    if (parameter_count > 0 && (c->source != BF_GEOMETRY_MESH || c->m_articulated))
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "tangents describe mesh frames: upload a mesh as frames first");
    c->adj.tangent_alloc.release();
    c->adj.dpoints = nullptr;
    c->adj.K = 0;
    c->adj.kinematic = false;
    if (parameter_count == 0) return BF_OK;
    const uint64_t count = 3ull * parameter_count * c->m_frames * c->m_vertices;
    for (uint64_t i = 0; i < count; ++i)
        if (!std::isfinite(point_tangents[i])) return fail(c, BF_ERROR_INVALID_ARGUMENT, "non-finite tangent");
    hipError_t e = hipSuccess;
    c->adj.dpoints = (float*)c->adj.tangent_alloc.take(e, 4 * count);
    if (e == hipSuccess) e = hipMemcpy(c->adj.dpoints, point_tangents, 4 * count, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(c, e, "tangent upload", false);
    c->adj.K = parameter_count;
    return BF_OK;
}

bf_status bf_adjoint_set_kinematic_parameters(bf_context* c, uint32_t parameter_count,
                                              const bf_kinematic_parameter* parameters) {
    BF_ENTER(c);
    if (parameter_count > 8) return fail(c, BF_ERROR_INVALID_ARGUMENT, "at most 8 parameters");
    if (parameter_count > 0 && !parameters) return fail(c, BF_ERROR_INVALID_ARGUMENT, "parameters are NULL");
    if (parameter_count > 0 && (c->source != BF_GEOMETRY_MESH || !c->m_articulated))
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "kinematic parameters need an articulated mesh");
    for (uint32_t k = 0; k < parameter_count; ++k) {
        const bf_kinematic_parameter& q = parameters[k];
        if (q.joint_mask == 0 || (c->m_joint_count < 32 && (q.joint_mask >> c->m_joint_count) != 0))
            return fail(c, BF_ERROR_INVALID_ARGUMENT, "joint_mask names no joint or a joint the mesh lacks");
        if (q.channel > 3 || q.component > 1 || q.harmonic >= c->m_harmonics)
            return fail(c, BF_ERROR_INVALID_ARGUMENT, "channel < 4, component < 2 and harmonic < harmonics");
    }
    c->adj.tangent_alloc.release();
    c->adj.dpoints = nullptr;
    c->adj.K = parameter_count;
    c->adj.kinematic = parameter_count > 0;
    c->adj.kin = KinParams{};
    for (uint32_t k = 0; k < parameter_count; ++k)
        c->adj.kin.p[k] = {parameters[k].joint_mask, parameters[k].channel, parameters[k].harmonic,
                           parameters[k].component};
    return BF_OK;
}

bf_status bf_adjoint_gradient(bf_context* c, uint32_t steps, const double* load_weights, uint32_t checkpoint_interval,
                              double* objective, double* gradient, double timings_ms[2]) {
    BF_ENTER(c);
    ++c->f_writes;
    if (steps == 0 || !load_weights || !objective || checkpoint_interval == 0)
        return fail(c, BF_ERROR_INVALID_ARGUMENT, "steps, load weights, objective and interval are required");
    if (c->adj.K > 0 && !gradient) return fail(c, BF_ERROR_INVALID_ARGUMENT, "gradient output is NULL");
    if (steps > HISTORY_CAPACITY) return fail(c, BF_ERROR_INVALID_ADVANCE_REQUEST, "at most 65536 steps per sweep");
    if (bf_status r = adjoint_supported(c)) return r;
    auto& a = c->adj;
    const uint32_t interval = std::min(checkpoint_interval, steps);
    const uint32_t segments = (steps + interval - 1) / interval;
    if (bf_status r = adjoint_alloc(c, interval, segments)) return r;
    const uint64_t n = c->g.n, pop = 19ull * n, popS = pop * c->pop_elem;   // elements; snapshot bytes
    const int64_t s0 = c->step;
    hipEvent_t t0, t1, t2;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    (void)hipEventCreate(&t2);
    a.have_sensitivity = false;

    // forward pass: the ordinary pipelined steps, a snapshot every interval
    hipError_t e = hipEventRecord(t0, c->sb);
    for (uint32_t seg = 0; seg < segments && e == hipSuccess; ++seg) {
        e = hipMemcpyAsync(a.checkpoints + seg * popS, c->f[c->cur], popS, hipMemcpyDeviceToDevice, c->sb);
        if (e != hipSuccess) break;
        const uint32_t todo = std::min<uint32_t>(interval, steps - seg * interval);
        if (bf_status r = bf_step(c, todo)) return r;
        e = drain(c);
    }
    if (e == hipSuccess) e = hipEventRecord(t1, c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "adjoint forward pass");
    std::vector<double> hist(8ull * steps);
    uint32_t got = 0;
    if (bf_status r = bf_read_load_history(c, hist.data(), steps, &got)) return r;
    double J = 0.0;
    for (uint32_t i = 0; i < got; ++i)
        for (int k = 0; k < 6; ++k) J += load_weights[6 * i + k] * hist[8 * i + k];
    *objective = J;

    // reverse sweep
    const int storage = (int)c->cfg.storage, coll = (int)c->cfg.collision, arith = arithmetic_of(c->cfg);
    const bool ray = c->cfg.wall_rule == BF_WALL_INTERPOLATED_RAY;
    StepParams pf = make_step_params(c);
    pf.accumulate_loads = 0;     // the recompute needs populations only
    const StepParams pa = make_step_params(c);
    a.lam_cur = 0;
    e = launch_fill_u32((uint32_t*)a.lam[0], 0u, (uint32_t)pop, c->sb);    // dJ/df at the last step: 0
    if (e == hipSuccess) e = launch_fill_u32((uint32_t*)a.grad_total, 0u, 16u, c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "adjoint init");
    // Segments in reverse order on stream B; the geometry of the segment
    // after the current one is built on stream A meanwhile, into the other
    // set (released by stream B once the segment that last used it is done).
    if (bf_status r = segment_geometry(c, s0, steps, interval, (int64_t)segments - 1)) return r;
    uint64_t tangent_steps = 0;
    for (int64_t seg = (int64_t)segments - 1; seg >= 0; --seg) {
        const int64_t sa_ = s0 + seg * interval;
        const int64_t sb_ = std::min<int64_t>(sa_ + interval, s0 + steps);
        const int set = (int)(seg & 1);
        e = hipStreamWaitEvent(c->sb, a.ev_geo[set], 0);
        if (e == hipSuccess && seg > 0) {
            if (seg + 1 < (int64_t)segments) e = hipStreamWaitEvent(c->sa, a.ev_done[1 - set], 0);
            if (e != hipSuccess) return hip_fail(c, e, "adjoint schedule");
            if (bf_status r = segment_geometry(c, s0, steps, interval, seg - 1)) return r;
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(a.seg_f, a.checkpoints + seg * popS, popS, hipMemcpyDeviceToDevice, c->sb);
        for (int64_t s = sa_ + 1; s < sb_ && e == hipSuccess; ++s) {
            const SlotGeometry gp = slot_geometry(c, (uint32_t)(s - 1 - sa_), set),
                               gc = slot_geometry(c, (uint32_t)(s - sa_), set);
            FluidBuffers fb{};
            fb.fin = a.seg_f + (s - 1 - sa_) * popS;
            fb.fout = a.seg_f + (s - sa_) * popS;
            fb.solid_prev = gp.mask;
            fb.solid_cur = gc.mask;
            fb.wall = gc.wall;
            fb.static_flags = c->static_flags;
            fb.dyn_prev = gp.dyn;
            fb.dyn_cur = gc.dyn;
            fb.general_list = a.slot_list[set] + (uint64_t)(s - sa_) * c->g.tiles;
            fb.general_count = a.slot_count[set] + (s - sa_);
            fb.partials = c->partials;
            fb.density = c->density;
            fb.velocity = c->velocity;
            fb.activation = &c->counters[8];
            fb.body = c->body;
            StepParams ps = pf;
            if (ray) {
                // the exact-root table of step s, as the forward step built it
                MeshBuffers mb = mesh_buffers(c);
                mb.pos = a.rpos;
                mb.vel = a.rvel;
                e = launch_mesh_prepare(mb, mesh_params(c, time_of(c, s), false), c->sb);
                if (e == hipSuccess) e = launch_ray_links(ray_buffers(c, gc, a.rpos, c->ray_list, c->ray_count,
                                                                      c->ray_dirs, c->ray_links[s & 1], nullptr),
                                                          ray_params(c), c->sb);
                ps.links = c->ray_links[s & 1];
            }
            if (e == hipSuccess) e = launch_fluid_fast(arith, storage, coll, fb, ps, c->sb);
            if (e == hipSuccess) e = launch_fluid_general(arith, storage, coll, fb, ps, c->general_blocks, c->sb);
        }
        if (e != hipSuccess) return hip_fail(c, e, "adjoint recompute");
        for (int64_t s = sb_; s > sa_; --s) {
            const uint32_t k = (uint32_t)(s - sa_);
            const SlotGeometry gp = slot_geometry(c, k - 1, set), gc = slot_geometry(c, k, set);
            const int j = (int)(s & 1);
            if (a.K > 0 || ray) {
                // Stream C, up to two steps ahead: vertex positions and
                // velocities of step s, the exact-root table and its entry
                // triangles, then the wall tangents into dW[s & 1], once the
                // step that last read these parity-j buffers (s + 2) is done.
                if (s == sb_) e = hipStreamWaitEvent(a.sc, a.ev_geo[set], 0);
                if (e == hipSuccess && tangent_steps >= 2) e = hipStreamWaitEvent(a.sc, a.ev_used[j], 0);
                MeshBuffers mb = mesh_buffers(c);
                mb.pos = a.vpos[j];
                mb.vel = a.vvel[j];
                const MeshParams mp = mesh_params(c, time_of(c, s), false);
                if (e == hipSuccess) e = launch_mesh_prepare(mb, mp, a.sc);
                if (e == hipSuccess && ray)
                    e = launch_ray_links(ray_buffers(c, gc, a.vpos[j], a.rlist, a.rcount, a.rdirs, a.rtab[j], a.rwin[j]),
                                         ray_params(c), a.sc);
                if (e == hipSuccess && a.K > 0)
                    e = launch_wall_tangent(a.slot_list[set] + (uint64_t)k * c->g.tiles, a.slot_count[set] + k,
                                            c->general_blocks, gc.tri, c->m_tri, a.vpos[j], a.vvel[j], mb, a.dpoints,
                                            a.kin, a.dpos[j], a.dvel[j], a.dW[j], mp, (int)a.K, a.sc);
                if (e == hipSuccess) e = hipEventRecord(a.ev_tan[j], a.sc);
                if (e == hipSuccess) e = hipStreamWaitEvent(c->sb, a.ev_tan[j], 0);
                if (e != hipSuccess) return hip_fail(c, e, "wall tangents");
                ++tangent_steps;
            }
            AdjointBuffers ab{};
            ab.f_prev = a.seg_f + (k - 1) * popS;
            ab.lam_in = a.lam[a.lam_cur];
            ab.lam_out = a.lam[1 - a.lam_cur];
            ab.solid_prev = gp.mask;
            ab.solid_cur = gc.mask;
            ab.wall = gc.wall;
            ab.static_flags = c->static_flags;
            ab.dyn_prev = gp.dyn;
            ab.dyn_cur = gc.dyn;
            ab.general_list = a.slot_list[set] + (uint64_t)k * c->g.tiles;
            ab.general_count = a.slot_count[set] + k;
            ab.nu = a.nu;
            ab.dW = a.dW[j];
            ab.grad_partials = a.grad_partials;
            ab.grad_total = a.grad_total;
            ab.body = c->body;
            ab.outlet_extra = a.outlet_extra;
            ab.links = a.rtab[j];
            ab.winners = a.rwin[j];
            ab.vpos = a.vpos[j];
            ab.dvpos = a.dpos[j];
            ab.tri = c->m_tri;
            ab.vertex_count = c->m_vertices;
            ab.half_thickness = c->m_half;
            AdjointParams w{};
            const double* lw = load_weights + 6 * (size_t)(s - s0 - 1);
            for (int i = 0; i < 3; ++i) {
                w.cF[i] = (float)lw[i];
                w.cT[i] = (float)lw[3 + i];
            }
            w.K = (int)a.K;
            e = launch_adjoint_step(coll, storage, ab, pa, w, c->general_blocks, c->sb);
            if (e == hipSuccess && (a.K > 0 || ray)) e = hipEventRecord(a.ev_used[j], c->sb);
            if (e != hipSuccess) return hip_fail(c, e, "adjoint step");
            a.lam_cur = 1 - a.lam_cur;
        }
        e = hipEventRecord(a.ev_done[set], c->sb);
        if (e != hipSuccess) return hip_fail(c, e, "adjoint schedule");
    }
    e = hipEventRecord(t2, c->sb);
    double g8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(g8, a.grad_total, 64, hipMemcpyDeviceToHost, c->sb);
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "adjoint result");
    for (uint32_t k = 0; k < a.K; ++k) gradient[k] = g8[k];
    if (timings_ms) {
        float f_ms = 0.0f, r_ms = 0.0f;
        (void)hipEventElapsedTime(&f_ms, t0, t1);
        (void)hipEventElapsedTime(&r_ms, t1, t2);
        timings_ms[0] = f_ms;
        timings_ms[1] = r_ms;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipEventDestroy(t2);
    a.have_sensitivity = true;
    ++c->f_writes;   // the reverse sweep leaves no step's inputs in place
    return BF_OK;
}

bf_status bf_mesh_freeze_masks(bf_context* c, const uint8_t* masks, uint32_t count) {
    BF_ENTER(c);
    c->frozen_alloc.release();
    c->frozen = nullptr;
    c->frozen_count = 0;
    if (count == 0) return BF_OK;
    if (!masks) return fail(c, BF_ERROR_INVALID_ARGUMENT, "masks are NULL");
    if (c->source != BF_GEOMETRY_MESH) return fail(c, BF_ERROR_INVALID_ARGUMENT, "mask replay needs a mesh");
    hipError_t e = hipSuccess;
    c->frozen = (uint8_t*)c->frozen_alloc.take(e, (uint64_t)count * c->g.n);
    if (e == hipSuccess) e = hipMemcpy(c->frozen, masks, (uint64_t)count * c->g.n, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(c, e, "mask replay upload", false);
    c->frozen_count = count;
    c->frozen_first = c->step;
    // the current step's geometry (and its copy as the previous step's)
    e = replay_mask(c, c->step, 0, mod3(c->step));
    if (e == hipSuccess) e = replay_mask(c, c->step - 1, 0, mod3(c->step - 1));
    if (e == hipSuccess && c->geometry_ready > c->step && count > 1)
        e = replay_mask(c, c->step + 1, 1, mod3(c->step + 1));
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "mask replay");
    return BF_OK;
}

bf_status bf_adjoint_read_sensitivity(bf_context* c, float* out, uint64_t count) {
    BF_ENTER(c);
    if (!out || count != 19ull * c->g.n) return fail(c, BF_ERROR_INVALID_ARGUMENT, "count must be 19 * cells");
    if (!c->adj.have_sensitivity) return fail(c, BF_ERROR_INVALID_ARGUMENT, "no adjoint sweep has run");
    // lambda lives in the row-interleaved layout of an FP32 lattice: unpack
    // it to the host's direction-major order through the staging buffer
    if (bf_status st = ensure_staging(c)) return st;
    hipError_t e = launch_convert_populations(STORE_FP32, false, c->adj.lam[c->adj.lam_cur], c->staging, count, c->g,
                                              c->sb);
    if (e == hipSuccess) e = hipMemcpyAsync(out, c->staging, 4 * count, hipMemcpyDeviceToHost, c->sb);
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "read sensitivity");
    return BF_OK;
}

bf_status bf_read_loads(bf_context* c, double out[8]) {
    BF_ENTER(c);
    if (!out) return BF_ERROR_INVALID_ARGUMENT;
    hipError_t e = hipMemcpyAsync(out, c->last_load, 64, hipMemcpyDeviceToHost, c->sb);
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "read loads");
    return BF_OK;
}

bf_status bf_read_load_history(bf_context* c, double* out, uint32_t max_records, uint32_t* records) {
    BF_ENTER(c);
    if (!out || !records) return BF_ERROR_INVALID_ARGUMENT;
    const uint64_t avail = std::min<uint64_t>(c->history_count, HISTORY_CAPACITY);
    const uint64_t want = std::min<uint64_t>(avail, max_records);
    std::vector<double> all(8ull * HISTORY_CAPACITY);
    hipError_t e = hipMemcpyAsync(all.data(), c->history, 64ull * HISTORY_CAPACITY, hipMemcpyDeviceToHost, c->sb);
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "read load history");
    for (uint64_t i = 0; i < want; ++i) {
        const uint64_t step = (uint64_t)c->step - want + 1 + i;
        std::memcpy(out + 8 * i, all.data() + 8 * (step % HISTORY_CAPACITY), 64);
    }
    *records = (uint32_t)want;
    return BF_OK;
}

bf_status bf_link_census(bf_context* c, uint64_t counts[21]) {
    BF_ENTER(c);
    if (!counts) return BF_ERROR_INVALID_ARGUMENT;
    const StepGeometry sg = geometry_of(c, c->step);
    const uint8_t* prev = sg.mask_prev;
    if (c->source == BF_GEOMETRY_STATIC) prev = c->static_prev_differs ? c->mask[1] : c->mask[0];
    int bc[3][2];
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < 2; ++s) bc[a][s] = (int)c->cfg.boundary[a][s];
    hipError_t e = launch_fill_u8((uint8_t*)c->census, 0, 8 * 21, c->sb);
    if (e == hipSuccess) e = launch_link_census(prev, sg.mask_cur, c->g, bc, c->census, c->sb);
    if (e == hipSuccess) e = hipMemcpyAsync(counts, c->census, 8 * 21, hipMemcpyDeviceToHost, c->sb);
    if (e == hipSuccess) e = hipStreamSynchronize(c->sb);
    if (e != hipSuccess) return hip_fail(c, e, "link census");
    return BF_OK;
}

bf_status bf_set_profiling(bf_context* c, int enabled) {
    BF_ENTER(c);
    if (enabled < 0 || enabled > 2) return fail(c, BF_ERROR_INVALID_ARGUMENT, "profiling level is 0, 1 or 2");
    c->profiling = enabled != 0;
    c->profiling_level = enabled;
    c->profiled = 0;
    return BF_OK;
}

bf_status bf_read_profile(bf_context* c, double out[6]) {
    BF_ENTER(c);
    if (!out) return BF_ERROR_INVALID_ARGUMENT;
    double geo = 0, fast = 0, general = 0;
    for (uint32_t i = 0; i < c->profiled; ++i) {
        float a = 0, b = 0, d = 0;
        const hipEvent_t* ev = &c->events[(size_t)i * PROFILE_EVENTS];
        (void)hipEventElapsedTime(&a, ev[0], ev[1]);   // fast kernel (stream A)
        if (c->profiling_level >= 2) {
            (void)hipEventElapsedTime(&d, ev[2], ev[3]);   // next step's geometry + tile list (stream B)
            (void)hipEventElapsedTime(&b, ev[4], ev[5]);   // general kernel (stream A, before fast)
        }
        fast += a;
        general += b;
        geo += d;
    }
    out[0] = geo;
    out[1] = fast;
    out[2] = general;
    out[3] = (double)c->profiled;
    uint32_t general_tiles = 0;
    hipError_t e = hipMemcpy(&general_tiles, &c->counters[par2(c->step)], 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(c, e, "read profile");
    out[4] = (double)general_tiles;
    out[5] = (double)c->g.tiles;
    c->profiled = 0;
    return BF_OK;
}

}  // extern "C"
