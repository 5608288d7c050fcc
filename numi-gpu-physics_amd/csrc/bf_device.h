// Device-side D3Q19 algebra shared by the fluid, geometry and reduction
// kernels.  gfx950 only; compiled with -ffp-contract=off so that every
// expression rounds exactly as written.
//
// Operation order follows Sources/BirdFlowMetal/Metal/BirdFlow.metal:
//   equilibrium            :478-484
//   sponge factor          :2427-2446
//   closest point          :1529-1583
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bf {

// D3Q19 velocity set, ordering 0; +-x; +-y; +-z; (+-1,+-1,0); (+-1,0,+-1);
// (0,+-1,+-1) (BirdFlow.metal:7-44).  Kept as constexpr functions so that
// fully unrolled loops see compile-time constants.
__host__ __device__ constexpr int cx(int q) {
    constexpr int v[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    return v[q];
}
__host__ __device__ constexpr int cy(int q) {
    constexpr int v[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
    return v[q];
}
__host__ __device__ constexpr int cz(int q) {
    constexpr int v[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
    return v[q];
}
__host__ __device__ constexpr int opp(int q) {
    constexpr int v[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
    return v[q];
}

template <typename R>
__host__ __device__ inline R weight(int q) {
    return q == 0 ? R(1) / R(3) : (q <= 6 ? R(1) / R(18) : R(1) / R(36));
}

template <typename R>
__host__ __device__ inline R cs2() { return R(1) / R(3); }

// Second-order equilibrium, W rho (1 + 3cu + 4.5cu^2 - 1.5u^2).
template <typename R>
__host__ __device__ inline R equilibrium(int q, R rho, R ux, R uy, R uz) {
    R cu = R(cx(q)) * ux + R(cy(q)) * uy + R(cz(q)) * uz;
    R u2 = ux * ux + uy * uy + uz * uz;
    return weight<R>(q) * rho * (R(1) + R(3) * cu + R(4.5) * cu * cu - R(1.5) * u2);
}

// Boundary treatment per face.
enum : int { BC_PERIODIC = 0, BC_FARFIELD = 1, BC_OUTLET = 2 };
enum : int { COLL_TRT = 0, COLL_TRT_SYMMETRIC_LIMITER = 1, COLL_REGULARIZED = 2, COLL_RR3 = 3 };
enum : int { WALL_HALFWAY = 0, WALL_INTERPOLATED_SDF = 1, WALL_INTERPOLATED_HALF = 2, WALL_INTERPOLATED_LINKS = 3 };
enum : int { STORE_FP32 = 0, STORE_FP16_OFFSET = 1, STORE_BF16_OFFSET = 2 };
enum : int { ARITH_FUSED = 0, ARITH_STRICT = 1 };

// Tile of cells owned by one 256-lane workgroup: 64 consecutive x (one
// wavefront, 256 contiguous bytes per population row) by 4 rows of y.
constexpr int TILE_X = 64;
constexpr int TILE_Y = 4;
constexpr int TILE_CELLS = TILE_X * TILE_Y;

struct Grid {
    int nx, ny, nz;
    int tiles_x, tiles_y;      // tiles per row / per plane; tiles_z == nz
    uint32_t n;                // cells (fits 32 bits for every supported grid)
    uint32_t tiles;            // tiles_x * tiles_y * nz
};

// Physical and numerical parameters of one fluid step (passed by value).
struct StepParams {
    Grid g;
    int bc[3][2];
    int wall_rule;
    int force_mode;
    int load_component;
    int selected_part;
    int accumulate_loads;
    int capture_fields;
    int fast_priority;         // s_setprio level of the fast kernel's wavefronts (0..3)
    int pair_cells;            // 2-byte storage: fast tiles two cells per lane (fluid_fast_pair)
    int general_priority;      // and of the general kernel's
    float omega_plus, omega_minus;
    float sponge_strength, sponge_width;
    float far_u[3], far_rho;
    float origin[3], dx;
    float force_scale;
    float far_eq[19];          // equilibrium(q, far_rho, far_u), FP32
    const float* links;        // WALL_INTERPOLATED_LINKS: fractions [tile][19][256] of fluid cells' links
    // non-finite population guard (optional): [0] flags (1 = a cell's
    // post-collision populations did not sum to a finite density), [1] the
    // first step that saw one; step_index = this step
    unsigned long long* health;
    unsigned long long step_index;
};

// A cell whose stored populations do not sum to a finite density (NaN or
// infinity among them) flags the step: sticky, the earliest step kept.
// Per-lane atomics on the rare bad lane only.
__device__ __forceinline__ void guard_finite(const StepParams& p, float density) {
    if (p.health && !isfinite(density)) {
        atomicOr(&p.health[0], 1ull);
        atomicMin(&p.health[1], p.step_index);
    }
}

// XCD-aware block order: the hardware deals workgroups round-robin over the
// 8 XCDs; remap so each XCD walks one contiguous run of tiles (its L2 then
// holds the cache lines that neighbouring tiles share).
__host__ __device__ inline uint32_t xcd_remap(uint32_t block, uint32_t blocks) {
    constexpr uint32_t XCDS = 8;
    uint32_t xcd = block % XCDS, local = block / XCDS;
    uint32_t per = blocks / XCDS, rem = blocks % XCDS;
    return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

__host__ __device__ inline void tile_origin(const Grid& g, uint32_t tile, int& x0, int& y0, int& z) {
    uint32_t per_plane = (uint32_t)g.tiles_x * (uint32_t)g.tiles_y;
    z = (int)(tile / per_plane);
    uint32_t r = tile - (uint32_t)z * per_plane;
    int ty = (int)(r / (uint32_t)g.tiles_x);
    int tx = (int)(r - (uint32_t)ty * (uint32_t)g.tiles_x);
    x0 = tx * TILE_X;
    y0 = ty * TILE_Y;
}

__host__ __device__ inline uint32_t tile_of(const Grid& g, int x, int y, int z) {
    return ((uint32_t)z * (uint32_t)g.tiles_y + (uint32_t)(y / TILE_Y)) * (uint32_t)g.tiles_x
        + (uint32_t)(x / TILE_X);
}

__device__ inline float sponge_factor(const StepParams& p, int x, int y, int z) {
    int dxm = min(x, p.g.nx - 1 - x);
    int dym = min(y, p.g.ny - 1 - y);
    int dzm = min(z, p.g.nz - 1 - z);
    float distance = (float)min(dxm, min(dym, dzm));
    float width = p.sponge_width;
    if (distance >= width) return 0.0f;
    float n = (width - distance) / fmaxf(width, 1.0f);
    n = fminf(fmaxf(n, 0.0f), 1.0f);
    return p.sponge_strength * n * n;
}

// Population storage: FP32 values, or FP16 / BF16 offsets f - w_q with the
// arithmetic in FP32.
template <int S>
struct Store;

// Position of direction q of cell (x, row), row = y + ny z, in a population
// buffer: the 19 directions interleaved row by row ([z][y][q][x]).  The 19
// loads and stores of a wavefront then fall in one stretch of 19 rows
// instead of 19 planes a lattice apart; on MI355X the fast kernel streams 7 %
// (FP32) and 9 % (FP16) faster than with direction-major planes.  The
// host-facing layout is direction-major (the pack / unpack kernels convert).
template <int S>
__host__ __device__ __forceinline__ size_t pop_index(const Grid& g, int q, int x, uint32_t row) {
    return ((size_t)row * 19u + (size_t)q) * (size_t)g.nx + (size_t)x;
}

template <>
struct Store<STORE_FP32> {
    using T = float;
    __device__ static inline float load(const T* __restrict__ p, size_t i, int) { return p[i]; }
    __device__ static inline void store(T* __restrict__ p, size_t i, int, float v) {
        __builtin_nontemporal_store(v, p + i);
    }
};

template <>
struct Store<STORE_FP16_OFFSET> {
    using T = _Float16;
    __device__ static inline float load(const T* __restrict__ p, size_t i, int q) {
        return (float)p[i] + weight<float>(q);
    }
    __device__ static inline void store(T* __restrict__ p, size_t i, int q, float v) {
        __builtin_nontemporal_store((T)(v - weight<float>(q)), p + i);
    }
};

template <>
struct Store<STORE_BF16_OFFSET> {
    using T = uint16_t;  // bfloat16 bits
    __device__ static inline float to_float(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
    // round to nearest even in hardware (v_cvt_pk_bf16_f32 on gfx950)
    __device__ static inline uint16_t from_float(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
    __device__ static inline float load(const T* __restrict__ p, size_t i, int q) {
        return to_float(p[i]) + weight<float>(q);
    }
    __device__ static inline void store(T* __restrict__ p, size_t i, int q, float v) {
        __builtin_nontemporal_store(from_float(v - weight<float>(q)), p + i);
    }
};

}  // namespace bf
