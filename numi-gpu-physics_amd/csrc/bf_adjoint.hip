// Hand-written adjoint of the moving-boundary step for gfx950.
//
// For J = sum_s c_s . L_s (L_s: force and torque of step s), the reverse
// sweep carries lambda_s = dJ/df_s back through each step with the masks of
// the step fixed, and accumulates dJ/dtheta through the wall field of every
// step (link fractions from the signed distances and wall velocities), the
// wall field's tangents d wall / d theta coming from the mesh tangents.
//
// The forward step pulls: g_q(x) = f_q(x - c_q), or the bounce-back value of
// a wall link, or the far-field equilibrium; then collides.  Its adjoint is
// split so that every write has exactly one writer (no atomics, results are
// deterministic):
//   adjoint_fast / adjoint_general_p1  each cell x computes
//        mu = (d post / d g)^T lambda_s(x)  and writes lambda_{s-1,q}(x - c_q),
//        whose only puller is x (zero where that pull is a wall link or x is
//        solid or refilled; entries nobody pulls are zeroed by their owner);
//   adjoint_general_p2a  adds what a cell's own populations feed besides the
//        regular pull: reflected and x-side populations of its links, the
//        cover / uncover impulses; and the wall-field gradient, contracted
//        with the tangents into per-tile FP64 partials;
//   adjoint_general_p2b  adds the far-side populations x + c_q of
//        interpolated links and of uncover impulses.
// Targets of p2a are the cell itself; a target of p2b has one contributor.
//
// Outlet (zero-gradient) faces on x: a cell of the face plane pulls the
// populations that would cross the face from its own plane (the clamp), so
// those targets have a second puller beside the regular one.  Its share goes
// to a face buffer indexed by the puller (one writer), and adjoint_outlet
// adds it to the target after the single-writer passes; the populations that
// leave through the face are pulled by nobody and are zeroed by their owner,
// as behind a far-field face.
//
// Supported: FP32 storage and FP16 / BF16 offsets (decoded on load; lambda
// stays FP32), TRT (with or without the symmetric limiter) / regularized /
// RR3, periodic and far-field faces and outlets on the x faces, sponge, every
// wall rule, force mode 6.
#include "bf_adjoint_math.h"
#include "bf_fluid.h"
#include "bf_rays.h"

namespace bf {

constexpr uint32_t NO_TRIANGLE = 0xFFFFFFFFu;

__device__ __forceinline__ void cross3f(const float a[3], const float b[3], float o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// c_q . (cF + cT x arm): d J / d (link or impulse momentum along c_q) per
// unit of force scale
template <int q>
__device__ __forceinline__ float load_weight(const AdjointParams& w, const float arm[3]) {
    float t[3];
    cross3f(w.cT, arm, t);
    return (float)cx(q) * (w.cF[0] + t[0]) + (float)cy(q) * (w.cF[1] + t[1]) + (float)cz(q) * (w.cF[2] + t[2]);
}

template <int COLL>
__device__ __forceinline__ void collision_adjoint(const float (&g)[19], float op, float om, const float (&lp)[19],
                                                  float (&mu)[19]) {
    adj::reverse<COLL, float>(g, op, om, lp, mu);
}

// Outlet x faces (BC_OUTLET): the pull of direction q from cell x is
// clamped into x's own plane, or (x, q) leaves through the face.
template <int q>
__device__ __forceinline__ int clamped_face(const StepParams& p, int x) {
    if (cx(q) < 0 && x == p.g.nx - 1 && p.bc[0][1] == BC_OUTLET) return 1;
    if (cx(q) > 0 && x == 0 && p.bc[0][0] == BC_OUTLET) return 0;
    return -1;
}
template <int q>
__device__ __forceinline__ bool leaves_outlet(const StepParams& p, int x) {
    return (cx(q) > 0 && x == p.g.nx - 1 && p.bc[0][1] == BC_OUTLET) ||
           (cx(q) < 0 && x == 0 && p.bc[0][0] == BC_OUTLET);
}
__device__ __forceinline__ size_t outlet_slot(const Grid& g, int face, int q, int y, int z) {
    return ((size_t)face * 19 + (size_t)q) * ((size_t)g.ny * (size_t)g.nz) + (size_t)y + (size_t)g.ny * (size_t)z;
}
// Adjoint populations use the forward lattice's row-interleaved layout:
// lambda_q of cell (x, row = y + ny z) at pop_index(q, x, row).
__device__ __forceinline__ size_t lam_at(const Grid& g, int q, int x, uint32_t row) {
    return pop_index<STORE_FP32>(g, q, x, row);
}
// A population of the step's input lattice (index i, direction q), decoded
// from its stored form: FP32, or FP16 / BF16 offsets from w_q.  The adjoint
// differentiates the FP32 arithmetic and takes the storage rounding as the
// identity.
template <int S>
__device__ __forceinline__ float fprev(const AdjointBuffers& b, size_t i, int q) {
    return Store<S>::load((const typename Store<S>::T*)b.f_prev, i, q);
}
// Rule and table fraction of link (x, q) (its source solid): under the
// exact-root rule the table's entry where a triangle's shell gave it, else
// the signed-distance ratio (the table holds that ratio there, computed from
// the same walls as link_value computes it).
__device__ __forceinline__ int link_rule(const AdjointBuffers& b, const StepParams& p, int x, int y, int z, int q,
                                         float& tlf, uint32_t& win) {
    tlf = 0.0f;
    win = NO_WINNER;
    if (p.wall_rule != WALL_INTERPOLATED_LINKS) return p.wall_rule;
    const size_t slot = table_slot(p.g, x, y, z, q);
    win = b.winners[slot];
    tlf = b.links[slot];
    return win == NO_WINNER ? WALL_INTERPOLATED_SDF : WALL_INTERPOLATED_LINKS;
}
// lambda_{s-1,q}(source) = v from the (only regular) puller x, or into the
// outlet face buffer when x's pull along q is clamped; nothing for a
// far-field source.  (xs, rows): the source's x and row.
template <int q>
__device__ __forceinline__ void write_pull(const AdjointBuffers& b, const StepParams& p, int x, int y, int z,
                                           int xs, uint32_t rows, bool far, float v) {
    if (far) return;
    const int face = clamped_face<q>(p, x);
    if (face >= 0) b.outlet_extra[outlet_slot(p.g, face, q, y, z)] = v;
    else b.lam_out[lam_at(p.g, q, xs, rows)] = v;
}
// (x, q) pulled by nobody: it leaves through a far-field or outlet face
template <int q>
__device__ __forceinline__ bool unpulled(const Sources& s, const StepParams& p, int x) {
    return orphan<q>(s) || leaves_outlet<q>(p, x);
}

// ---------------------------------------------------------------------------
// Wall-free tiles.
// ---------------------------------------------------------------------------
// At least 4 wavefronts per SIMD (128 VGPRs): the RR3 instance spills 22
// registers to scratch there and still runs 3 % faster than at its natural
// 140 VGPRs and 3 wavefronts (193 against 200 us at 256 x 128 x 128).
template <int COLL, int S>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void adjoint_fast(AdjointBuffers b,
                                                                                           StepParams p) {
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    if ((b.static_flags[tile] | b.dyn_prev[tile] | b.dyn_cur[tile]) != 0) return;
    int x0, y0, z;
    tile_origin(p.g, tile, x0, y0, z);
    const int x = x0 + (int)(threadIdx.x & (TILE_X - 1));
    const int y = y0 + (int)(threadIdx.x / TILE_X);
    const uint32_t row = (uint32_t)y + (uint32_t)p.g.ny * (uint32_t)z;
    const Sources s = sources(p, x, y, z);
    // the sponge blend f_out = post + (f_far - post) s scales the adjoint by 1 - s
    const float keep = 1.0f - (p.sponge_strength > 0.0f ? sponge_factor(p, x, y, z) : 0.0f);
    float g[19], lp[19], mu[19];
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        int xs;
        uint32_t rows;
        bool far;
        source_xr<q>(s, p.g, x, y, z, xs, rows, far);
        // unconditional load (a far-field source reads this cell)
        const float v = fprev<S>(b, far ? pop_index<S>(p.g, q, x, row) : pop_index<S>(p.g, q, xs, rows), q);
        g[q] = far ? p.far_eq[q] : v;
        lp[q] = b.lam_in[lam_at(p.g, q, x, row)] * keep;
    });
    collision_adjoint<COLL>(g, p.omega_plus, p.omega_minus, lp, mu);
    // The sources again, from coordinates the compiler cannot match with the
    // ones above: otherwise it keeps the 19 load addresses (the store
    // addresses are the same numbers) live across the collision, 38 VGPRs.
    int xr = x, yr = y, zr = z;
    asm volatile("" : "+v"(xr), "+v"(yr), "+v"(zr));
    const Sources sr = sources(p, xr, yr, zr);
    const uint32_t rowr = (uint32_t)yr + (uint32_t)p.g.ny * (uint32_t)zr;
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        int xs;
        uint32_t rows;
        bool far;
        source_xr<q>(sr, p.g, xr, yr, zr, xs, rows, far);
        write_pull<q>(b, p, xr, yr, zr, xs, rows, far, mu[q]);
        if (unpulled<q>(sr, p, xr)) b.lam_out[lam_at(p.g, q, xr, rowr)] = 0.0f;
    });
}

// ---------------------------------------------------------------------------
// Wall tiles, pass 1: collision adjoint and the single-writer pull targets.
// nu (per listed tile, direction-major, 256 lanes): for a fluid cell the
// sensitivity of J to each wall-link value (collision part plus the link's
// own load), for a refilled cell mu.
// ---------------------------------------------------------------------------
template <int COLL, int S>
__device__ __forceinline__ void general_p1_cell(const AdjointBuffers& b, const StepParams& p,
                                                const AdjointParams& w, int x, int y, int z, float bx, float by,
                                                float bz, float* __restrict__ nu_cell) {
    const Grid& g = p.g;
    const uint32_t row = (uint32_t)y + (uint32_t)g.ny * (uint32_t)z;
    const uint32_t gid = (uint32_t)x + (uint32_t)g.nx * row;
    const bool is_solid = b.solid_cur[gid] != 0;
    const bool was_solid = b.solid_prev[gid] != 0;
    const Sources s = sources(p, x, y, z);
    uint32_t src[19];
    int sx[19];
    uint32_t srow[19];
    bool far[19];
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        source_xr<q>(s, g, x, y, z, sx[q], srow[q], far[q]);
        src[q] = (uint32_t)sx[q] + (uint32_t)g.nx * srow[q];
        if (unpulled<q>(s, p, x)) b.lam_out[lam_at(g, q, x, row)] = 0.0f;
    });
    if (is_solid) {
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            write_pull<q>(b, p, x, y, z, sx[q], srow[q], far[q], 0.0f);
            nu_cell[q * 256] = 0.0f;
        });
        return;
    }
    const float sponge = p.sponge_strength > 0.0f ? sponge_factor(p, x, y, z) : 0.0f;
    float lp[19], gq[19], mu[19];
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        lp[q] = b.lam_in[lam_at(g, q, x, row)] * (1.0f - sponge);
    });
    if (was_solid) {
        // refilled with feq(1, u_wall): no dependence on the pulled populations
        const float4 rv = b.wall[gid];
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            gq[q] = equilibrium<float>(q, 1.0f, rv.x, rv.y, rv.z);
        });
        collision_adjoint<COLL>(gq, p.omega_plus, p.omega_minus, lp, mu);
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            write_pull<q>(b, p, x, y, z, sx[q], srow[q], far[q], 0.0f);
            nu_cell[q * 256] = mu[q];
        });
        return;
    }
    const float4 wg4 = b.wall[gid];
    const float wgv[4] = {wg4.x, wg4.y, wg4.z, wg4.w};
    const float wx = p.origin[0] + ((float)x + 0.5f) * p.dx;
    const float wy = p.origin[1] + ((float)y + 0.5f) * p.dx;
    const float wz = p.origin[2] + ((float)z + 0.5f) * p.dx;
    bool link[19];
    float load_dv[19];
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        link[q] = false;
        load_dv[q] = 0.0f;
        if (far[q]) {
            gq[q] = p.far_eq[q];
            return;
        }
        const uint8_t part = b.solid_cur[src[q]];
        if (part == 0) {
            int xs;
            uint32_t rows;
            bool fq;
            source_xr<q>(s, g, x, y, z, xs, rows, fq);
            gq[q] = fprev<S>(b, pop_index<S>(g, q, xs, rows), q);
            return;
        }
        link[q] = true;
        const float4 ws4 = b.wall[src[q]];
        const float wsv[4] = {ws4.x, ws4.y, ws4.z, ws4.w};
        const int fx = x + cx(q), fy = y + cy(q), fz = z + cz(q);
        const bool fin_ok = inside(g, fx, fy, fz);
        const uint32_t fid = fin_ok ? flat(g, fx, fy, fz) : gid;
        const bool farther_ok = fin_ok && b.solid_cur[fid] == 0;
        const float refl = fprev<S>(b, pop_index<S>(g, opp(q), x, row), opp(q));
        const float farther =
            farther_ok ? fprev<S>(b, pop_index<S>(g, opp(q), fx, (uint32_t)fy + (uint32_t)g.ny * (uint32_t)fz), opp(q))
                       : 0.0f;
        const float pin = fprev<S>(b, pop_index<S>(g, q, x, row), q);
        float tlf;
        uint32_t win;
        const int rule = link_rule(b, p, x, y, z, q, tlf, win);
        const adj::LinkD<float> d = adj::link_value<float>(cx(q), cy(q), cz(q), weight<float>(q), p.far_rho, rule,
                                                           farther_ok, refl, farther, pin, wgv, wsv, tlf);
        gq[q] = d.value;
        const int sel = p.selected_part;
        if (p.accumulate_loads && p.load_component != 2 && (sel == 0 || sel == (int)part)) {
            const float arm[3] = {wx - d.lf * (float)cx(q) * p.dx - bx, wy - d.lf * (float)cy(q) * p.dx - by,
                                  wz - d.lf * (float)cz(q) * p.dx - bz};
            // J_link = -fscale (value + reflected) c_q . e(arm)
            load_dv[q] = -p.force_scale * load_weight<q>(w, arm);
        }
    });
    collision_adjoint<COLL>(gq, p.omega_plus, p.omega_minus, lp, mu);
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        write_pull<q>(b, p, x, y, z, sx[q], srow[q], far[q], link[q] ? 0.0f : mu[q]);
        nu_cell[q * 256] = link[q] ? mu[q] + load_dv[q] : 0.0f;
    });
}

template <int COLL, int S>
__global__ __launch_bounds__(256) void adjoint_general_p1(AdjointBuffers b, StepParams p, AdjointParams w) {
    const uint32_t count = *b.general_count;
    const float bx = (float)b.body[0], by = (float)b.body[1], bz = (float)b.body[2];
    for (uint32_t i = blockIdx.x; i < count; i += gridDim.x) {
        int x0, y0, z;
        tile_origin(p.g, b.general_list[i], x0, y0, z);
        const int x = x0 + (int)(threadIdx.x & (TILE_X - 1));
        const int y = y0 + (int)(threadIdx.x / TILE_X);
        if (x < p.g.nx && y < p.g.ny)
            general_p1_cell<COLL, S>(b, p, w, x, y, z, bx, by, bz, b.nu + (size_t)i * 19 * 256 + threadIdx.x);
    }
}

// Outlet x faces: adds each face cell's clamped pull (held in the face
// buffer by the puller) to its target in the same plane, after the regular
// pullers have set the targets.  One thread per (face cell, direction).
__global__ __launch_bounds__(256) void adjoint_outlet(AdjointBuffers b, StepParams p) {
    const Grid& g = p.g;
    const uint32_t plane = (uint32_t)g.ny * (uint32_t)g.nz;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2u * 19u * plane) return;
    const int face = (int)(i / (19u * plane));
    const uint32_t r = i - (uint32_t)face * 19u * plane;
    const int q = (int)(r / plane);
    const uint32_t yz = r - (uint32_t)q * plane;
    const int y = (int)(yz % (uint32_t)g.ny), z = (int)(yz / (uint32_t)g.ny);
    const int x = face == 1 ? g.nx - 1 : 0;
    if (p.bc[0][face] != BC_OUTLET || cx(q) != (face == 1 ? -1 : 1)) return;
    bool far = false;
    const int ys = shift_axis(y, cy(q), g.ny, p.bc[1][0], p.bc[1][1], far);
    const int zs = shift_axis(z, cz(q), g.nz, p.bc[2][0], p.bc[2][1], far);
    if (far) return;   // the pull was replaced by the far-field equilibrium
    b.lam_out[lam_at(g, q, x, (uint32_t)ys + (uint32_t)g.ny * (uint32_t)zs)] +=
        b.outlet_extra[outlet_slot(g, face, q, y, z)];
}

// Sum of 8 FP64 lanes over the workgroup (fixed tree).
__device__ __forceinline__ void block_sum8(double (&v)[8], double* __restrict__ out) {
    __shared__ double part[4][8];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] += __shfl_down(v[i], off, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < 8; ++i) part[wave][i] = v[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            double s = part[0][i];
            for (int k = 1; k < (int)(blockDim.x / 64); ++k) s += part[k][i];
            out[i] = s;
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void contract(double (&acc)[8], int K, const float4* __restrict__ dW, size_t n,
                                         uint32_t cell, float gx, float gy, float gz, float gw) {
    for (int k = 0; k < K; ++k) {
        const float4 t = dW[(size_t)k * n + cell];
        acc[k] += (double)gx * t.x + (double)gy * t.y + (double)gz * t.z + (double)gw * t.w;
    }
}

// d lf / d theta_k of an exact-root link (x, q) whose entry triangle t lies
// in the link (defined below, with the dual numbers).
struct RayDQ {
    float v[8];
};
__device__ __noinline__ RayDQ ray_link_dq(const AdjointBuffers& b, const StepParams& p, int x, int y, int z, int q,
                                          uint32_t t, int K);

// ---------------------------------------------------------------------------
// Wall tiles, pass 2a: the cell's own populations and the wall gradient.
// ---------------------------------------------------------------------------
// RAY instantiates the exact-root contraction; without it the kernel holds no
// call (a call site reserves a 504-byte stack per lane and halves occupancy).
template <int S, bool RAY>
__device__ __forceinline__ void general_p2a_cell(const AdjointBuffers& b, const StepParams& p,
                                                 const AdjointParams& w, int x, int y, int z, float bx, float by,
                                                 float bz, const float* __restrict__ nu_cell, double (&acc)[8]) {
    const Grid& g = p.g;
    const size_t n = g.n;
    const uint32_t row = (uint32_t)y + (uint32_t)g.ny * (uint32_t)z;
    const uint32_t gid = (uint32_t)x + (uint32_t)g.nx * row;
    const bool is_solid = b.solid_cur[gid] != 0;
    const bool was_solid = b.solid_prev[gid] != 0;
    const bool topology = p.accumulate_loads && p.load_component != 1 && p.force_mode == 6;
    const float wx = p.origin[0] + ((float)x + 0.5f) * p.dx;
    const float wy = p.origin[1] + ((float)y + 0.5f) * p.dx;
    const float wz = p.origin[2] + ((float)z + 0.5f) * p.dx;
    const float arm[3] = {wx - bx, wy - by, wz - bz};
    const float fs = p.force_scale;
    if (is_solid) {
        // f_s(x) = feq(1, u_wall(x))
        const float4 wv = b.wall[gid];
        float l[19];
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            l[q] = b.lam_in[lam_at(g, q, x, row)];
        });
        float gx, gy, gz;
        adj::feq_velocity_adjoint<float>(l, wv.x, wv.y, wv.z, gx, gy, gz);
        contract(acc, w.K, b.dW, n, gid, gx, gy, gz, 0.0f);
        if (topology && !was_solid) {
            // fresh cover: F += sum_q f_q c_q (the preserved momentum)
            for_q(AllQ{}, [&](auto qc) {
                constexpr int q = decltype(qc)::value;
                if constexpr (q != 0) b.lam_out[lam_at(g, q, x, row)] += fs * load_weight<q>(w, arm);
            });
        }
        return;
    }
    if (was_solid) {
        const float4 rv = b.wall[gid];
        float l[19];
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            l[q] = nu_cell[q * 256];
        });
        float gx, gy, gz;
        adj::feq_velocity_adjoint<float>(l, rv.x, rv.y, rv.z, gx, gy, gz);
        if (topology) {
            // fresh uncover: F -= P_refill (= u_wall) + sum over persistent
            // fluid neighbours of (f_q(x) + f_opp(x + c_q)) c_q
            float e[3];
            cross3f(w.cT, arm, e);
            gx -= fs * (w.cF[0] + e[0]);
            gy -= fs * (w.cF[1] + e[1]);
            gz -= fs * (w.cF[2] + e[2]);
            for_q(AllQ{}, [&](auto qc) {
                constexpr int q = decltype(qc)::value;
                if constexpr (q != 0) {
                    const int fx = x + cx(q), fy = y + cy(q), fz = z + cz(q);
                    if (inside(g, fx, fy, fz)) {
                        const uint32_t nb = flat(g, fx, fy, fz);
                        if (b.solid_prev[nb] == 0 && b.solid_cur[nb] == 0)
                            b.lam_out[lam_at(g, q, x, row)] += -fs * load_weight<q>(w, arm);
                    }
                }
            });
        }
        contract(acc, w.K, b.dW, n, gid, gx, gy, gz, 0.0f);
        return;
    }
    const Sources s = sources(p, x, y, z);
    const float4 wg4 = b.wall[gid];
    const float wgv[4] = {wg4.x, wg4.y, wg4.z, wg4.w};
    float gxw[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    // Load pass 1, unconditional (a far-field source or a far side outside the
    // domain reads this cell and is ignored): the part ids of every source
    // and far side, the cell's previous populations, its adjoint slots (kept
    // in registers and stored once at the end: only this cell writes them
    // in this pass) and nu.
    auto src_of = [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        bool far;
        const uint32_t src = source_of<q>(s, g, x, y, z, far);
        return far ? gid : src;
    };
    auto fid_of = [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const int fx = x + cx(q), fy = y + cy(q), fz = z + cz(q);
        return inside(g, fx, fy, fz) ? flat(g, fx, fy, fz) : gid;
    };
    uint32_t part[19], fpart[19];
    float fown[19], lam[19], nuv[19];
    uint32_t farbits = 0, okbits = 0;
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        bool far;
        (void)source_of<q>(s, g, x, y, z, far);
        farbits |= (uint32_t)far << q;
        okbits |= (uint32_t)inside(g, x + cx(q), y + cy(q), z + cz(q)) << q;
        part[q] = b.solid_cur[src_of(qc)];
        fpart[q] = b.solid_cur[fid_of(qc)];
        fown[q] = fprev<S>(b, pop_index<S>(g, q, x, row), q);
        lam[q] = b.lam_out[lam_at(g, q, x, row)];
        nuv[q] = nu_cell[q * 256];
    });
    asm volatile("" : BF_ASM19(part), BF_ASM19(fpart), BF_ASM19(fown));
    asm volatile("" : BF_ASM19(lam), BF_ASM19(nuv));
    // links, links whose far side is an in-domain fluid cell, links whose
    // part counts toward the loads
    uint32_t linkbits = 0, fartherbits = 0, selbits = 0;
    const int sel = p.selected_part;
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        linkbits |= (uint32_t)(part[q] != 0 && !((farbits >> q) & 1u)) << q;
        fartherbits |= (uint32_t)(((okbits >> q) & 1u) && fpart[q] == 0) << q;
        selbits |= (uint32_t)(sel == 0 || sel == (int)part[q]) << q;
    });
    // Load pass 2, three directions at a time: each link's source wall entry
    // and far-side population loaded together; then the wall-field tangents
    // of the group's link sources.
    auto link_group = [&](auto seq, auto first_c, uint32_t gmask) {
        constexpr int q0 = decltype(first_c)::value;
        float4 ws[3];
        float fo[3];
        uint32_t srcs[3];
        float gws[3][4];
        for (int j = 0; j < 3; ++j) {
            ws[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            fo[j] = 0.0f;
            srcs[j] = gid;
            for (int i = 0; i < 4; ++i) gws[j][i] = 0.0f;
        }
        if (!(linkbits & gmask)) return;
        for_q(seq, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const bool lk = (linkbits >> q) & 1u;
            const uint32_t src = lk ? src_of(qc) : gid;
            srcs[q - q0] = src;
            ws[q - q0] = b.wall[src];
            const int fx = x + cx(q), fy = y + cy(q), fz = z + cz(q);
            fo[q - q0] = fprev<S>(b, lk && inside(g, fx, fy, fz)
                                      ? pop_index<S>(g, opp(q), fx, (uint32_t)fy + (uint32_t)g.ny * (uint32_t)fz)
                                      : pop_index<S>(g, opp(q), x, row), opp(q));
        });
        asm volatile("" : "+v"(ws[0].x), "+v"(ws[0].y), "+v"(ws[0].z), "+v"(ws[0].w), "+v"(ws[1].x),
                     "+v"(ws[1].y), "+v"(ws[1].z), "+v"(ws[1].w), "+v"(ws[2].x), "+v"(ws[2].y), "+v"(ws[2].z),
                     "+v"(ws[2].w), "+v"(fo[0]), "+v"(fo[1]), "+v"(fo[2]));
        for_q(seq, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if (!((linkbits >> q) & 1u)) return;
            const float4 ws4 = ws[q - q0];
            const float wsv[4] = {ws4.x, ws4.y, ws4.z, ws4.w};
            const bool farther_ok = (fartherbits >> q) & 1u;
            const float refl = fown[opp(q)];
            const float farther = farther_ok ? fo[q - q0] : 0.0f;
            const float pin = fown[q];
            float tlf = 0.0f;
            uint32_t win = NO_WINNER;
            int rule = p.wall_rule;
            if constexpr (RAY)
                rule = link_rule(b, p, x, y, z, q, tlf, win);
            else
                __builtin_assume(rule != WALL_INTERPOLATED_LINKS);
            const adj::LinkD<float> d = adj::link_value<float>(cx(q), cy(q), cz(q), weight<float>(q), p.far_rho,
                                                               rule, farther_ok, refl, farther, pin, wgv, wsv, tlf);
            const float nu = nuv[q];
            // exact-root fraction of an unclamped entry: d lf / d theta from
            // the entry triangle's shell
            if constexpr (RAY) {
                if (rule == WALL_INTERPOLATED_LINKS && w.K > 0 && tlf > 1.0e-4f && tlf < 1.0f && d.dv_dlf != 0.0f) {
                    const RayDQ r = ray_link_dq(b, p, x, y, z, q, win, w.K);
                    const double sc = (double)(nu * d.dv_dlf);
                    for (int k = 0; k < w.K; ++k) acc[k] += sc * (double)r.v[k];
                }
            }
            float refl_load = 0.0f;
            if (p.accumulate_loads && p.load_component != 2 && ((selbits >> q) & 1u)) {
                const float larm[3] = {wx - d.lf * (float)cx(q) * p.dx - bx, wy - d.lf * (float)cy(q) * p.dx - by,
                                       wz - d.lf * (float)cz(q) * p.dx - bz};
                refl_load = -fs * load_weight<q>(w, larm);
            }
            lam[opp(q)] += nu * d.d_refl + refl_load;
            if (d.uses_pin) lam[q] += nu * d.d_pin;
            for (int i = 0; i < 4; ++i) gxw[i] += nu * d.d_wg[i];
            for (int i = 0; i < 4; ++i) gws[q - q0][i] = nu * d.d_ws[i];
        });
        // contract in ascending q per parameter (a non-link direction adds
        // zero times this cell's tangent)
        for (int k = 0; k < w.K; ++k) {
            float4 t[3];
            for (int j = 0; j < 3; ++j) t[j] = b.dW[(size_t)k * n + srcs[j]];
            for (int j = 0; j < 3; ++j)
                if ((linkbits >> (q0 + j)) & 1u)
                    acc[k] += (double)gws[j][0] * t[j].x + (double)gws[j][1] * t[j].y + (double)gws[j][2] * t[j].z +
                              (double)gws[j][3] * t[j].w;
        }
    };
    link_group(QSeq<1, 2, 3>{}, std::integral_constant<int, 1>{}, 0xEu);
    link_group(QSeq<4, 5, 6>{}, std::integral_constant<int, 4>{}, 0x70u);
    link_group(QSeq<7, 8, 9>{}, std::integral_constant<int, 7>{}, 0x380u);
    link_group(QSeq<10, 11, 12>{}, std::integral_constant<int, 10>{}, 0x1C00u);
    link_group(QSeq<13, 14, 15>{}, std::integral_constant<int, 13>{}, 0xE000u);
    link_group(QSeq<16, 17, 18>{}, std::integral_constant<int, 16>{}, 0x70000u);
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        b.lam_out[lam_at(g, q, x, row)] = lam[q];
    });
    contract(acc, w.K, b.dW, n, gid, gxw[0], gxw[1], gxw[2], gxw[3]);
}

template <int S, bool RAY>
__global__ __launch_bounds__(256) void adjoint_general_p2a(AdjointBuffers b, StepParams p, AdjointParams w) {
    const uint32_t count = *b.general_count;
    const float bx = (float)b.body[0], by = (float)b.body[1], bz = (float)b.body[2];
    for (uint32_t i = blockIdx.x; i < count; i += gridDim.x) {
        int x0, y0, z;
        tile_origin(p.g, b.general_list[i], x0, y0, z);
        const int x = x0 + (int)(threadIdx.x & (TILE_X - 1));
        const int y = y0 + (int)(threadIdx.x / TILE_X);
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (x < p.g.nx && y < p.g.ny)
            general_p2a_cell<S, RAY>(b, p, w, x, y, z, bx, by, bz, b.nu + (size_t)i * 19 * 256 + threadIdx.x, acc);
        block_sum8(acc, b.grad_partials + (size_t)i * 8);
    }
}

// ---------------------------------------------------------------------------
// Wall tiles, pass 2b: populations one link beyond the cell.
// ---------------------------------------------------------------------------
template <int S>
__global__ __launch_bounds__(256) void adjoint_general_p2b(AdjointBuffers b, StepParams p, AdjointParams w) {
    const uint32_t count = *b.general_count;
    const float bx = (float)b.body[0], by = (float)b.body[1], bz = (float)b.body[2];
    const Grid& g = p.g;
    for (uint32_t i = blockIdx.x; i < count; i += gridDim.x) {
        int x0, y0, z;
        tile_origin(g, b.general_list[i], x0, y0, z);
        const int x = x0 + (int)(threadIdx.x & (TILE_X - 1));
        const int y = y0 + (int)(threadIdx.x / TILE_X);
        if (x >= g.nx || y >= g.ny) continue;
        const uint32_t row = (uint32_t)y + (uint32_t)g.ny * (uint32_t)z;
        const uint32_t gid = (uint32_t)x + (uint32_t)g.nx * row;
        const float* nu_cell = b.nu + (size_t)i * 19 * 256 + threadIdx.x;
        if (b.solid_cur[gid] != 0) continue;
        if (b.solid_prev[gid] != 0) {
            if (!(p.accumulate_loads && p.load_component != 1 && p.force_mode == 6)) continue;
            const float arm[3] = {p.origin[0] + ((float)x + 0.5f) * p.dx - bx,
                                  p.origin[1] + ((float)y + 0.5f) * p.dx - by,
                                  p.origin[2] + ((float)z + 0.5f) * p.dx - bz};
            for_q(AllQ{}, [&](auto qc) {
                constexpr int q = decltype(qc)::value;
                if constexpr (q != 0) {
                    const int fx = x + cx(q), fy = y + cy(q), fz = z + cz(q);
                    if (inside(g, fx, fy, fz)) {
                        const uint32_t nb = flat(g, fx, fy, fz);
                        if (b.solid_prev[nb] == 0 && b.solid_cur[nb] == 0)
                            b.lam_out[lam_at(g, opp(q), fx, (uint32_t)fy + (uint32_t)g.ny * (uint32_t)fz)] +=
                                -p.force_scale * load_weight<q>(w, arm);
                    }
                }
            });
            continue;
        }
        if (p.wall_rule == WALL_HALFWAY) continue;
        const Sources s = sources(p, x, y, z);
        const float4 wg4 = b.wall[gid];
        const float wgv[4] = {wg4.x, wg4.y, wg4.z, wg4.w};
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            bool far;
            const uint32_t src = source_of<q>(s, g, x, y, z, far);
            if (far || b.solid_cur[src] == 0) return;
            const int fx = x + cx(q), fy = y + cy(q), fz = z + cz(q);
            if (!inside(g, fx, fy, fz)) return;
            const uint32_t fid = flat(g, fx, fy, fz);
            if (b.solid_cur[fid] != 0) return;
            const float4 ws4 = b.wall[src];
            const float wsv[4] = {ws4.x, ws4.y, ws4.z, ws4.w};
            float tlf;
            uint32_t win;
            const int rule = link_rule(b, p, x, y, z, q, tlf, win);
            const adj::LinkD<float> d = adj::link_value<float>(
                cx(q), cy(q), cz(q), weight<float>(q), p.far_rho, rule, true,
                fprev<S>(b, pop_index<S>(g, opp(q), x, row), opp(q)),
                fprev<S>(b, pop_index<S>(g, opp(q), fx, (uint32_t)fy + (uint32_t)g.ny * (uint32_t)fz), opp(q)),
                fprev<S>(b, pop_index<S>(g, q, x, row), q), wgv, wsv, tlf);
            if (d.uses_farther)
                b.lam_out[lam_at(g, opp(q), fx, (uint32_t)fy + (uint32_t)g.ny * (uint32_t)fz)] +=
                    nu_cell[q * 256] * d.d_farther;
        });
    }
}

// Fixed-order sum of the listed tiles' gradient partials into the running
// FP64 gradient.
__global__ __launch_bounds__(256) void reduce_gradient(const double* __restrict__ partials,
                                                       const uint32_t* __restrict__ count,
                                                       double* __restrict__ total) {
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint32_t c = *count;
    for (uint32_t i = threadIdx.x; i < c; i += blockDim.x) {
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] += partials[(size_t)i * 8 + k];
    }
    __shared__ double out[8];
    block_sum8(v, out);
    if (threadIdx.x < 8) total[threadIdx.x] += out[threadIdx.x];
}

// ---------------------------------------------------------------------------
// Wall-field tangents.  Mesh frames carry per-parameter vertex tangents;
// vertex positions and velocities at the query time are the frame
// interpolation (prepare), so their tangents interpolate the same way.
// ---------------------------------------------------------------------------
// Tangents of a body-frame vertex placed by the body pose (place_vertex in
// bf_geometry.hip: p = x_b + R lp, v = s (v_b + R (w_b x lp) + R lv)): the
// pose does not depend on theta, so dp = R dlp, dv = s (R (w_b x dlp) + R dlv).
__device__ __forceinline__ void rotate_tangent(const double* __restrict__ body, const float v[3], float o[3]) {
    double qd[4] = {body[3], body[4], body[5], body[6]};
    const double n2 = qd[0] * qd[0] + qd[1] * qd[1] + qd[2] * qd[2] + qd[3] * qd[3];
    if (n2 > 1.0e-20) {
        const double sc = 1.0 / sqrt(n2);
        for (int i = 0; i < 4; ++i) qd[i] *= sc;
    } else {
        qd[0] = qd[1] = qd[2] = 0.0;
        qd[3] = 1.0;
    }
    const float q[4] = {(float)qd[0], (float)qd[1], (float)qd[2], (float)qd[3]};
    const float t[3] = {2.0f * (q[1] * v[2] - q[2] * v[1]), 2.0f * (q[2] * v[0] - q[0] * v[2]),
                        2.0f * (q[0] * v[1] - q[1] * v[0])};
    const float c[3] = {q[1] * t[2] - q[2] * t[1], q[2] * t[0] - q[0] * t[2], q[0] * t[1] - q[1] * t[0]};
    for (int i = 0; i < 3; ++i) o[i] = v[i] + q[3] * t[i] + c[i];
}

__device__ __forceinline__ void place_tangent(const float dlp[3], const float dlv[3], const double* __restrict__ body,
                                              const MeshParams& m, float4& dp, float4& dv) {
    if (!m.move_with_body) {
        dp = make_float4(dlp[0], dlp[1], dlp[2], 0.0f);
        dv = make_float4(dlv[0] * m.velocity_scale, dlv[1] * m.velocity_scale, dlv[2] * m.velocity_scale, 0.0f);
        return;
    }
    const float w[3] = {(float)body[10], (float)body[11], (float)body[12]};
    const float wxp[3] = {w[1] * dlp[2] - w[2] * dlp[1], w[2] * dlp[0] - w[0] * dlp[2], w[0] * dlp[1] - w[1] * dlp[0]};
    float r[3], rw[3], ru[3];
    rotate_tangent(body, dlp, r);
    rotate_tangent(body, wxp, rw);
    rotate_tangent(body, dlv, ru);
    dp = make_float4(r[0], r[1], r[2], 0.0f);
    dv = make_float4((rw[0] + ru[0]) * m.velocity_scale, (rw[1] + ru[1]) * m.velocity_scale,
                     (rw[2] + ru[2]) * m.velocity_scale, 0.0f);
}

__global__ void tangent_vertices(const float* __restrict__ dpoints, const float* __restrict__ times,
                                 const double* __restrict__ body, float4* __restrict__ dpos,
                                 float4* __restrict__ dvel, MeshParams m, int K) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= m.vertex_count) return;
    const uint32_t first = min(m.first, m.frame_count - 2u);
    const uint32_t second = first + 1u;
    const float duration = times[second] - times[first];
    const float blend = fminf(fmaxf((m.query_time - times[first]) / duration, 0.0f), 1.0f);
    for (int k = 0; k < K; ++k) {
        const float* a = dpoints + (((size_t)k * m.frame_count + first) * m.vertex_count + v) * 3;
        const float* b = dpoints + (((size_t)k * m.frame_count + second) * m.vertex_count + v) * 3;
        const float d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const float dlp[3] = {a[0] + blend * d[0], a[1] + blend * d[1], a[2] + blend * d[2]};
        const float dlv[3] = {d[0] / duration, d[1] / duration, d[2] / duration};
        place_tangent(dlp, dlv, body, m, dpos[(size_t)k * m.vertex_count + v], dvel[(size_t)k * m.vertex_count + v]);
    }
}

// Dual number (value, derivative along one parameter).
struct D {
    float v, d;
};
__device__ __forceinline__ D operator+(D a, D b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ D operator-(D a, D b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ D operator*(D a, D b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ D operator/(D a, D b) {
    const float q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ D dconst(float v) { return {v, 0.0f}; }
__device__ __forceinline__ D ddot(const D a[3], const D b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Closest point on triangle (a, b, c) to the fixed point P with the Voronoi
// region chosen on the values (the region is held fixed, like the link set),
// derivatives along one parameter.  Same regions and formulas as
// closest_point in bf_geometry.hip.
__device__ void closest_point_dual(const float P[3], const D a[3], const D b[3], const D c[3], D cp[3], D bc[3]) {
    D ab[3], ac[3], ap[3], bp[3], cpv[3];
    for (int i = 0; i < 3; ++i) {
        ab[i] = b[i] - a[i];
        ac[i] = c[i] - a[i];
        ap[i] = dconst(P[i]) - a[i];
        bp[i] = dconst(P[i]) - b[i];
        cpv[i] = dconst(P[i]) - c[i];
    }
    const D d1 = ddot(ab, ap), d2 = ddot(ac, ap);
    auto set = [&](const D p[3], D b0, D b1, D b2) {
        for (int i = 0; i < 3; ++i) cp[i] = p[i];
        bc[0] = b0;
        bc[1] = b1;
        bc[2] = b2;
    };
    if (d1.v <= 0.0f && d2.v <= 0.0f) return set(a, dconst(1), dconst(0), dconst(0));
    const D d3 = ddot(ab, bp), d4 = ddot(ac, bp);
    if (d3.v >= 0.0f && d4.v <= d3.v) return set(b, dconst(0), dconst(1), dconst(0));
    const D vc = d1 * d4 - d3 * d2;
    if (vc.v <= 0.0f && d1.v >= 0.0f && d3.v <= 0.0f) {
        const D v = d1 / (d1 - d3);
        D p[3];
        for (int i = 0; i < 3; ++i) p[i] = a[i] + v * ab[i];
        return set(p, dconst(1) - v, v, dconst(0));
    }
    const D d5 = ddot(ab, cpv), d6 = ddot(ac, cpv);
    if (d6.v >= 0.0f && d5.v <= d6.v) return set(c, dconst(0), dconst(0), dconst(1));
    const D vb = d5 * d2 - d1 * d6;
    if (vb.v <= 0.0f && d2.v >= 0.0f && d6.v <= 0.0f) {
        const D w = d2 / (d2 - d6);
        D p[3];
        for (int i = 0; i < 3; ++i) p[i] = a[i] + w * ac[i];
        return set(p, dconst(1) - w, dconst(0), w);
    }
    const D va = d3 * d6 - d5 * d4;
    if (va.v <= 0.0f && (d4.v - d3.v) >= 0.0f && (d5.v - d6.v) >= 0.0f) {
        const D w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        D p[3];
        for (int i = 0; i < 3; ++i) p[i] = b[i] + w * (c[i] - b[i]);
        return set(p, dconst(0), dconst(1) - w, w);
    }
    D den = va + vb + vc;
    if (den.v < 1.0e-20f) den = dconst(1.0e-20f);
    const D v = vb / den, w = vc / den;
    D p[3];
    for (int i = 0; i < 3; ++i) p[i] = a[i] + ab[i] * v + ac[i] * w;
    return set(p, dconst(1) - v - w, v, w);
}

// Tangents of an articulated mesh (articulated_prepare in bf_geometry.hip)
// with respect to Fourier coefficients of its joints, by forward-mode dual
// numbers through the same kinematics: the angle of series c of a joint is
//   A_c(t) = s_a sum_h (a_h sin((h+1) phi) + b_h cos((h+1) phi)),  phi = omega (t mod T),
// its rate A'_c = s_a omega sum_h (h+1) (a_h cos - b_h sin), so
// dA_c/da_h = s_a sin((h+1) phi), dA'_c/da_h = s_a omega (h+1) cos((h+1) phi),
// dA_c/db_h = s_a cos((h+1) phi), dA'_c/db_h = -s_a omega (h+1) sin((h+1) phi);
// the rotations u3 = Ry(pitch + twist span) l, u2 = Rz(side dev) u3,
// u1 = Rx(side stroke) u2 and the analytic velocity carry the derivative.
__device__ __forceinline__ D dsin_(D a, float s, float c) { return {s, c * a.d}; }
__device__ __forceinline__ D dcos_(D a, float s, float c) { return {c, -(s * a.d)}; }
__device__ __forceinline__ void dsincos(D a, D& sn, D& cs) {
    float s, c;
    sincosf(a.v, &s, &c);
    sn = dsin_(a, s, c);
    cs = dcos_(a, s, c);
}

__global__ void articulated_tangents(const float* __restrict__ rest, const uint8_t* __restrict__ vertex_joint,
                                     const float* __restrict__ span_fraction, const Kinematics* __restrict__ kin,
                                     KinParams kp, const double* __restrict__ body, float4* __restrict__ dpos,
                                     float4* __restrict__ dvel, MeshParams m, int K) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= m.vertex_count) return;
    const Kinematics& k = *kin;
    const uint32_t jv = vertex_joint[v];
    const bool articulated = jv >= 1 && jv <= k.joint_count;
    const uint32_t j = articulated ? jv - 1 : 0;
    const KinJoint& J = k.joints[j];
    const uint32_t harmonics = min(k.harmonics, 4u);
    const float t = m.query_time;
    const float period = k.period;
    const float tm = t - floorf(t / period) * period;
    const float omega = 6.28318548f / period;
    const float phi = omega * tm;
    float sh[4] = {0.0f, 0.0f, 0.0f, 0.0f}, ch[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t h = 0; h < harmonics; ++h) sincosf((float)(h + 1) * phi, &sh[h], &ch[h]);
    const float(*series[4])[2] = {J.stroke, J.deviation, J.pitch, J.twist};
    float ang[4], rate[4];
    for (int c = 0; c < 4; ++c) {
        float a = 0.0f, r = 0.0f;
        for (uint32_t h = 0; h < harmonics; ++h) {
            a += series[c][h][0] * sh[h] + series[c][h][1] * ch[h];
            r += (float)(h + 1) * (series[c][h][0] * ch[h] - series[c][h][1] * sh[h]);
        }
        ang[c] = a * k.amplitude_scale;
        rate[c] = r * omega * k.amplitude_scale;
    }
    const float* rp = rest + 3 * (size_t)v;
    const float side = J.side;
    const float span = span_fraction[v];
    for (int p = 0; p < K; ++p) {
        const KinParam P = kp.p[p];
        float dlp[3] = {0.0f, 0.0f, 0.0f}, dlv[3] = {0.0f, 0.0f, 0.0f};
        if (articulated && ((P.joint_mask >> j) & 1u) && P.harmonic < harmonics) {
            const uint32_t h = P.harmonic;
            const float hs = (float)(h + 1);
            const float da = k.amplitude_scale * (P.component == 0 ? sh[h] : ch[h]);
            const float dr = k.amplitude_scale * omega * hs * (P.component == 0 ? ch[h] : -sh[h]);
            D A[4], R[4];
            for (int c = 0; c < 4; ++c) {
                A[c] = {ang[c], c == (int)P.channel ? da : 0.0f};
                R[c] = {rate[c], c == (int)P.channel ? dr : 0.0f};
            }
            const D sd = dconst(side);
            D sa, ca, sb, cb, sp, cp;
            dsincos(sd * A[0], sa, ca);
            dsincos(sd * A[1], sb, cb);
            const D l[3] = {dconst(rp[0]), dconst(side * rp[1]), dconst(rp[2])};
            const D a = A[2] + A[3] * dconst(span);
            const D adot = R[2] + R[3] * dconst(span);
            dsincos(a, sp, cp);
            const D u3[3] = {cp * l[0] + sp * l[2], l[1], cp * l[2] - sp * l[0]};
            const D u2[3] = {cb * u3[0] - sb * u3[1], sb * u3[0] + cb * u3[1], u3[2]};
            const D u1[3] = {u2[0], ca * u2[1] - sa * u2[2], sa * u2[1] + ca * u2[2]};
            const D ad = sd * R[0], bd = sd * R[1];
            const D zero = dconst(0.0f);
            const D t1[3] = {zero, zero - ad * u1[2], ad * u1[1]};
            const D w2[3] = {zero - bd * u2[1], bd * u2[0], zero};
            const D t2[3] = {w2[0], ca * w2[1] - sa * w2[2], sa * w2[1] + ca * w2[2]};
            const D w3[3] = {adot * u3[2], zero, zero - adot * u3[0]};
            const D z3[3] = {cb * w3[0] - sb * w3[1], sb * w3[0] + cb * w3[1], w3[2]};
            const D t3[3] = {z3[0], ca * z3[1] - sa * z3[2], sa * z3[1] + ca * z3[2]};
            for (int i = 0; i < 3; ++i) {
                dlp[i] = u1[i].d;
                dlv[i] = (t1[i] + t2[i] + t3[i]).d;
            }
        }
        place_tangent(dlp, dlv, body, m, dpos[(size_t)p * m.vertex_count + v], dvel[(size_t)p * m.vertex_count + v]);
    }
}

// The shell entry of ray_shell (bf_rays.hip) with the triangle's vertices
// carrying derivatives along one parameter: the same operations on the
// values (so the same branch: face, edge cylinder or vertex sphere) and
// their derivatives; the ray F + t D itself does not move.
__device__ __forceinline__ D dsqrt(D a) {
    const float r = sqrtf(a.v);
    return {r, r > 0.0f ? a.d / (2.0f * r) : 0.0f};
}
__device__ __forceinline__ void dsub3(const D a[3], const D b[3], D o[3]) {
    for (int i = 0; i < 3; ++i) o[i] = a[i] - b[i];
}
__device__ __forceinline__ void dcross3(const D a[3], const D b[3], D o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ bool dedge_in(const D U[3], const D V[3], const D P[3], const D n[3]) {
    D e[3], m[3], c[3];
    dsub3(V, U, e);
    dsub3(P, U, m);
    dcross3(e, m, c);
    return ddot(c, n).v >= 0.0f;
}

__device__ D ray_shell_dual(const float Ff[3], const float Df[3], const D A[3], const D B[3], const D C[3], float h) {
    const D F[3] = {dconst(Ff[0]), dconst(Ff[1]), dconst(Ff[2])};
    const D Dv[3] = {dconst(Df[0]), dconst(Df[1]), dconst(Df[2])};
    D best = dconst(2.0f);
    D e1[3], e2[3], n[3];
    dsub3(B, A, e1);
    dsub3(C, A, e2);
    dcross3(e1, e2, n);
    const D nl = dsqrt(ddot(n, n));
    if (nl.v > 0.0f) {
        for (int i = 0; i < 3; ++i) n[i] = n[i] / nl;
        D fa[3];
        dsub3(F, A, fa);
        const D fn = ddot(fa, n);
        if (fn.v <= h && fn.v >= -h && dedge_in(A, B, F, n) && dedge_in(B, C, F, n) && dedge_in(C, A, F, n))
            return dconst(0.0f);
        const D dn = ddot(Dv, n);
        if (dn.v != 0.0f) {
            for (int side = 0; side < 2; ++side) {
                const float off = side == 0 ? h : -h;
                const D t = (dconst(off) - fn) / dn;
                if (t.v >= 0.0f && t.v < best.v) {
                    const D P[3] = {F[0] + Dv[0] * t, F[1] + Dv[1] * t, F[2] + Dv[2] * t};
                    if (dedge_in(A, B, P, n) && dedge_in(B, C, P, n) && dedge_in(C, A, P, n)) best = t;
                }
            }
        }
    }
    const D dd = ddot(Dv, Dv);
    for (int k = 0; k < 3; ++k) {
        const D* u = k == 0 ? A : (k == 1 ? B : C);
        const D* v = k == 0 ? B : (k == 1 ? C : A);
        D e[3], m[3];
        dsub3(v, u, e);
        dsub3(F, u, m);
        const D ee = ddot(e, e), de = ddot(Dv, e), me = ddot(m, e), md = ddot(m, Dv), mm = ddot(m, m);
        const D a = ee * dd - de * de;
        const D hb = ee * md - me * de;
        const D c = (ee * mm - me * me) - dconst(h * h) * ee;
        if (c.v <= 0.0f && me.v >= 0.0f && me.v <= ee.v) return dconst(0.0f);
        if (a.v > 0.0f) {
            const D disc = hb * hb - a * c;
            if (disc.v >= 0.0f) {
                const D t = (dconst(0.0f) - hb - dsqrt(disc)) / a;
                if (t.v >= 0.0f && t.v < best.v) {
                    const D uu = (me + t * de) / ee;
                    if (uu.v >= 0.0f && uu.v <= 1.0f) best = t;
                }
            }
        }
    }
    for (int k = 0; k < 3; ++k) {
        const D* v = k == 0 ? A : (k == 1 ? B : C);
        D m[3];
        dsub3(F, v, m);
        const D md = ddot(m, Dv);
        const D c = ddot(m, m) - dconst(h * h);
        if (c.v <= 0.0f) return dconst(0.0f);
        const D disc = md * md - dd * c;
        if (disc.v >= 0.0f) {
            const D t = (dconst(0.0f) - md - dsqrt(disc)) / dd;
            if (t.v >= 0.0f && t.v < best.v) best = t;
        }
    }
    return best;
}

__device__ __noinline__ RayDQ ray_link_dq(const AdjointBuffers& b, const StepParams& p, int x, int y, int z, int q,
                                          uint32_t t, int K) {
    RayDQ r;
    for (int k = 0; k < 8; ++k) r.v[k] = 0.0f;
    // the link's segment as the ray pass forms it: from the fluid node of
    // the source S along c toward S (S wrapped / clamped by the faces)
    const int c[3] = {cx(q), cy(q), cz(q)};
    const int xyz[3] = {x, y, z};
    const int n[3] = {p.g.nx, p.g.ny, p.g.nz};
    float F[3], Dr[3];
    for (int a = 0; a < 3; ++a) {
        int sa = xyz[a];
        if (!ray_source(p.bc[a], xyz[a] - c[a], n[a], sa)) return r;
        const float Sa = p.origin[a] + ((float)sa + 0.5f) * p.dx;
        F[a] = Sa + (float)c[a] * p.dx;
        Dr[a] = (float)(-c[a]) * p.dx;
    }
    if (t == NO_WINNER) return r;
    const uint32_t iv[3] = {b.tri[3 * (size_t)t], b.tri[3 * (size_t)t + 1], b.tri[3 * (size_t)t + 2]};
    const float4 P0 = b.vpos[iv[0]], P1 = b.vpos[iv[1]], P2 = b.vpos[iv[2]];
    for (int k = 0; k < K; ++k) {
        const float4 d0 = b.dvpos[(size_t)k * b.vertex_count + iv[0]];
        const float4 d1 = b.dvpos[(size_t)k * b.vertex_count + iv[1]];
        const float4 d2 = b.dvpos[(size_t)k * b.vertex_count + iv[2]];
        const D A[3] = {{P0.x, d0.x}, {P0.y, d0.y}, {P0.z, d0.z}};
        const D B[3] = {{P1.x, d1.x}, {P1.y, d1.y}, {P1.z, d1.z}};
        const D C[3] = {{P2.x, d2.x}, {P2.y, d2.y}, {P2.z, d2.z}};
        r.v[k] = ray_shell_dual(F, Dr, A, B, C, b.half_thickness).d;
    }
    return r;
}

// d wall / d theta_k of every keyed cell: wall = (sum_i b_i v_i, (|P - cp| - h) / dx)
// Cells of the step's listed tiles only: the reverse step reads dW at its
// listed cells and at the solid sources of their links, whose tiles are
// listed too (a solid cell flags its own tile).
__global__ __launch_bounds__(256) void wall_tangent(const uint32_t* __restrict__ list,
                                                   const uint32_t* __restrict__ count,
                                                   const uint32_t* __restrict__ tri_of_cell,
                                                   const uint32_t* __restrict__ tri, const float4* __restrict__ pos,
                                                   const float4* __restrict__ vel, const float4* __restrict__ dpos,
                                                   const float4* __restrict__ dvel, float4* __restrict__ dW,
                                                   MeshParams m, int K) {
    const uint32_t tiles = *count;
    for (uint32_t i = blockIdx.x; i < tiles; i += gridDim.x) {
    int x0, y0, z;
    tile_origin(m.g, list[i], x0, y0, z);
    const int x = x0 + (int)(threadIdx.x & (TILE_X - 1));
    const int y = y0 + (int)(threadIdx.x / TILE_X);
    if (x >= m.g.nx || y >= m.g.ny) continue;
    const uint32_t gid = (uint32_t)x + (uint32_t)m.g.nx * ((uint32_t)y + (uint32_t)m.g.ny * (uint32_t)z);
    const uint32_t t = tri_of_cell[gid];
    if (t == NO_TRIANGLE || t >= m.triangle_count) continue;
    const float P[3] = {m.origin[0] + ((float)x + 0.5f) * m.dx, m.origin[1] + ((float)y + 0.5f) * m.dx,
                        m.origin[2] + ((float)z + 0.5f) * m.dx};
    const uint32_t iv[3] = {tri[3 * (size_t)t], tri[3 * (size_t)t + 1], tri[3 * (size_t)t + 2]};
    const size_t nv = m.vertex_count;
    for (int k = 0; k < K; ++k) {
        D vtx[3][3], vv[3][3];
        for (int j = 0; j < 3; ++j) {
            const float4 p4 = pos[iv[j]], dp4 = dpos[(size_t)k * nv + iv[j]];
            const float4 v4 = vel[iv[j]], dv4 = dvel[(size_t)k * nv + iv[j]];
            vtx[j][0] = {p4.x, dp4.x};
            vtx[j][1] = {p4.y, dp4.y};
            vtx[j][2] = {p4.z, dp4.z};
            vv[j][0] = {v4.x, dv4.x};
            vv[j][1] = {v4.y, dv4.y};
            vv[j][2] = {v4.z, dv4.z};
        }
        D cp[3], bc[3];
        closest_point_dual(P, vtx[0], vtx[1], vtx[2], cp, bc);
        D u[3];
        for (int i = 0; i < 3; ++i) u[i] = bc[0] * vv[0][i] + bc[1] * vv[1][i] + bc[2] * vv[2][i];
        const float r[3] = {P[0] - cp[0].v, P[1] - cp[1].v, P[2] - cp[2].v};
        const float len = sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        const float dlen = len > 0.0f ? -(r[0] * cp[0].d + r[1] * cp[1].d + r[2] * cp[2].d) / len : 0.0f;
        dW[(size_t)k * m.g.n + gid] = make_float4(u[0].d, u[1].d, u[2].d, dlen / m.dx);
    }
    }
}

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
template <int COLL, int S>
static hipError_t adjoint_step_t(const AdjointBuffers& b, const StepParams& p, const AdjointParams& w,
                                 int general_blocks, hipStream_t s) {
    hipLaunchKernelGGL((adjoint_fast<COLL, S>), dim3(p.g.tiles), dim3(256), 0, s, b, p);
    hipLaunchKernelGGL((adjoint_general_p1<COLL, S>), dim3(general_blocks), dim3(256), 0, s, b, p, w);
    if (p.bc[0][0] == BC_OUTLET || p.bc[0][1] == BC_OUTLET) {
        if (!b.outlet_extra) return hipErrorInvalidValue;
        const uint32_t threads = 2u * 19u * (uint32_t)p.g.ny * (uint32_t)p.g.nz;
        hipLaunchKernelGGL(adjoint_outlet, dim3((threads + 255) / 256), dim3(256), 0, s, b, p);
    }
    if (p.wall_rule == WALL_INTERPOLATED_LINKS)
        hipLaunchKernelGGL((adjoint_general_p2a<S, true>), dim3(general_blocks), dim3(256), 0, s, b, p, w);
    else
        hipLaunchKernelGGL((adjoint_general_p2a<S, false>), dim3(general_blocks), dim3(256), 0, s, b, p, w);
    hipLaunchKernelGGL(adjoint_general_p2b<S>, dim3(general_blocks), dim3(256), 0, s, b, p, w);
    hipLaunchKernelGGL(reduce_gradient, dim3(1), dim3(256), 0, s, b.grad_partials, b.general_count, b.grad_total);
    return hipGetLastError();
}

template <int S>
static hipError_t adjoint_step_s(int coll, const AdjointBuffers& b, const StepParams& p, const AdjointParams& w,
                                 int general_blocks, hipStream_t s) {
    switch (coll) {
        case COLL_TRT: return adjoint_step_t<COLL_TRT, S>(b, p, w, general_blocks, s);
        case COLL_TRT_SYMMETRIC_LIMITER: return adjoint_step_t<COLL_TRT_SYMMETRIC_LIMITER, S>(b, p, w, general_blocks, s);
        case COLL_REGULARIZED: return adjoint_step_t<COLL_REGULARIZED, S>(b, p, w, general_blocks, s);
        case COLL_RR3: return adjoint_step_t<COLL_RR3, S>(b, p, w, general_blocks, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_adjoint_step(int coll, int storage, const AdjointBuffers& b, const StepParams& p,
                               const AdjointParams& w, int general_blocks, hipStream_t s) {
    switch (storage) {
        case STORE_FP32: return adjoint_step_s<STORE_FP32>(coll, b, p, w, general_blocks, s);
        case STORE_FP16_OFFSET: return adjoint_step_s<STORE_FP16_OFFSET>(coll, b, p, w, general_blocks, s);
        case STORE_BF16_OFFSET: return adjoint_step_s<STORE_BF16_OFFSET>(coll, b, p, w, general_blocks, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_wall_tangent(const uint32_t* list, const uint32_t* count, int blocks, const uint32_t* tri_of_cell,
                               const uint32_t* tri, const float4* pos, const float4* vel, const MeshBuffers& mb,
                               const float* dpoints, const KinParams& kp, float4* dpos, float4* dvel, float4* dW,
                               const MeshParams& m, int K, hipStream_t s) {
    if (K < 1 || K > 8) return hipErrorInvalidValue;
    const dim3 grid((m.vertex_count + 255) / 256);
    if (mb.kinematics)
        hipLaunchKernelGGL(articulated_tangents, grid, dim3(256), 0, s, mb.rest, mb.vertex_joint, mb.span_fraction,
                           mb.kinematics, kp, mb.body, dpos, dvel, m, K);
    else
        hipLaunchKernelGGL(tangent_vertices, grid, dim3(256), 0, s, dpoints, mb.times, mb.body, dpos, dvel, m, K);
    hipLaunchKernelGGL(wall_tangent, dim3(blocks), dim3(256), 0, s, list, count, tri_of_cell, tri, pos, vel, dpos,
                       dvel, dW, m, K);
    return hipGetLastError();
}

}  // namespace bf
