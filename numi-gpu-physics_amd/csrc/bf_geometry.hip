// Moving-wall geometry for gfx950: indexed-mesh vertex preparation, raster to
// nearest-triangle keys, resolve to part flags / wall velocity / signed
// distance, and the canonical translating sphere.
//
// Reference: Sources/BirdFlowMetal/Metal/BirdFlow.metal
//   closest point :1529-1583   prepare :1589-1622   raster :1638-1686
//   resolve :1697-1738         clear   :1898-1911   sphere :183-195, 2307-2350
//
// The reference clears and resolves the whole grid every step.  Here the
// raster marks the 64x4-cell tiles it keys (one byte per tile per step, plain
// stores); one tile pass then resolves the marked tiles and clears the tiles
// whose mask / wall entries are stale from earlier steps, so clear + resolve
// touch only the neighbourhood of the surface.  Keys are (distance bin << 32
// | triangle), ordering identical to the reference's (bin << 12 | triangle)
// for meshes of up to 4,096 triangles and without its 4,096-triangle limit.
#include "bf_launch.h"

#include <algorithm>

namespace bf {

constexpr uint64_t EMPTY_KEY = ~0ull;

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return ax * bx + ay * by + az * bz;
}

struct Closest {
    float px, py, pz;
    float b0, b1, b2;
};

// Voronoi-region closest point on triangle (a, b, c), BirdFlow.metal:1529-1583.
__device__ inline Closest closest_point(float px, float py, float pz, float3 a, float3 b, float3 c) {
    const float abx = b.x - a.x, aby = b.y - a.y, abz = b.z - a.z;
    const float acx = c.x - a.x, acy = c.y - a.y, acz = c.z - a.z;
    const float apx = px - a.x, apy = py - a.y, apz = pz - a.z;
    const float d1 = dot3(abx, aby, abz, apx, apy, apz);
    const float d2 = dot3(acx, acy, acz, apx, apy, apz);
    if (d1 <= 0.0f && d2 <= 0.0f) return {a.x, a.y, a.z, 1.0f, 0.0f, 0.0f};
    const float bpx = px - b.x, bpy = py - b.y, bpz = pz - b.z;
    const float d3 = dot3(abx, aby, abz, bpx, bpy, bpz);
    const float d4 = dot3(acx, acy, acz, bpx, bpy, bpz);
    if (d3 >= 0.0f && d4 <= d3) return {b.x, b.y, b.z, 0.0f, 1.0f, 0.0f};
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        const float v = d1 / (d1 - d3);
        return {a.x + v * abx, a.y + v * aby, a.z + v * abz, 1.0f - v, v, 0.0f};
    }
    const float cpx = px - c.x, cpy = py - c.y, cpz = pz - c.z;
    const float d5 = dot3(abx, aby, abz, cpx, cpy, cpz);
    const float d6 = dot3(acx, acy, acz, cpx, cpy, cpz);
    if (d6 >= 0.0f && d5 <= d6) return {c.x, c.y, c.z, 0.0f, 0.0f, 1.0f};
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        const float w = d2 / (d2 - d6);
        return {a.x + w * acx, a.y + w * acy, a.z + w * acz, 1.0f - w, 0.0f, w};
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return {b.x + w * (c.x - b.x), b.y + w * (c.y - b.y), b.z + w * (c.z - b.z), 0.0f, 1.0f - w, w};
    }
    const float inv = 1.0f / fmaxf(va + vb + vc, 1.0e-20f);
    const float v = vb * inv, w = vc * inv;
    return {a.x + abx * v + acx * w, a.y + aby * v + acy * w, a.z + abz * v + acz * w, 1.0f - v - w, v, w};
}


// BirdFlow.metal:1589-1622.  points: frame-major float3.
// v' = v + w t + q x t with t = 2 q x v (q = x, y, z, w), FP32
__device__ __forceinline__ void qrotate_f(const float q[4], const float v[3], float o[3]) {
    const float t[3] = {2.0f * (q[1] * v[2] - q[2] * v[1]), 2.0f * (q[2] * v[0] - q[0] * v[2]),
                        2.0f * (q[0] * v[1] - q[1] * v[0])};
    const float c[3] = {q[1] * t[2] - q[2] * t[1], q[2] * t[0] - q[0] * t[2], q[0] * t[1] - q[1] * t[0]};
    for (int i = 0; i < 3; ++i) o[i] = v[i] + q[3] * t[i] + c[i];
}

// Places vertex v from its local position lp and local velocity lv (per
// time unit).  With move_with_body the local frame is the body frame and the
// device body pose places it: p = x_b + R lp, v = s (v_b + R (w_b x lp) + R lv).
__device__ __forceinline__ void place_vertex(uint32_t v, const float lp[3], const float lv[3],
                                             const double* __restrict__ body, const MeshParams& m,
                                             float4* __restrict__ pos, float4* __restrict__ vel) {
    if (!m.move_with_body) {
        pos[v] = make_float4(lp[0], lp[1], lp[2], 0.0f);
        vel[v] = make_float4(lv[0] * m.velocity_scale, lv[1] * m.velocity_scale, lv[2] * m.velocity_scale, 0.0f);
        return;
    }
    // normalised pose in FP64, then rounded once to FP32
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
    const float w[3] = {(float)body[10], (float)body[11], (float)body[12]};
    const float wxp[3] = {w[1] * lp[2] - w[2] * lp[1], w[2] * lp[0] - w[0] * lp[2], w[0] * lp[1] - w[1] * lp[0]};
    float r[3], rw[3], ru[3];
    qrotate_f(q, lp, r);
    qrotate_f(q, wxp, rw);
    qrotate_f(q, lv, ru);
    float po[3], ve[3];
    for (int i = 0; i < 3; ++i) {
        po[i] = (float)body[i] + r[i];
        ve[i] = ((float)body[7 + i] + (rw[i] + ru[i])) * m.velocity_scale;
    }
    pos[v] = make_float4(po[0], po[1], po[2], 0.0f);
    vel[v] = make_float4(ve[0], ve[1], ve[2], 0.0f);
}

// Vertex positions and lattice velocities at the query time
// (BirdFlow.metal:1589-1622): frames interpolated, velocity from the frame
// difference.
__global__ void mesh_prepare(const float* __restrict__ points, const float* __restrict__ times,
                             float4* __restrict__ pos, float4* __restrict__ vel, const double* __restrict__ body,
                             MeshParams m) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= m.vertex_count) return;
    const uint32_t first = min(m.first, m.frame_count - 2u);
    const uint32_t second = first + 1u;
    const float duration = times[second] - times[first];
    const float blend = fminf(fmaxf((m.query_time - times[first]) / duration, 0.0f), 1.0f);
    const float* a = points + ((size_t)first * m.vertex_count + v) * 3;
    const float* b = points + ((size_t)second * m.vertex_count + v) * 3;
    const float dxv = b[0] - a[0], dyv = b[1] - a[1], dzv = b[2] - a[2];
    const float lp[3] = {m.translation[0] + (a[0] + blend * dxv), m.translation[1] + (a[1] + blend * dyv),
                         m.translation[2] + (a[2] + blend * dzv)};
    const float lv[3] = {dxv / duration, dyv / duration, dzv / duration};
    place_vertex(v, lp, lv, body, m, pos, vel);
}

// Articulated meshes (kinematics evaluated per vertex; defined below).
struct Kinematics;
__global__ void articulated_prepare(const float* __restrict__ rest, const uint8_t* __restrict__ vertex_joint,
                                    const float* __restrict__ span_fraction, const Kinematics* __restrict__ kin,
                                    float4* __restrict__ pos, float4* __restrict__ vel,
                                    const double* __restrict__ body, MeshParams m);

__global__ void fill_u8(uint8_t* __restrict__ p, uint8_t v, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void fill_u32(uint32_t* __restrict__ p, uint32_t v, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void fill_u64(uint64_t* __restrict__ p, uint64_t v, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void fill_wall(float4* __restrict__ p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 16.0f);
}

// Raster to nearest-triangle keys (BirdFlow.metal:1638-1686 bounds and key).
// One single-wavefront workgroup per batch of up to 64 triangles
// (m.raster_batch; grid-stride over batches), or, for a mesh of few large
// triangles, m.raster_split wavefronts per triangle (grid y), each taking
// every raster_split-th 64-column chunk of the box:
//  0. lane j loads triangle (batch + j) and sets up its expanded voxel AABB,
//     plane and bounding sphere, so one gather round trip serves the batch;
//     the wave then walks the batch one triangle at a time with the
//     triangle's values broadcast from lane j (readlane, scalar registers);
//  1. a conservative filter over the voxels of the AABB (cut to the bounding
//     sphere's box when the band is finite): lanes take the (x, y) columns of
//     the box and step through z (through the sphere's chord of the column
//     when the band is finite); a cell can lie in the candidate band
//     (m.cull_cells) only if its distance to the triangle's plane and to the
//     bounding sphere are within the band (both are lower bounds of the
//     distance to the triangle; a margin of 1e-3 cells absorbs rounding);
//  2. the exact closest point and key (the reference's arithmetic) for the
//     survivors, compacted into LDS so that every lane has work;
//  3. for cells within the band: the key is merged into a per-batch LDS
//     table (cell -> minimum key; a batch's triangles are neighbours on the
//     mesh and key largely the same cells); at the end of the batch each
//     table entry goes to the grid as one 64-bit atomicMin whose result is
//     unused (the wave does not wait for it), and marks the cell's 64x4 tile
//     for this step's tile pass (a crowded table sends a key straight to the
//     grid).  A cell within the band always ends with a key, so marking on
//     every attempt marks the same tiles as marking on a win.
// With a triangle list the batches walk the listed ids (the keys of the other
// triangles come from seed_keys, below).  Keys are a minimum over (cell,
// triangle) pairs: the result does not depend on the order of the atomics.  With m.cull_cells = +inf (the reference's
// full AABB) the filter passes every cell of the AABB.
// LDS per single-wavefront workgroup: candidates 2 KB + key table 8 KB.  At
// 10 KB the compiler targets 4 waves per SIMD and keeps the raster under 100
// VGPRs; at 24 KB it targeted 2 and reserved 169, which took register room
// from the concurrent fluid kernel.
constexpr int RASTER_CAP = 256;
constexpr int RASTER_TABLE_BITS = 9;
constexpr uint32_t RASTER_TABLE = 1u << RASTER_TABLE_BITS;
constexpr uint32_t RASTER_PROBES = 16;

__device__ __forceinline__ uint32_t lane_rank(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ float bcast(float v, uint32_t j) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)j));
}

__device__ __forceinline__ int bcast(int v, uint32_t j) { return __builtin_amdgcn_readlane(v, (int)j); }

__global__ __launch_bounds__(64) void mesh_raster(const float4* __restrict__ pos, const uint32_t* __restrict__ tri,
                                                  const uint32_t* __restrict__ list,
                                                  unsigned long long* __restrict__ keys,
                                                  uint8_t* __restrict__ tiles_new, MeshParams m) {
    __shared__ uint2 cand[RASTER_CAP];
    // per-batch minimum key of each cell the batch's triangles key (open
    // addressing, linear probing); flushed to the grid at the end of the batch
    __shared__ unsigned long long tkey[RASTER_TABLE];
    __shared__ uint32_t tcell[RASTER_TABLE];
    __shared__ uint32_t ttile[RASTER_TABLE];
    const uint32_t lane = threadIdx.x;
    for (uint32_t j = lane; j < RASTER_TABLE; j += 64) {
        tkey[j] = EMPTY_KEY;
        tcell[j] = 0xFFFFFFFFu;
    }
    __syncthreads();
    const float dx = m.dx;
    const float guard = m.half_thickness + 2.0f * dx;
    const int n[3] = {m.g.nx, m.g.ny, m.g.nz};
    const bool filter = isfinite(m.cull_cells);
    const float reach = (filter ? m.cull_cells : 0.0f) * dx + 1.0e-3f * dx;
    const uint32_t bsz = m.raster_batch;  // triangles per batch, 1..64
    for (uint32_t batch = blockIdx.x * bsz; batch < m.raster_count; batch += gridDim.x * bsz) {
        // 0. per-lane set-up of triangle batch + lane
        const bool mine = lane < bsz && batch + lane < m.raster_count;
        const uint32_t tl = !mine ? 0u : (list ? list[batch + lane] : batch + lane);
        float3 a = make_float3(0.0f, 0.0f, 0.0f), b = a, c = a;
        if (mine) {
            const float4 a4 = pos[tri[3 * (size_t)tl + 0]];
            const float4 b4 = pos[tri[3 * (size_t)tl + 1]];
            const float4 c4 = pos[tri[3 * (size_t)tl + 2]];
            a = make_float3(a4.x, a4.y, a4.z);
            b = make_float3(b4.x, b4.y, b4.z);
            c = make_float3(c4.x, c4.y, c4.z);
        }
        float nx_ = (b.y - a.y) * (c.z - a.z) - (b.z - a.z) * (c.y - a.y);
        float ny_ = (b.z - a.z) * (c.x - a.x) - (b.x - a.x) * (c.z - a.z);
        float nz_ = (b.x - a.x) * (c.y - a.y) - (b.y - a.y) * (c.x - a.x);
        const float nl = sqrtf(nx_ * nx_ + ny_ * ny_ + nz_ * nz_);
        if (nl > 0.0f) {
            nx_ /= nl;
            ny_ /= nl;
            nz_ /= nl;
        }
        const float cm[3] = {(a.x + b.x + c.x) / 3.0f, (a.y + b.y + c.y) / 3.0f, (a.z + b.z + c.z) / 3.0f};
        auto d2 = [](float x, float y, float z) { return x * x + y * y + z * z; };
        const float r = sqrtf(fmaxf(d2(a.x - cm[0], a.y - cm[1], a.z - cm[2]),
                                    fmaxf(d2(b.x - cm[0], b.y - cm[1], b.z - cm[2]),
                                          d2(c.x - cm[0], c.y - cm[1], c.z - cm[2]))));
        const float R = reach + r * 1.0001f;
        const float sphere2 = R * R;
        const float av[3] = {a.x, a.y, a.z}, bv[3] = {b.x, b.y, b.z}, cv[3] = {c.x, c.y, c.z};
        int lo[3], ext[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float mn = fminf(av[i], fminf(bv[i], cv[i]));
            const float mx = fmaxf(av[i], fmaxf(bv[i], cv[i]));
            const float ls = ((mn - guard) - m.origin[i]) / dx - 0.5f;
            const float us = ((mx + guard) - m.origin[i]) / dx - 0.5f;
            int l = min(max((int)floorf(ls), 0), n[i] - 1);
            int h = min(max((int)ceilf(us), 0), n[i] - 1);
            if (filter) {
                // cells whose centre lies within R of the sphere's centre along
                // this axis (the sphere test fails for every other cell: its
                // squared distance is at least this axis' term); 0.01 cell of
                // margin absorbs rounding
                const float top = (float)n[i] + 2.0f;
                const float sl = fminf(fmaxf(((cm[i] - R) - m.origin[i]) / dx - 0.5f, -2.0f), top);
                const float su = fminf(fmaxf(((cm[i] + R) - m.origin[i]) / dx - 0.5f, -2.0f), top);
                l = max(l, (int)floorf(sl - 0.01f));
                h = min(h, (int)ceilf(su + 0.01f));
            }
            lo[i] = l;
            ext[i] = mine ? max(h - l + 1, 0) : 0;
        }
        const uint32_t nb = min(bsz, m.raster_count - batch);
        for (uint32_t j = 0; j < nb; ++j) {
            const uint32_t t = (uint32_t)bcast((int)tl, j);
            const float3 ta = make_float3(bcast(a.x, j), bcast(a.y, j), bcast(a.z, j));
            const float3 tb = make_float3(bcast(b.x, j), bcast(b.y, j), bcast(b.z, j));
            const float3 tc = make_float3(bcast(c.x, j), bcast(c.y, j), bcast(c.z, j));
            const float tnx = bcast(nx_, j), tny = bcast(ny_, j), tnz = bcast(nz_, j), tnl = bcast(nl, j);
            const float tcx = bcast(cm[0], j), tcy = bcast(cm[1], j), tcz = bcast(cm[2], j);
            const float ts2 = bcast(sphere2, j);
            const int lx = bcast(lo[0], j), ly = bcast(lo[1], j), lz = bcast(lo[2], j);
            const int ex = bcast(ext[0], j), ey = bcast(ext[1], j), ez = bcast(ext[2], j);
            const uint32_t cols = (uint32_t)ex * (uint32_t)ey;
            uint32_t count = 0;  // wave-uniform
            // 2./3. the exact key for the compacted candidates
            auto flush = [&]() {
                __syncthreads();
                for (uint32_t k = lane; k < count; k += 64) {
                    const uint2 e = cand[k];
                    const int cx = (int)(e.x & 0xFFFFu), cy = (int)(e.x >> 16), cz = (int)e.y;
                    const float px = m.origin[0] + ((float)cx + 0.5f) * dx;
                    const float py = m.origin[1] + ((float)cy + 0.5f) * dx;
                    const float pz = m.origin[2] + ((float)cz + 0.5f) * dx;
                    const Closest cp = closest_point(px, py, pz, ta, tb, tc);
                    const float ddx = px - cp.px, ddy = py - cp.py, ddz = pz - cp.pz;
                    const float dist_cells = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz) / dx;
                    if (dist_cells <= m.cull_cells) {
                        const float scaled = roundf(dist_cells * 65536.0f);
                        const uint32_t bin = scaled >= (float)0xFFFFF ? 0xFFFFFu : (uint32_t)scaled;
                        const unsigned long long key = ((unsigned long long)bin << 32) | (unsigned long long)t;
                        const uint32_t g = (uint32_t)cx + (uint32_t)m.g.nx * ((uint32_t)cy + (uint32_t)m.g.ny * (uint32_t)cz);
                        const uint32_t tile = tile_of(m.g, cx, cy, cz);
                        const uint32_t h = (g * 2654435761u) >> (32 - RASTER_TABLE_BITS);
                        bool done = false;
#pragma unroll 1
                        for (uint32_t probe = 0; probe < RASTER_PROBES && !done; ++probe) {
                            const uint32_t slot = (h + probe) & (RASTER_TABLE - 1);
                            const uint32_t prev = atomicCAS(&tcell[slot], 0xFFFFFFFFu, g);
                            if (prev == 0xFFFFFFFFu) ttile[slot] = tile;
                            if (prev == 0xFFFFFFFFu || prev == g) {
                                (void)atomicMin(&tkey[slot], key);
                                done = true;
                            }
                        }
                        if (!done) {
                            // table crowded: straight to the grid
                            (void)__hip_atomic_fetch_min(&keys[g], key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            tiles_new[tile] = 1;
                        }
                    }
                }
                __syncthreads();
                count = 0;
            };
            for (uint32_t c0 = blockIdx.y * 64u; c0 < cols; c0 += gridDim.y * 64u) {
                const uint32_t col = c0 + lane;
                const int y = ly + (int)(col / (uint32_t)max(ex, 1));
                const int x = lx + (int)(col - (uint32_t)(y - ly) * (uint32_t)ex);
                const float wx = m.origin[0] + ((float)x + 0.5f) * dx;
                const float wy = m.origin[1] + ((float)y + 0.5f) * dx;
                // this lane's z range: the whole box, or with a finite band the
                // chord of the bounding sphere through the column (a cell whose
                // (x, y) distance alone exceeds the sphere fails the sphere test
                // for every z; the chord is widened by 1e-5 R^2 under the root
                // and 0.01 cell for rounding)
                int zlo = 0, zhi = ez - 1;
                if (col >= cols) {
                    zlo = 1 << 30;
                    zhi = -1;
                } else if (filter) {
                    const float ddx = wx - tcx, ddy = wy - tcy;
                    const float dxy2 = ddx * ddx + ddy * ddy;
                    if (dxy2 > ts2) {
                        zlo = 1 << 30;
                        zhi = -1;
                    } else {
                        const float half = sqrtf((ts2 - dxy2) + 1.0e-5f * ts2);
                        const float top = (float)n[2] + 2.0f;
                        const float kl = fminf(fmaxf(((tcz - half) - m.origin[2]) / dx - 0.5f, -2.0f), top);
                        const float kh = fminf(fmaxf(((tcz + half) - m.origin[2]) / dx - 0.5f, -2.0f), top);
                        zlo = max(0, (int)floorf(kl - 0.01f) - lz);
                        zhi = min(ez - 1, (int)ceilf(kh + 0.01f) - lz);
                    }
                }
                int wlo = zlo, whi = zhi;
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    wlo = min(wlo, __shfl_xor(wlo, off, 64));
                    whi = max(whi, __shfl_xor(whi, off, 64));
                }
                for (int zi = wlo; zi <= whi; ++zi) {
                    const int z = lz + zi;
                    bool pass = zi >= zlo && zi <= zhi;
                    if (pass && filter) {
                        const float wz = m.origin[2] + ((float)z + 0.5f) * dx;
                        const float plane = fabsf(tnx * (wx - ta.x) + tny * (wy - ta.y) + tnz * (wz - ta.z));
                        pass = (tnl == 0.0f || plane <= reach) && d2(wx - tcx, wy - tcy, wz - tcz) <= ts2;
                    }
                    const unsigned long long ballot = __ballot(pass);
                    if (pass) cand[count + lane_rank(ballot)] = make_uint2((uint32_t)x | ((uint32_t)y << 16), (uint32_t)z);
                    count += (uint32_t)__popcll(ballot);
                    if (count > RASTER_CAP - 64) flush();
                }
            }
            if (count > 0) flush();
        }
        // the batch's minimum keys to the grid (64-bit atomicMin whose result
        // is unused) and their tiles marked; the table is emptied
        __syncthreads();
        for (uint32_t j = lane; j < RASTER_TABLE; j += 64) {
            const uint32_t g = tcell[j];
            if (g != 0xFFFFFFFFu) {
                (void)__hip_atomic_fetch_min(&keys[g], tkey[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                tiles_new[ttile[j]] = 1;
                tkey[j] = EMPTY_KEY;
                tcell[j] = 0xFFFFFFFFu;
            }
        }
        __syncthreads();
    }
}

// Seeds the keys of the triangles the raster skips (static triangles: the
// same vertex positions in every frame, so the same keys every step) from
// their cached (cell, key) pairs, and marks the cells' tiles.  Runs before the
// raster of the moving triangles, when every key is empty (each tile pass
// empties the keys it resolves), so plain stores suffice; the raster's
// atomicMin then takes the minimum over both sets.
__global__ void seed_keys(const uint32_t* __restrict__ cells, const unsigned long long* __restrict__ cached,
                          uint32_t count, unsigned long long* __restrict__ keys, uint8_t* __restrict__ tiles_new,
                          Grid g) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t gid = cells[i];
    keys[gid] = cached[i];
    const uint32_t row = gid / (uint32_t)g.nx;
    tiles_new[tile_of(g, (int)(gid - row * (uint32_t)g.nx), (int)(row % (uint32_t)g.ny), (int)(row / (uint32_t)g.ny))] = 1;
}

// Number of keyed cells (capacity 0), or: moves every key into the next
// (cell, key) slot and empties the cell.  One atomic per wavefront.
__global__ void collect_keys(unsigned long long* __restrict__ keys, uint32_t n, uint32_t* __restrict__ cells,
                             unsigned long long* __restrict__ out, uint32_t capacity,
                             unsigned long long* __restrict__ counter) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned long long key = gid < n ? keys[gid] : EMPTY_KEY;
    const bool keyed = key != EMPTY_KEY;
    const unsigned long long b = __ballot(keyed);
    if (b == 0) return;
    unsigned long long base = 0;
    if ((threadIdx.x & 63) == 0) base = atomicAdd(counter, (unsigned long long)__popcll(b));
    base = __shfl(base, 0, 64);
    if (!keyed || capacity == 0) return;
    const unsigned long long slot = base + lane_rank(b);
    if (slot < capacity) {
        cells[slot] = gid;
        out[slot] = key;
        keys[gid] = EMPTY_KEY;
    }
}

// Marks the tiles whose cells pull from, or are, solid cell (x, y, z).
__device__ inline void mark_tiles(const Grid& g, const int periodic[3], uint8_t* __restrict__ dyn, int x, int y,
                                  int z) {
    int xs[2] = {x - 1, x + 1}, ys[2] = {y - 1, y + 1};
    int zs[3] = {z - 1, z, z + 1};
    for (int i = 0; i < 2; ++i) {
        if (xs[i] < 0) xs[i] = periodic[0] ? g.nx - 1 : 0;
        if (xs[i] >= g.nx) xs[i] = periodic[0] ? 0 : g.nx - 1;
        if (ys[i] < 0) ys[i] = periodic[1] ? g.ny - 1 : 0;
        if (ys[i] >= g.ny) ys[i] = periodic[1] ? 0 : g.ny - 1;
    }
    for (int i = 0; i < 3; ++i) {
        if (zs[i] < 0) zs[i] = periodic[2] ? g.nz - 1 : 0;
        if (zs[i] >= g.nz) zs[i] = periodic[2] ? 0 : g.nz - 1;
    }
    // x-1, x, x+1 lie in at most two tiles (the wrapped ones included)
    const int txs[3] = {xs[0] / TILE_X, x / TILE_X, xs[1] / TILE_X};
    const int tys[3] = {ys[0] / TILE_Y, y / TILE_Y, ys[1] / TILE_Y};
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i)
                dyn[((uint32_t)zs[k] * (uint32_t)g.tiles_y + (uint32_t)tys[j]) * (uint32_t)g.tiles_x
                    + (uint32_t)txs[i]] = 1;
}

// Tiles the geometry pass visits this step: keyed now, or holding mask /
// wall entries written by the steps that last used those buffers.
__global__ __launch_bounds__(256) void geometry_tile_list(const uint8_t* __restrict__ tiles_new,
                                                          const uint8_t* __restrict__ tiles_mask_old,
                                                          const uint8_t* __restrict__ tiles_wall_old, uint32_t tiles,
                                                          uint32_t* __restrict__ list, uint32_t* __restrict__ count) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool flagged = t < tiles && (tiles_new[t] | tiles_mask_old[t] | tiles_wall_old[t]) != 0;
    const unsigned long long b = __ballot(flagged);
    if (b == 0) return;
    uint32_t base = 0;
    if ((threadIdx.x & 63) == 0) base = atomicAdd(count, (uint32_t)__popcll(b));
    base = __shfl(base, 0, 64);
    if (flagged) list[base + lane_rank(b)] = t;
}

// BirdFlow.metal:1697-1738 on the listed tiles.  In a tile keyed this step
// every cell is written: keyed cells get the part flag (signed distance <=
// 0), barycentric wall velocity and signed distance / dx and their key is
// reset; the others are cleared (mask 0, wall (0,0,0,16)).  Other listed
// tiles only clear the stale buffer(s).  Solid cells mark the tiles whose
// cells pull from them.
__global__ __launch_bounds__(256) void geometry_tile_pass(
    const float4* __restrict__ pos, const float4* __restrict__ vel, const uint32_t* __restrict__ tri,
    const uint8_t* __restrict__ tri_part, unsigned long long* __restrict__ keys, const uint32_t* __restrict__ list,
    const uint32_t* __restrict__ count, const uint8_t* __restrict__ tiles_new,
    const uint8_t* __restrict__ tiles_mask_old, const uint8_t* __restrict__ tiles_wall_old, uint8_t* __restrict__ mask,
    float4* __restrict__ wall, uint8_t* __restrict__ dyn, uint32_t* __restrict__ tri_out, MeshParams m) {
    const uint32_t n_tiles = *count;
    for (uint32_t i = blockIdx.x; i < n_tiles; i += gridDim.x) {
        const uint32_t tile = list[i];
        int x0, y0, z;
        tile_origin(m.g, tile, x0, y0, z);
        const int x = x0 + (int)(threadIdx.x & (TILE_X - 1));
        const int y = y0 + (int)(threadIdx.x / TILE_X);
        if (x >= m.g.nx || y >= m.g.ny) continue;
        const uint32_t gid = (uint32_t)x + (uint32_t)m.g.nx * ((uint32_t)y + (uint32_t)m.g.ny * (uint32_t)z);
        const float4 cleared = make_float4(0.0f, 0.0f, 0.0f, 16.0f);
        if (!tiles_new[tile]) {
            if (tiles_mask_old[tile]) mask[gid] = 0;
            if (tiles_wall_old[tile]) {
                wall[gid] = cleared;
                if (tri_out) tri_out[gid] = 0xFFFFFFFFu;
            }
            continue;
        }
        const unsigned long long key = keys[gid];
        const uint32_t t = (uint32_t)(key & 0xFFFFFFFFull);
        if (key == EMPTY_KEY || t >= m.triangle_count) {
            mask[gid] = 0;
            wall[gid] = cleared;
            if (tri_out) tri_out[gid] = 0xFFFFFFFFu;
            continue;
        }
        keys[gid] = EMPTY_KEY;
        if (tri_out) tri_out[gid] = t;   // nearest triangle, for wall tangents
        const float wx = m.origin[0] + ((float)x + 0.5f) * m.dx;
        const float wy = m.origin[1] + ((float)y + 0.5f) * m.dx;
        const float wz = m.origin[2] + ((float)z + 0.5f) * m.dx;
        const uint32_t ia = tri[3 * (size_t)t], ib = tri[3 * (size_t)t + 1], ic = tri[3 * (size_t)t + 2];
        const float4 a4 = pos[ia], b4 = pos[ib], c4 = pos[ic];
        const Closest cp = closest_point(wx, wy, wz, make_float3(a4.x, a4.y, a4.z), make_float3(b4.x, b4.y, b4.z),
                                         make_float3(c4.x, c4.y, c4.z));
        const float4 va = vel[ia], vb = vel[ib], vc = vel[ic];
        const float ux = cp.b0 * va.x + cp.b1 * vb.x + cp.b2 * vc.x;
        const float uy = cp.b0 * va.y + cp.b1 * vb.y + cp.b2 * vc.y;
        const float uz = cp.b0 * va.z + cp.b1 * vb.z + cp.b2 * vc.z;
        const float ddx = wx - cp.px, ddy = wy - cp.py, ddz = wz - cp.pz;
        const float sd = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz) - m.half_thickness;
        const bool solid = sd <= 0.0f;
        mask[gid] = solid ? tri_part[t] : 0;
        wall[gid] = make_float4(ux, uy, uz, sd / m.dx);
        if (solid) mark_tiles(m.g, m.periodic, dyn, x, y, z);
    }
}

// Whole-grid tile marking for host-written or canonical geometry.
__global__ void mark_tiles_from_mask(const uint8_t* __restrict__ mask, uint8_t* __restrict__ dyn, Grid g,
                                     int px, int py, int pz) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= g.n || mask[gid] == 0) return;
    const int periodic[3] = {px, py, pz};
    const int x = (int)(gid % (uint32_t)g.nx);
    const uint32_t yz = gid / (uint32_t)g.nx;
    mark_tiles(g, periodic, dyn, x, (int)(yz % (uint32_t)g.ny), (int)(yz / (uint32_t)g.ny));
}

// Translating sphere (BirdFlow.metal:183-195, 2307-2350): occupancy at
// center(t) = c0 + v_geom t, uniform / tangential / normal wall velocity.
__global__ void sphere_geometry(uint8_t* __restrict__ mask, float4* __restrict__ wall, uint8_t* __restrict__ dyn,
                                Grid g, float4 center_radius, float3 geometry_velocity, float4 wall_velocity, float t,
                                int px, int py, int pz) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= g.n) return;
    const int x = (int)(gid % (uint32_t)g.nx);
    const uint32_t yz = gid / (uint32_t)g.nx;
    const int y = (int)(yz % (uint32_t)g.ny), z = (int)(yz / (uint32_t)g.ny);
    const float cxv = center_radius.x + geometry_velocity.x * t;
    const float cyv = center_radius.y + geometry_velocity.y * t;
    const float czv = center_radius.z + geometry_velocity.z * t;
    const float rx = (float)x + 0.5f - cxv, ry = (float)y + 0.5f - cyv, rz = (float)z + 0.5f - czv;
    const float r = center_radius.w;
    const bool in = dot3(rx, ry, rz, rx, ry, rz) <= r * r;
    float wx = wall_velocity.x, wy = wall_velocity.y, wz = wall_velocity.z;
    const int mode = (int)(fmaxf(wall_velocity.w, 0.0f) + 0.5f);
    if (mode != 0) {
        const float l = fmaxf(sqrtf(rx * rx + ry * ry + rz * rz), 1.0e-6f);
        const float nxv = rx / l, nyv = ry / l, nzv = rz / l;
        const float dn = dot3(wx, wy, wz, nxv, nyv, nzv);
        const float nwx = dn * nxv, nwy = dn * nyv, nwz = dn * nzv;
        if (mode == 1) {
            wx = wx - nwx;
            wy = wy - nwy;
            wz = wz - nwz;
        } else {
            wx = nwx;
            wy = nwy;
            wz = nwz;
        }
    }
    mask[gid] = in ? 1 : 0;
    wall[gid] = make_float4(wx, wy, wz, 0.0f);
    if (in) {
        const int periodic[3] = {px, py, pz};
        mark_tiles(g, periodic, dyn, x, y, z);
    }
}

// Link census: per direction, (fluid, persistent) cells whose pull source is
// solid; fresh cover and uncover counts.  counts[0..18] links, [19] covered,
// [20] uncovered.
__global__ void link_census(const uint8_t* __restrict__ prev, const uint8_t* __restrict__ cur, Grid g, int bcx0,
                            int bcx1, int bcy0, int bcy1, int bcz0, int bcz1,
                            unsigned long long* __restrict__ counts) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= g.n) return;
    const bool s = cur[gid] != 0, w = prev[gid] != 0;
    if (s && !w) atomicAdd(&counts[19], 1ull);
    if (!s && w) atomicAdd(&counts[20], 1ull);
    if (s || w) return;
    const int x = (int)(gid % (uint32_t)g.nx);
    const uint32_t yz = gid / (uint32_t)g.nx;
    const int y = (int)(yz % (uint32_t)g.ny), z = (int)(yz / (uint32_t)g.ny);
    const int bc[3][2] = {{bcx0, bcx1}, {bcy0, bcy1}, {bcz0, bcz1}};
    const int n[3] = {g.nx, g.ny, g.nz};
    for (int q = 1; q < 19; ++q) {
        int sv[3] = {x - cx(q), y - cy(q), z - cz(q)};
        bool far = false;
        for (int a = 0; a < 3; ++a) {
            if (sv[a] < 0) {
                if (bc[a][0] == BC_PERIODIC) sv[a] = n[a] - 1;
                else if (bc[a][0] == BC_OUTLET) sv[a] = 0;
                else far = true;
            } else if (sv[a] >= n[a]) {
                if (bc[a][1] == BC_PERIODIC) sv[a] = 0;
                else if (bc[a][1] == BC_OUTLET) sv[a] = n[a] - 1;
                else far = true;
            }
        }
        if (far) continue;
        const uint32_t src = (uint32_t)sv[0] + (uint32_t)g.nx * ((uint32_t)sv[1] + (uint32_t)g.ny * (uint32_t)sv[2]);
        if (cur[src] != 0) atomicAdd(&counts[q], 1ull);
    }
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------

hipError_t launch_mesh_prepare(const MeshBuffers& b, const MeshParams& m, hipStream_t s) {
    if (b.kinematics)
        hipLaunchKernelGGL(articulated_prepare, dim3((m.vertex_count + 255) / 256), dim3(256), 0, s, b.rest,
                           b.vertex_joint, b.span_fraction, b.kinematics, b.pos, b.vel, b.body, m);
    else
        hipLaunchKernelGGL(mesh_prepare, dim3((m.vertex_count + 255) / 256), dim3(256), 0, s, b.points, b.times,
                           b.pos, b.vel, b.body, m);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Analytic signed-distance bird.  prepareBirdGeometry (BirdFlow.metal:686-814,
// analytic kinematics) runs in one lane; buildBirdGeometry (:816-938) runs on
// the tiles that hold cells within the geometry radius, cells that were
// solid last step, or stale entries.  Conventions shared with the oracle:
// rsqrt(x) = 1 / sqrt(x), normalize(v) = v * (1 / |v|), mix(a, b, t) =
// a + (b - a) t, and bird_sincos below for sin / cos; no contraction, so the
// flags and wall velocities equal the oracle's bit for bit.
namespace {

struct V3 {
    float x, y, z;
};
__device__ inline V3 v3(float x, float y, float z) { return {x, y, z}; }
__device__ inline V3 vadd(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline V3 vmul(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ inline float vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float vlen(V3 a) { return sqrtf(vdot(a, a)); }
__device__ inline V3 vcross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
__device__ inline V3 vnormalize(V3 a) { return vmul(a, 1.0f / sqrtf(vdot(a, a))); }
__device__ inline float smin(float a, float b) { return a < b ? a : b; }
__device__ inline float smax(float a, float b) { return a > b ? a : b; }
__device__ inline float clamp01(float v) { return smin(smax(v, 0.0f), 1.0f); }
__device__ inline float mixf(float a, float b, float t) { return a + (b - a) * t; }
__device__ inline V3 row3(const float4& r) { return v3(r.x, r.y, r.z); }
__device__ inline float4 row4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

// x = k pi/2 + r, k = rint(x 2/pi), pi/2 in three floats (fused
// multiply-add reduction); Taylor series of sin / cos on |r| <= pi/4 through
// r^11 / r^12 by Horner with fused multiply-adds; quadrant k mod 4.
__device__ inline void bird_sincos(float x, float& s, float& c) {
    const float k = __builtin_rintf(x * 0x1.45f306p-1f);
    float r = __builtin_fmaf(-k, 0x1.921fb6p+0f, x);
    r = __builtin_fmaf(-k, -0x1.777a5cp-25f, r);
    r = __builtin_fmaf(-k, -0x1.ee59dap-50f, r);
    const float r2 = r * r;
    float ps = __builtin_fmaf(-1.0f / 39916800.0f, r2, 1.0f / 362880.0f);
    ps = __builtin_fmaf(ps, r2, -1.0f / 5040.0f);
    ps = __builtin_fmaf(ps, r2, 1.0f / 120.0f);
    ps = __builtin_fmaf(ps, r2, -1.0f / 6.0f);
    const float sr = __builtin_fmaf(r * r2, ps, r);
    float pc = __builtin_fmaf(1.0f / 479001600.0f, r2, -1.0f / 3628800.0f);
    pc = __builtin_fmaf(pc, r2, 1.0f / 40320.0f);
    pc = __builtin_fmaf(pc, r2, -1.0f / 720.0f);
    pc = __builtin_fmaf(pc, r2, 1.0f / 24.0f);
    pc = __builtin_fmaf(pc, r2, -0.5f);
    const float cr = __builtin_fmaf(r2, pc, 1.0f);
    switch ((int)(long long)k & 3) {
        case 0: s = sr; c = cr; break;
        case 1: s = cr; c = -sr; break;
        case 2: s = -sr; c = -cr; break;
        default: s = -cr; c = sr; break;
    }
}

struct Quat {
    float x, y, z, w;
};

__device__ inline V3 qrotate(Quat q, V3 v) {
    const V3 u = v3(q.x, q.y, q.z);
    const V3 t = vmul(vcross(u, v), 2.0f);
    return vadd(vadd(v, vmul(t, q.w)), vcross(u, t));
}

__device__ inline V3 rotate_axis(V3 value, V3 axis, float angle) {
    float s, c;
    bird_sincos(angle, s, c);
    return vadd(vadd(vmul(value, c), vmul(vcross(axis, value), s)), vmul(vmul(axis, vdot(axis, value)), 1.0f - c));
}

__device__ inline float sd_box(float dx, float dy, float dz) {
    return vlen(v3(smax(dx, 0.0f), smax(dy, 0.0f), smax(dz, 0.0f))) + smin(smax(dx, smax(dy, dz)), 0.0f);
}

__device__ inline float sd_ellipsoid(V3 p, const BirdShape& b) {
    const float rx = b.body_radii[0], ry = b.body_radii[1], rz = b.body_radii[2];
    const float k0 = vlen(v3(p.x / rx, p.y / ry, p.z / rz));
    const float k1 = vlen(v3(p.x / (rx * rx), p.y / (ry * ry), p.z / (rz * rz)));
    return k1 > 1.0e-12f ? k0 * (k0 - 1.0f) / k1 : -smin(rx, smin(ry, rz));
}

__device__ inline float sd_wing(V3 l, const BirdShape& b) {
    const float frac = clamp01(l.y / smax(b.wing_span, 1.0e-6f));
    const float chord = mixf(b.wing_root_chord, b.wing_tip_chord, frac);
    const float center = -b.wing_sweep * frac;
    return sd_box(fabsf(l.x - center) - 0.5f * chord, smax(-l.y, l.y - b.wing_span),
                  fabsf(l.z) - 0.5f * b.wing_thickness);
}

__device__ inline float sd_tail(V3 l, const BirdShape& b) {
    const float from_root = -(l.x + b.body_radii[0]);
    const float frac = clamp01(from_root / smax(b.tail_length, 1.0e-6f));
    const float half_width = mixf(0.35f * b.tail_half_width, b.tail_half_width, frac);
    return sd_box(smax(-from_root, from_root - b.tail_length), fabsf(l.y) - half_width,
                  fabsf(l.z + 0.15f * b.body_radii[2]) - 0.5f * b.tail_thickness);
}

// makeWingFrame (BirdFlow.metal:611-684); writes root, chord (w: tip
// twist), span, normal, angular velocity (w: tip twist rate).
__device__ inline void wing_frame(float side, const float (&a)[4], const float (&r)[4], const BirdShape& b, Quat q,
                                  V3 body_pos, float4* out) {
    const float stroke = a[0], deviation = a[1], pitch = a[2], tip_twist = a[3];
    const float stroke_rate = r[0], deviation_rate = r[1], pitch_rate = r[2], tip_twist_rate = r[3];
    const V3 ex = qrotate(q, v3(1.0f, 0.0f, 0.0f));
    const V3 ey = qrotate(q, v3(0.0f, 1.0f, 0.0f));
    const V3 ez = qrotate(q, v3(0.0f, 0.0f, 1.0f));
    const V3 root = vadd(body_pos, qrotate(q, v3(b.wing_root_offset[0], side * b.wing_root_offset[1],
                                                 b.wing_root_offset[2])));
    const float signed_stroke = side * stroke;
    const V3 span_stroke = rotate_axis(vmul(ey, side), ex, signed_stroke);
    const V3 normal_stroke = rotate_axis(vmul(ez, side), ex, signed_stroke);
    const V3 sp = rotate_axis(span_stroke, normal_stroke, deviation);
    const V3 chord_pre = rotate_axis(ex, normal_stroke, deviation);
    const float signed_pitch = side * pitch;
    const V3 ch = rotate_axis(chord_pre, sp, signed_pitch);
    const V3 nm = rotate_axis(normal_stroke, sp, signed_pitch);
    const V3 spn = vnormalize(sp);
    const V3 w = vadd(vadd(vmul(ex, side * stroke_rate), vmul(normal_stroke, deviation_rate)),
                      vmul(spn, side * pitch_rate));
    out[0] = row4(root, 0.0f);
    out[1] = row4(vnormalize(ch), side * tip_twist);
    out[2] = row4(spn, 0.0f);
    out[3] = row4(vnormalize(nm), 0.0f);
    out[4] = row4(w, side * tip_twist_rate);
}

// measuredWingKinematics (BirdFlow.metal:550-609): cubic Hermite between the
// keyframes around the cycle phase.  A keyframe is 17 floats: phase, left
// angles (stroke, deviation, pitch, tip twist), left rates (per second),
// right angles, right rates.
__device__ inline void measured_angles(const float* __restrict__ kf, int count, float frequency, float phase,
                                       int off, float (&ang)[4], float (&rate)[4]) {
    int low = 0, high = count;
    while (low < high) {
        const int middle = low + (high - low) / 2;
        if (kf[17 * middle] <= phase) low = middle + 1;
        else high = middle;
    }
    const int upper = low == count ? 0 : low;
    const int lower = upper == 0 ? count - 1 : upper - 1;
    const float phase0 = kf[17 * lower];
    const float phase1 = upper == 0 ? 1.0f : kf[17 * upper];
    const float adjusted = upper == 0 && phase < phase0 ? phase + 1.0f : phase;
    const float interval = phase1 - phase0;
    const float f = (adjusted - phase0) / interval;
    const float seconds = interval / frequency;
    const float t2 = f * f, t3 = t2 * f;
    const float h00 = 2.0f * t3 - 3.0f * t2 + 1.0f;
    const float h10 = t3 - 2.0f * t2 + f;
    const float h01 = -2.0f * t3 + 3.0f * t2;
    const float h11 = t3 - t2;
    const float d00 = 6.0f * t2 - 6.0f * f;
    const float d10 = 3.0f * t2 - 4.0f * f + 1.0f;
    const float d01 = -6.0f * t2 + 6.0f * f;
    const float d11 = 3.0f * t2 - 2.0f * f;
    const float* a0 = kf + 17 * lower + off;
    const float* a1 = kf + 17 * upper + off;
    for (int i = 0; i < 4; ++i) {
        ang[i] = h00 * a0[i] + h10 * seconds * a0[4 + i] + h01 * a1[i] + h11 * seconds * a1[4 + i];
        rate[i] = (d00 * a0[i] + d10 * seconds * a0[4 + i] + d01 * a1[i] + d11 * seconds * a1[4 + i]) / seconds;
    }
}

}  // namespace

// One lane: the frame rows body position, orientation, linear velocity,
// (world angular velocity, geometry radius), then left and right wing
// (root, chord, span, normal, angular velocity).
__global__ void bird_prepare(BirdShape b, const double* __restrict__ body, const float* __restrict__ keyframes,
                             int keyframe_count, float time, float dx, float4* __restrict__ frame) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const V3 pos = v3((float)body[0], (float)body[1], (float)body[2]);
    Quat q = {(float)body[3], (float)body[4], (float)body[5], (float)body[6]};
    const float n2 = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
    if (n2 > 1.0e-20f) {
        const float inv = 1.0f / sqrtf(n2);
        q = {q.x * inv, q.y * inv, q.z * inv, q.w * inv};
    } else {
        q = {0.0f, 0.0f, 0.0f, 1.0f};
    }
    const V3 lin = v3((float)body[7], (float)body[8], (float)body[9]);
    const V3 wb = v3((float)body[10], (float)body[11], (float)body[12]);
    float la[4], lr[4], ra[4], rr[4];
    if (keyframe_count > 0) {
        const float cycles = time * b.frequency;
        float phase = cycles - truncf(cycles);
        phase = phase < 0.0f ? phase + 1.0f : phase;
        measured_angles(keyframes, keyframe_count, b.frequency, phase, 1, la, lr);
        measured_angles(keyframes, keyframe_count, b.frequency, phase, 9, ra, rr);
    } else {
        const float omega = 2.0f * 3.14159265358979323846f * b.frequency;
        const float phase = omega * time;
        float s0, c0, s1, c1;
        bird_sincos(phase, s0, c0);
        bird_sincos(phase + b.pitch_phase, s1, c1);
        la[0] = b.stroke_bias + b.stroke_amplitude * s0;
        lr[0] = omega * b.stroke_amplitude * c0;
        la[2] = b.pitch_mean + b.pitch_amplitude * s1;
        lr[2] = omega * b.pitch_amplitude * c1;
        la[1] = lr[1] = la[3] = lr[3] = 0.0f;
        for (int i = 0; i < 4; ++i) {
            ra[i] = la[i];
            rr[i] = lr[i];
        }
    }
    float4 rows[BIRD_FRAME_ROWS];
    wing_frame(1.0f, la, lr, b, q, pos, rows + 4);
    wing_frame(-1.0f, ra, rr, b, q, pos, rows + 9);
    rows[0] = row4(pos, 0.0f);
    rows[1] = make_float4(q.x, q.y, q.z, q.w);
    rows[2] = row4(lin, 0.0f);
    const float body_radius = vlen(v3(b.body_radii[0], b.body_radii[1], b.body_radii[2]));
    const float tail_radius = vlen(v3(b.body_radii[0] + b.tail_length, smax(b.body_radii[1], b.tail_half_width),
                                      b.body_radii[2] + 0.5f * b.tail_thickness));
    const float root_radius = vlen(v3(b.wing_root_offset[0], b.wing_root_offset[1], b.wing_root_offset[2]));
    const float wing_radius = root_radius + b.wing_span + fabsf(b.wing_sweep)
                              + 0.5f * smax(b.wing_root_chord, b.wing_tip_chord) + 0.5f * b.wing_thickness;
    const float radius = smax(body_radius, smax(tail_radius, wing_radius)) + 2.0f * dx;
    rows[3] = row4(qrotate(q, wb), radius);
    for (int i = 0; i < BIRD_FRAME_ROWS; ++i) frame[i] = rows[i];
}

// Squared distance from point c to the box of a tile's cell centres.
__device__ inline float tile_box_dist2(float lx, float hx, float ly, float hy, float zc, float cx_, float cy_,
                                       float cz_) {
    const float dx = fmaxf(fmaxf(lx - cx_, cx_ - hx), 0.0f);
    const float dy = fmaxf(fmaxf(ly - cy_, cy_ - hy), 0.0f);
    const float dz = zc - cz_;
    return dx * dx + dy * dy + dz * dz;
}

// Tiles evaluated this step: those whose box of cell centres comes within
// two cells of a bounding sphere of the body ellipsoid (largest radius), of
// the tail box or of either wing box (centre at mid-span, radius half the
// span plus sweep, half the largest chord and half the thickness), so every
// cell that can be solid lies in one; and, with a previous geometry, the
// tiles near last step's solid cells (evaluated even beyond the radius, as
// the reference does).  Every tile's flag is written.
__global__ __launch_bounds__(256) void bird_mark_tiles(BirdShape b, const float4* __restrict__ frame,
                                                       const uint8_t* __restrict__ dyn_prev,
                                                       uint8_t* __restrict__ tiles_new, BirdParams p) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= p.g.tiles) return;
    int x0, y0, z;
    tile_origin(p.g, t, x0, y0, z);
    const int x1 = min(x0 + TILE_X, p.g.nx) - 1, y1 = min(y0 + TILE_Y, p.g.ny) - 1;
    const float lx = p.origin[0] + ((float)x0 + 0.5f) * p.dx, hx = p.origin[0] + ((float)x1 + 0.5f) * p.dx;
    const float ly = p.origin[1] + ((float)y0 + 0.5f) * p.dx, hy = p.origin[1] + ((float)y1 + 0.5f) * p.dx;
    const float zc = p.origin[2] + ((float)z + 0.5f) * p.dx;
    const float guard = 2.0f * p.dx;
    const float4 c = frame[0];
    const Quat q = {frame[1].x, frame[1].y, frame[1].z, frame[1].w};
    bool hit = false;
    {
        const float r = smax(b.body_radii[0], smax(b.body_radii[1], b.body_radii[2])) + guard;
        hit = hit || tile_box_dist2(lx, hx, ly, hy, zc, c.x, c.y, c.z) <= r * r;
    }
    {
        const V3 lc = v3(-b.body_radii[0] - 0.5f * b.tail_length, 0.0f, -0.15f * b.body_radii[2]);
        const V3 wc = vadd(v3(c.x, c.y, c.z), qrotate(q, lc));
        const float r = vlen(v3(0.5f * b.tail_length, b.tail_half_width, 0.5f * b.tail_thickness)) + guard;
        hit = hit || tile_box_dist2(lx, hx, ly, hy, zc, wc.x, wc.y, wc.z) <= r * r;
    }
    const float rw = 0.5f * b.wing_span + fabsf(b.wing_sweep) + 0.5f * smax(b.wing_root_chord, b.wing_tip_chord)
                     + 0.5f * b.wing_thickness + guard;
#pragma unroll
    for (int wing = 0; wing < 2; ++wing) {
        const V3 wc = vadd(row3(frame[4 + 5 * wing]), vmul(row3(frame[6 + 5 * wing]), 0.5f * b.wing_span));
        hit = hit || tile_box_dist2(lx, hx, ly, hy, zc, wc.x, wc.y, wc.z) <= rw * rw;
    }
    tiles_new[t] = (hit || (dyn_prev && dyn_prev[t])) ? 1 : 0;
}

// buildBirdGeometry on the tiles marked this step: every cell gets the
// reference's value (mask 0 and wall 0 beyond the radius unless solid last
// step).  Other listed tiles (stale entries of earlier steps) are cleared to
// mask 0, wall 0: their cells lie outside every component sphere and were
// not solid last step, so the flags equal the reference's everywhere and the
// wall velocity wherever a step reads it (cells solid this step or the
// previous one).  Solid cells mark the tiles whose cells pull from them.
__global__ __launch_bounds__(256) void bird_tile_pass(BirdShape b, const float4* __restrict__ frame,
                                                      const uint32_t* __restrict__ list,
                                                      const uint32_t* __restrict__ count,
                                                      const uint8_t* __restrict__ tiles_new,
                                                      const uint8_t* __restrict__ mask_prev,
                                                      uint8_t* __restrict__ mask, float4* __restrict__ wall,
                                                      uint8_t* __restrict__ dyn, BirdParams p) {
    __shared__ float4 fr[BIRD_FRAME_ROWS];
    if (threadIdx.x < BIRD_FRAME_ROWS) fr[threadIdx.x] = frame[threadIdx.x];
    __syncthreads();
    const V3 body_pos = row3(fr[0]);
    const Quat q = {fr[1].x, fr[1].y, fr[1].z, fr[1].w};
    const Quat qc = {-q.x, -q.y, -q.z, q.w};
    const V3 lin = row3(fr[2]);
    const V3 omega = row3(fr[3]);
    const float radius = fr[3].w;
    const uint32_t n_tiles = *count;
    for (uint32_t i = blockIdx.x; i < n_tiles; i += gridDim.x) {
        const uint32_t tile = list[i];
        int x0, y0, z;
        tile_origin(p.g, tile, x0, y0, z);
        const int x = x0 + (int)(threadIdx.x & (TILE_X - 1));
        const int y = y0 + (int)(threadIdx.x / TILE_X);
        if (x >= p.g.nx || y >= p.g.ny) continue;
        const uint32_t gid = (uint32_t)x + (uint32_t)p.g.nx * ((uint32_t)y + (uint32_t)p.g.ny * (uint32_t)z);
        const V3 world = v3(p.origin[0] + ((float)x + 0.5f) * p.dx, p.origin[1] + ((float)y + 0.5f) * p.dx,
                            p.origin[2] + ((float)z + 0.5f) * p.dx);
        const V3 rel = vsub(world, body_pos);
        const bool was_solid = mask_prev && mask_prev[gid] != 0;
        if (!tiles_new[tile] || (!was_solid && vdot(rel, rel) > radius * radius)) {
            mask[gid] = 0;
            wall[gid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            continue;
        }
        const V3 local = qrotate(qc, rel);
        const V3 base = vadd(lin, vcross(omega, rel));
        float best = sd_ellipsoid(local, b);
        V3 velocity = base;
        uint8_t part = 1;
#pragma unroll
        for (int wing = 0; wing < 2; ++wing) {
            const float4* f = fr + 4 + 5 * wing;
            const V3 root = row3(f[0]), chord = row3(f[1]), span = row3(f[2]), normal = row3(f[3]);
            const V3 ang = row3(f[4]);
            const float tip_twist = f[1].w, tip_twist_rate = f[4].w;
            const V3 r = vsub(world, root);
            const V3 lu = v3(vdot(r, chord), vdot(r, span), vdot(r, normal));
            const float frac = clamp01(lu.y / smax(b.wing_span, 1.0e-6f));
            float ts, tc;
            bird_sincos(tip_twist * frac, ts, tc);
            const V3 l = v3(tc * lu.x + ts * lu.z, lu.y, -ts * lu.x + tc * lu.z);
            const float d = sd_wing(l, b);
            if (d < best) {
                best = d;
                const V3 twist_axis = vmul(span, frac * tip_twist_rate);
                velocity = vadd(vadd(base, vcross(ang, r)), vcross(twist_axis, vsub(world, vadd(root, vmul(span, lu.y)))));
                part = wing == 0 ? 2 : 3;
            }
        }
        const float tail = sd_tail(local, b);
        if (tail < best) {
            best = tail;
            velocity = base;
            part = 4;
        }
        const bool solid = best <= 0.0f;
        mask[gid] = solid ? part : 0;
        wall[gid] = make_float4(velocity.x * p.velocity_scale, velocity.y * p.velocity_scale,
                                velocity.z * p.velocity_scale, 0.0f);
        if (solid) mark_tiles(p.g, p.periodic, dyn, x, y, z);
    }
}

// Momentum of one prescribed rigid wing about the body position from its
// prepared frame rows (prescribedWingMomentum, BirdFlow.metal:6461-6490):
// centre offset c = chord m_c + span m_s + normal m_n from the hinge,
// P = m (w x c), L = (root + c - body) x P + sum over the axes a of
// a (I_a (w . a)).  Same operations in the same order as the oracle.
__device__ void wing_momentum(const float4 root, const float4 chord, const float4 span, const float4 normal,
                              const float4 omega, float mass, const float com[3], const float inertia[3],
                              const float4 body, float linear[3], float angular[3]) {
    const float ch[3] = {chord.x, chord.y, chord.z}, sp[3] = {span.x, span.y, span.z};
    const float nm[3] = {normal.x, normal.y, normal.z}, rt[3] = {root.x, root.y, root.z};
    const float bd[3] = {body.x, body.y, body.z};
    float off[3], arm[3], spin[3];
    for (int i = 0; i < 3; ++i) off[i] = ch[i] * com[0] + sp[i] * com[1] + nm[i] * com[2];
    const float cv[3] = {omega.y * off[2] - omega.z * off[1], omega.z * off[0] - omega.x * off[2],
                         omega.x * off[1] - omega.y * off[0]};
    for (int i = 0; i < 3; ++i) linear[i] = mass * cv[i];
    const float wc = omega.x * ch[0] + omega.y * ch[1] + omega.z * ch[2];
    const float ws = omega.x * sp[0] + omega.y * sp[1] + omega.z * sp[2];
    const float wn = omega.x * nm[0] + omega.y * nm[1] + omega.z * nm[2];
    for (int i = 0; i < 3; ++i) {
        spin[i] = ch[i] * (inertia[0] * wc) + sp[i] * (inertia[1] * ws) + nm[i] * (inertia[2] * wn);
        arm[i] = rt[i] + off[i] - bd[i];
    }
    const float ax[3] = {arm[1] * linear[2] - arm[2] * linear[1], arm[2] * linear[0] - arm[0] * linear[2],
                         arm[0] * linear[1] - arm[1] * linear[0]};
    for (int i = 0; i < 3; ++i) angular[i] = ax[i] + spin[i];
}

// updateWingInertialReaction (BirdFlow.metal:6493-6572), one lane: the hinge
// reaction -(momentum - previous) / dt of each wing from the step's prepared
// frame (0 on the first frame), and the momenta kept for the next step.
__global__ void wing_inertial_reaction(const float4* __restrict__ frame, WingDynamics d,
                                       float* __restrict__ previous, float* __restrict__ reaction) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float m[4][3];
    // frame rows: 0 body position, 4..8 left root / chord / span / normal /
    // angular velocity, 9..13 the right wing's
    wing_momentum(frame[4], frame[5], frame[6], frame[7], frame[8], d.mass[0], d.center_of_mass[0], d.inertia[0],
                  frame[0], m[0], m[1]);
    wing_momentum(frame[9], frame[10], frame[11], frame[12], frame[13], d.mass[1], d.center_of_mass[1], d.inertia[1],
                  frame[0], m[2], m[3]);
    const float inverse_dt = 1.0f / d.dt;
    for (int r = 0; r < 4; ++r) {
        for (int i = 0; i < 3; ++i) {
            reaction[4 * r + i] = d.initialize ? 0.0f : -(m[r][i] - previous[4 * r + i]) * inverse_dt;
            previous[4 * r + i] = m[r][i];
        }
        reaction[4 * r + 3] = 0.0f;
        previous[4 * r + 3] = 0.0f;
    }
}

hipError_t launch_bird_geometry(const BirdShape& shape, const BirdBuffers& b, const BirdParams& p, hipStream_t s,
                                const WingDynamics* wing_dynamics) {
    const unsigned tile_blocks = (p.g.tiles + 255) / 256;
    hipLaunchKernelGGL(bird_prepare, dim3(1), dim3(64), 0, s, shape, b.body, b.keyframes, (int)b.keyframe_count,
                       p.time, p.dx, b.frame);
    if (wing_dynamics && b.wing_momentum && b.wing_reaction)
        hipLaunchKernelGGL(wing_inertial_reaction, dim3(1), dim3(64), 0, s, (const float4*)b.frame, *wing_dynamics,
                           b.wing_momentum, b.wing_reaction);
    hipLaunchKernelGGL(bird_mark_tiles, dim3(tile_blocks), dim3(256), 0, s, shape, (const float4*)b.frame,
                       b.dyn_prev, b.tiles_new, p);
    hipLaunchKernelGGL(fill_u8, dim3(tile_blocks), dim3(256), 0, s, b.dyn, (uint8_t)0, p.g.tiles);
    hipLaunchKernelGGL(fill_u32, dim3(1), dim3(64), 0, s, b.list_count, 0u, 1u);
    hipLaunchKernelGGL(geometry_tile_list, dim3(tile_blocks), dim3(256), 0, s, b.tiles_new, b.tiles_mask_old,
                       b.tiles_wall_old, p.g.tiles, b.list, b.list_count);
    hipLaunchKernelGGL(bird_tile_pass, dim3(p.tile_pass_blocks), dim3(256), 0, s, shape, (const float4*)b.frame,
                       b.list, b.list_count, (const uint8_t*)b.tiles_new, b.mask_prev, b.mask, b.wall, b.dyn, p);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Prescribed beta-planform flapping wing.  preparePrescribedFlappingWing
// (BirdFlow.metal:1040-1082) in one lane; buildPrescribedFlappingWing
// (:1178-1245) on the tiles within the guard sphere (and near last step's
// solid cells); then the link fractions (:1146-1174, :1271-1296) of every
// fluid cell with a solid source, into a table indexed by tile, direction
// and lane (the reference keeps them in the solid cells' dormant population
// slots, which the next step's geometry may not touch here while the fluid
// step runs).  Conventions shared with the oracle: fmod(a, 1) = a - trunc(a),
// pow(x > 0, y) = exp2(y log2 x) with the log2 / exp2 below, pow(0, y) = 0,
// mix(a, b, t) = a + (b - a) t, sin / cos by bird_sincos.
namespace {

__device__ inline float wing_log2(float x) {
    int k;
    float m = frexpf(x, &k);
    if (m < 0.70710678f) {
        m = m * 2.0f;
        k = k - 1;
    }
    const float t = (m - 1.0f) / (m + 1.0f);
    const float t2 = t * t;
    float s = __builtin_fmaf(t2, 1.0f / 11.0f, 1.0f / 9.0f);
    s = __builtin_fmaf(s, t2, 1.0f / 7.0f);
    s = __builtin_fmaf(s, t2, 1.0f / 5.0f);
    s = __builtin_fmaf(s, t2, 1.0f / 3.0f);
    s = __builtin_fmaf(s, t2, 1.0f);
    const float ln_m = 2.0f * t * s;
    return (float)k + ln_m * 1.44269504f;
}

__device__ inline float wing_exp2(float v) {
    const float n = __builtin_rintf(v);
    const float f = (v - n) * 0.693147181f;
    float e = __builtin_fmaf(f, 1.0f / 40320.0f, 1.0f / 5040.0f);
    e = __builtin_fmaf(e, f, 1.0f / 720.0f);
    e = __builtin_fmaf(e, f, 1.0f / 120.0f);
    e = __builtin_fmaf(e, f, 1.0f / 24.0f);
    e = __builtin_fmaf(e, f, 1.0f / 6.0f);
    e = __builtin_fmaf(e, f, 0.5f);
    e = __builtin_fmaf(e, f, 1.0f);
    e = __builtin_fmaf(e, f, 1.0f);
    return ldexpf(e, (int)n);
}

__device__ inline float wing_pow(float x, float y) {
    if (!(x > 0.0f)) return 0.0f;
    return wing_exp2(y * wing_log2(x));
}

__device__ inline void wing_stroke(float phase, const WingShape& w, float& angle, float& rate) {
    const float pi = 3.14159265358979323846f;
    const float amplitude = w.stroke_amplitude, duration = w.stroke_duration;
    const float half = 0.5f * duration;
    const float rmax = w.stroke_rate_max;
    float s, c;
    if (phase < half) {
        bird_sincos(pi * (phase + half) / duration, s, c);
        angle = amplitude + rmax * duration / pi * (s - 1.0f);
        rate = rmax * c;
    } else if (phase < 0.5f - half) {
        const float end = amplitude - rmax * duration / pi;
        angle = end - rmax * (phase - half);
        rate = -rmax;
    } else if (phase < 0.5f + half) {
        const float start = 0.5f - half;
        const float at = -amplitude + rmax * duration / pi;
        bird_sincos(pi * (phase - start) / duration, s, c);
        angle = at - rmax * duration / pi * s;
        rate = -rmax * c;
    } else if (phase < 1.0f - half) {
        const float end = -amplitude + rmax * duration / pi;
        angle = end + rmax * (phase - (0.5f + half));
        rate = rmax;
    } else {
        const float start = 1.0f - half;
        const float at = amplitude - rmax * duration / pi;
        bird_sincos(pi * (phase - start) / duration, s, c);
        angle = at + rmax * duration / pi * s;
        rate = rmax * c;
    }
}

__device__ inline void wing_pitch(float phase, const WingShape& w, float& angle, float& rate) {
    const float pi = 3.14159265358979323846f;
    const float duration = w.pitch_duration;
    const float half = 0.5f * duration;
    const float low = w.pitch_low, high = w.pitch_high;
    const float delta = high - low;
    float x, start, change;
    if (phase < half || phase >= 1.0f - half) {
        const float wrapped = phase < half ? phase + 1.0f : phase;
        x = (wrapped - (1.0f - half)) / duration;
        start = high;
        change = -delta;
    } else if (phase >= 0.5f - half && phase < 0.5f + half) {
        x = (phase - (0.5f - half)) / duration;
        start = low;
        change = delta;
    } else {
        angle = phase < 0.5f ? low : high;
        rate = 0.0f;
        return;
    }
    float s, c;
    bird_sincos(2.0f * pi * x, s, c);
    const float blend = x - s / (2.0f * pi);
    rate = change / duration * (1.0f - c);
    angle = start + change * blend;
}

// largest violation of the planform's bounds (<= 0 inside)
__device__ inline float wing_violation(const WingShape& w, const float4* f, float dx, V3 world) {
    const V3 rel = vsub(world, row3(f[0]));
    const float radial = vdot(rel, row3(f[2]));
    const float normal = vdot(rel, row3(f[3]));
    const float chord = vdot(rel, row3(f[1]));
    const float radius = w.radius;
    const float chord_mean = f[0].w;
    const float half = 0.5f * smax(w.thickness, dx);
    const float rf = clamp01(radial / radius);
    const float base = smax(rf * (1.0f - rf), 0.0f);
    const float local = chord_mean * w.beta_normalization * wing_pow(base, w.beta_exponent);
    const float lead = -0.25f * chord_mean;
    const float v_root = -radial, v_tip = radial - radius;
    const float v_normal = fabsf(normal) - half;
    const float v_lead = lead - chord, v_trail = chord - (lead + local);
    return smax(smax(v_root, v_tip), smax(v_normal, smax(v_lead, v_trail)));
}

}  // namespace

__global__ void wing_prepare(WingShape w, float time, float4* __restrict__ frame) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const float a = time / w.cycle;
    float phase = a - truncf(a);
    phase = phase < 0.0f ? phase + 1.0f : phase;
    float sa, sr, pa, pr;
    wing_stroke(phase, w, sa, sr);
    wing_pitch(phase, w, pa, pr);
    float ss, sc;
    bird_sincos(sa, ss, sc);
    const V3 axis = v3(0.0f, 0.0f, 1.0f);
    const V3 span = v3(sc, ss, 0.0f);
    const V3 tangent = vcross(axis, span);
    const V3 chord = rotate_axis(tangent, span, -pa);
    const V3 normal = vnormalize(vcross(span, chord));
    const float inv = 1.0f / w.cycle;
    const V3 omega = vsub(vmul(axis, sr * inv), vmul(span, pr * inv));
    frame[0] = make_float4(w.root[0], w.root[1], w.root[2], w.chord_mean);
    frame[1] = row4(vnormalize(chord), 0.0f);
    frame[2] = row4(vnormalize(span), 0.0f);
    frame[3] = row4(normal, 0.0f);
    frame[4] = row4(omega, 0.0f);
    frame[5] = make_float4(sa, pa, sr * inv, pr * inv);
}

// Tiles evaluated this step: within two cells of the guard sphere (radius
// span + 1.5 mean chords + 2 cells), or near last step's solid cells.
__global__ __launch_bounds__(256) void wing_mark_tiles(WingShape w, const float4* __restrict__ frame,
                                                       const uint8_t* __restrict__ dyn_prev,
                                                       uint8_t* __restrict__ tiles_new, BirdParams p) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= p.g.tiles) return;
    int x0, y0, z;
    tile_origin(p.g, t, x0, y0, z);
    const int x1 = min(x0 + TILE_X, p.g.nx) - 1, y1 = min(y0 + TILE_Y, p.g.ny) - 1;
    const float lx = p.origin[0] + ((float)x0 + 0.5f) * p.dx, hx = p.origin[0] + ((float)x1 + 0.5f) * p.dx;
    const float ly = p.origin[1] + ((float)y0 + 0.5f) * p.dx, hy = p.origin[1] + ((float)y1 + 0.5f) * p.dx;
    const float zc = p.origin[2] + ((float)z + 0.5f) * p.dx;
    const float4 c = frame[0];
    bool hit;
    if (w.beta_exponent > 0.0f) {
        // The wing lies in the box chord [lead, lead + c_max], span [0, R],
        // normal [-h, h] of its frame (c_max at mid-span, where r(1 - r) is
        // largest).  A tile (the segment of cell centres [lx, hx] x [ly, hy]
        // at z) is marked unless one of the frame's axes separates it from
        // that box grown by two cells: every wing cell, and every cell that
        // links to one, lies in a marked tile.
        const float m = 2.0f * p.dx;
        const float lead = -0.25f * c.w;
        const float cmax = 1.01f * c.w * w.beta_normalization * wing_pow(0.25f, w.beta_exponent);
        const float h = 0.5f * smax(w.thickness, p.dx);
        const V3 mid = v3(0.5f * (lx + hx) - c.x, 0.5f * (ly + hy) - c.y, zc - c.z);
        const float ex = 0.5f * (hx - lx), ey = 0.5f * (hy - ly);
        const float lo[3] = {lead - m, -m, -h - m};
        const float hi[3] = {lead + cmax + m, w.radius + m, h + m};
        hit = true;
        for (int a = 0; a < 3; ++a) {
            const V3 axis = row3(frame[1 + a]);
            const float centre = vdot(mid, axis);
            const float reach = fabsf(axis.x) * ex + fabsf(axis.y) * ey;
            hit = hit && centre + reach >= lo[a] && centre - reach <= hi[a];
        }
    } else {
        const float r = w.radius + 1.5f * c.w + 2.0f * p.dx + 2.0f * p.dx;
        hit = tile_box_dist2(lx, hx, ly, hy, zc, c.x, c.y, c.z) <= r * r;
    }
    tiles_new[t] = (hit || (dyn_prev && dyn_prev[t])) ? 1 : 0;
}

// buildPrescribedFlappingWing on the marked tiles (every cell the
// reference's value); other listed tiles cleared to the reference's far
// value (mask 0, wall (0, 0, 0, radius)).
__global__ __launch_bounds__(256) void wing_tile_pass(WingShape w, const float4* __restrict__ frame,
                                                      const uint32_t* __restrict__ list,
                                                      const uint32_t* __restrict__ count,
                                                      const uint8_t* __restrict__ tiles_new,
                                                      const uint8_t* __restrict__ mask_prev,
                                                      uint8_t* __restrict__ mask, float4* __restrict__ wall,
                                                      uint8_t* __restrict__ dyn, BirdParams p) {
    __shared__ float4 fr[WING_FRAME_ROWS];
    if (threadIdx.x < WING_FRAME_ROWS) fr[threadIdx.x] = frame[threadIdx.x];
    __syncthreads();
    const V3 root = row3(fr[0]);
    const V3 omega = row3(fr[4]);
    const float chord_mean = fr[0].w;
    const float radius = w.radius + 1.5f * chord_mean + 2.0f * p.dx;
    const uint32_t n_tiles = *count;
    for (uint32_t i = blockIdx.x; i < n_tiles; i += gridDim.x) {
        const uint32_t tile = list[i];
        int x0, y0, z;
        tile_origin(p.g, tile, x0, y0, z);
        const int x = x0 + (int)(threadIdx.x & (TILE_X - 1));
        const int y = y0 + (int)(threadIdx.x / TILE_X);
        if (x >= p.g.nx || y >= p.g.ny) continue;
        const uint32_t gid = (uint32_t)x + (uint32_t)p.g.nx * ((uint32_t)y + (uint32_t)p.g.ny * (uint32_t)z);
        const V3 world = v3(p.origin[0] + ((float)x + 0.5f) * p.dx, p.origin[1] + ((float)y + 0.5f) * p.dx,
                            p.origin[2] + ((float)z + 0.5f) * p.dx);
        const V3 rel = vsub(world, root);
        const bool was_solid = mask_prev && mask_prev[gid] != 0;
        if (!tiles_new[tile] || (!was_solid && vdot(rel, rel) > radius * radius)) {
            mask[gid] = 0;
            wall[gid] = make_float4(0.0f, 0.0f, 0.0f, radius);
            continue;
        }
        const float cc = vdot(rel, row3(fr[1]));
        const float rc = vdot(rel, row3(fr[2]));
        const float nc = vdot(rel, row3(fr[3]));
        const float half = 0.5f * smax(w.thickness, p.dx);
        const bool in_radial = rc >= 0.0f && rc <= w.radius;
        const bool in_normal = fabsf(nc) <= half;
        bool is_wing = false;
        if (in_radial && in_normal) {
            const float rf = clamp01(rc / w.radius);
            const float base = smax(rf * (1.0f - rf), 0.0f);
            const float local = chord_mean * w.beta_normalization * wing_pow(base, w.beta_exponent);
            const float lead = -0.25f * chord_mean;
            const float trail = lead + local;
            is_wing = cc >= lead && cc <= trail;
        }
        const V3 v = vcross(omega, rel);
        mask[gid] = is_wing ? 1 : 0;
        wall[gid] = make_float4(v.x * p.velocity_scale, v.y * p.velocity_scale, v.z * p.velocity_scale,
                                is_wing ? -1.0f : 1.0f);
        if (is_wing) mark_tiles(p.g, p.periodic, dyn, x, y, z);
    }
}

// Link fractions of the fluid cells of the tiles near the wing: for each
// direction whose (in-domain) source is solid, the bisection from the
// source's centre plus c_q dx (the reference's fluid point) to the source.
// Source coordinate of a link along one axis, as the fluid kernel's pull
// treats the faces: periodic wraps, an outlet clamps into the domain, a far
// field has no in-domain source (false).
__device__ inline bool link_source(const int bc[2], int s, int n, int& out) {
    if (s < 0) {
        if (bc[0] == BC_FARFIELD) return false;
        s = bc[0] == BC_PERIODIC ? n - 1 : 0;
    } else if (s >= n) {
        if (bc[1] == BC_FARFIELD) return false;
        s = bc[1] == BC_PERIODIC ? 0 : n - 1;
    }
    out = s;
    return true;
}

__global__ __launch_bounds__(256) void wing_links(WingShape w, const float4* __restrict__ frame,
                                                  const uint32_t* __restrict__ list,
                                                  const uint32_t* __restrict__ count,
                                                  const uint8_t* __restrict__ dyn, const uint8_t* __restrict__ mask,
                                                  float* __restrict__ links, BirdParams p) {
    __shared__ float4 fr[WING_FRAME_ROWS];
    if (threadIdx.x < WING_FRAME_ROWS) fr[threadIdx.x] = frame[threadIdx.x];
    __syncthreads();
    // Tiles with wing cells or their neighbours (dyn) are marked tiles, so
    // they are on the step's geometry list.
    const uint32_t n_tiles = *count;
    for (uint32_t i = blockIdx.x; i < n_tiles; i += gridDim.x) {
        const uint32_t tile = list[i];
        if (!dyn[tile]) continue;
        int x0, y0, z;
        tile_origin(p.g, tile, x0, y0, z);
        const int x = x0 + (int)(threadIdx.x & (TILE_X - 1));
        const int y = y0 + (int)(threadIdx.x / TILE_X);
        if (x >= p.g.nx || y >= p.g.ny) continue;
        const uint32_t gid = (uint32_t)x + (uint32_t)p.g.nx * ((uint32_t)y + (uint32_t)p.g.ny * (uint32_t)z);
        if (mask[gid] != 0) continue;
        float* out = links + (size_t)tile * 19 * TILE_CELLS + threadIdx.x;
        for (int q = 1; q < 19; ++q) {
            // the source the fluid kernel pulls from: wrapped across a
            // periodic face, clamped at an outlet, none beyond a far field
            int sx, sy, sz;
            if (!link_source(p.bc[0], x - cx(q), p.g.nx, sx) || !link_source(p.bc[1], y - cy(q), p.g.ny, sy) ||
                !link_source(p.bc[2], z - cz(q), p.g.nz, sz))
                continue;
            const uint32_t src = (uint32_t)sx + (uint32_t)p.g.nx * ((uint32_t)sy + (uint32_t)p.g.ny * (uint32_t)sz);
            if (mask[src] == 0) continue;
            const V3 solid = v3(p.origin[0] + ((float)sx + 0.5f) * p.dx, p.origin[1] + ((float)sy + 0.5f) * p.dx,
                                p.origin[2] + ((float)sz + 0.5f) * p.dx);
            const V3 fluid = v3(solid.x + (float)cx(q) * p.dx, solid.y + (float)cy(q) * p.dx,
                                solid.z + (float)cz(q) * p.dx);
            float lower = 0.0f, upper = 1.0f;
            for (int it = 0; it < 10; ++it) {
                const float mid = 0.5f * (lower + upper);
                const V3 sample = v3(fluid.x + (solid.x - fluid.x) * mid, fluid.y + (solid.y - fluid.y) * mid,
                                     fluid.z + (solid.z - fluid.z) * mid);
                if (wing_violation(w, fr, p.dx, sample) > 0.0f) lower = mid;
                else upper = mid;
            }
            out[(size_t)q * TILE_CELLS] = smin(smax(0.5f * (lower + upper), 1.0e-4f), 1.0f);
        }
    }
}

hipError_t launch_wing_geometry(const WingShape& shape, const WingBuffers& b, const BirdParams& p, hipStream_t s) {
    const unsigned tile_blocks = (p.g.tiles + 255) / 256;
    hipLaunchKernelGGL(wing_prepare, dim3(1), dim3(64), 0, s, shape, p.time, b.frame);
    hipLaunchKernelGGL(wing_mark_tiles, dim3(tile_blocks), dim3(256), 0, s, shape, (const float4*)b.frame, b.dyn_prev,
                       b.tiles_new, p);
    hipLaunchKernelGGL(fill_u8, dim3(tile_blocks), dim3(256), 0, s, b.dyn, (uint8_t)0, p.g.tiles);
    hipLaunchKernelGGL(fill_u32, dim3(1), dim3(64), 0, s, b.list_count, 0u, 1u);
    hipLaunchKernelGGL(geometry_tile_list, dim3(tile_blocks), dim3(256), 0, s, b.tiles_new, b.tiles_mask_old,
                       b.tiles_wall_old, p.g.tiles, b.list, b.list_count);
    hipLaunchKernelGGL(wing_tile_pass, dim3(p.tile_pass_blocks), dim3(256), 0, s, shape, (const float4*)b.frame,
                       b.list, b.list_count, (const uint8_t*)b.tiles_new, b.mask_prev, b.mask, b.wall, b.dyn, p);
    hipLaunchKernelGGL(wing_links, dim3(p.tile_pass_blocks), dim3(256), 0, s, shape, (const float4*)b.frame,
                       (const uint32_t*)b.list, (const uint32_t*)b.list_count, (const uint8_t*)b.dyn,
                       (const uint8_t*)b.mask, b.links, p);
    return hipGetLastError();
}

hipError_t launch_mesh_static_keys(const MeshBuffers& b, const MeshParams& m, unsigned long long* count,
                                   hipStream_t s) {
    if (m.g.nx > 65535 || m.g.ny > 65535) return hipErrorInvalidValue;   // raster packs x | y << 16
    if (m.raster_batch < 1 || m.raster_batch > 64 || m.raster_blocks < 1 || m.raster_split < 1 || m.raster_split > 64)
        return hipErrorInvalidValue;
    if (b.kinematics)
        hipLaunchKernelGGL(articulated_prepare, dim3((m.vertex_count + 255) / 256), dim3(256), 0, s, b.rest,
                           b.vertex_joint, b.span_fraction, b.kinematics, b.pos, b.vel, b.body, m);
    else
        hipLaunchKernelGGL(mesh_prepare, dim3((m.vertex_count + 255) / 256), dim3(256), 0, s, b.points, b.times,
                           b.pos, b.vel, b.body, m);
    hipLaunchKernelGGL(fill_u8, dim3((m.g.tiles + 255) / 256), dim3(256), 0, s, b.tiles_new, (uint8_t)0, m.g.tiles);
    if (m.raster_count > 0)
        hipLaunchKernelGGL(mesh_raster, dim3(std::min((m.raster_count + m.raster_batch - 1) / m.raster_batch, m.raster_blocks), m.raster_split), dim3(64), 0, s, b.pos,
                           b.tri, b.raster_list, b.keys, b.tiles_new, m);
    hipLaunchKernelGGL(fill_u64, dim3(1), dim3(64), 0, s, (uint64_t*)count, (uint64_t)0, 1u);
    hipLaunchKernelGGL(collect_keys, dim3((m.g.n + 255) / 256), dim3(256), 0, s, b.keys, m.g.n, (uint32_t*)nullptr,
                       (unsigned long long*)nullptr, 0u, count);
    return hipGetLastError();
}

hipError_t launch_collect_keys(unsigned long long* keys, uint32_t n, uint32_t* cells, unsigned long long* out,
                               uint32_t capacity, unsigned long long* counter, hipStream_t s) {
    hipLaunchKernelGGL(fill_u64, dim3(1), dim3(64), 0, s, (uint64_t*)counter, (uint64_t)0, 1u);
    hipLaunchKernelGGL(collect_keys, dim3((n + 255) / 256), dim3(256), 0, s, keys, n, cells, out, capacity, counter);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Articulated mesh: vertices on rigid joints driven by Fourier series
// (bf_mesh_upload_articulated).  Per joint: stroke, deviation, pitch and
// twist angles and their rates at the query time; per vertex: the rest
// position mirrored by the joint's side, pitch + span-weighted twist about
// y, deviation about z, stroke about x, the hinge and the mesh centre; the
// velocity is the time derivative.  FP32 with bird_sincos, in the oracle's
// order (oracle/bfref_kin.inc): bit-identical positions and velocities.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void kin_series(const float (&c)[4][2], uint32_t harmonics, const float* sh,
                                           const float* ch, float omega, float scale, float& angle, float& rate) {
    float a = 0.0f, r = 0.0f;
    for (uint32_t h = 0; h < harmonics; ++h) {
        a = a + (c[h][0] * sh[h] + c[h][1] * ch[h]);
        r = r + (float)(h + 1) * (c[h][0] * ch[h] - c[h][1] * sh[h]);
    }
    angle = a * scale;
    rate = (r * omega) * scale;
}

__global__ void articulated_prepare(const float* __restrict__ rest, const uint8_t* __restrict__ vertex_joint,
                                    const float* __restrict__ span_fraction, const Kinematics* __restrict__ kin,
                                    float4* __restrict__ pos, float4* __restrict__ vel,
                                    const double* __restrict__ body, MeshParams m) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= m.vertex_count) return;
    const Kinematics& k = *kin;
    const float* r = rest + 3 * (size_t)v;
    const uint32_t jv = vertex_joint[v];
    float lp[3], lv[3];
    if (jv == 0 || jv > k.joint_count) {
        for (int i = 0; i < 3; ++i) {
            lp[i] = k.center[i] + r[i];
            lv[i] = 0.0f;
        }
    } else {
        const KinJoint& J = k.joints[jv - 1];
        const uint32_t harmonics = min(k.harmonics, 4u);
        const float t = m.query_time;
        const float period = k.period;
        const float cycles = floorf(t / period);
        const float tm = t - cycles * period;
        const float omega = 6.28318548f / period;
        const float phi = omega * tm;
        float sh[4] = {0.0f, 0.0f, 0.0f, 0.0f}, ch[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (uint32_t h = 0; h < harmonics; ++h) bird_sincos((float)(h + 1) * phi, sh[h], ch[h]);
        float ang[4], rate[4];
        kin_series(J.stroke, harmonics, sh, ch, omega, k.amplitude_scale, ang[0], rate[0]);
        kin_series(J.deviation, harmonics, sh, ch, omega, k.amplitude_scale, ang[1], rate[1]);
        kin_series(J.pitch, harmonics, sh, ch, omega, k.amplitude_scale, ang[2], rate[2]);
        kin_series(J.twist, harmonics, sh, ch, omega, k.amplitude_scale, ang[3], rate[3]);
        const float side = J.side;
        float sa, ca, sb, cb;
        bird_sincos(side * ang[0], sa, ca);
        bird_sincos(side * ang[1], sb, cb);
        const float l[3] = {r[0], side * r[1], r[2]};
        const float span = span_fraction[v];
        const float a = ang[2] + ang[3] * span;
        const float adot = rate[2] + rate[3] * span;
        float sp, cp;
        bird_sincos(a, sp, cp);
        const float u3[3] = {cp * l[0] + sp * l[2], l[1], -sp * l[0] + cp * l[2]};
        const float u2[3] = {cb * u3[0] - sb * u3[1], sb * u3[0] + cb * u3[1], u3[2]};
        const float u1[3] = {u2[0], ca * u2[1] - sa * u2[2], sa * u2[1] + ca * u2[2]};
        const float root[3] = {J.root[0], side * J.root[1], J.root[2]};
        for (int i = 0; i < 3; ++i) lp[i] = (k.center[i] + root[i]) + u1[i];
        const float ad = side * rate[0], bd = side * rate[1];
        const float t1[3] = {0.0f, -(ad * u1[2]), ad * u1[1]};
        const float w2[3] = {-(bd * u2[1]), bd * u2[0], 0.0f};
        const float t2[3] = {w2[0], ca * w2[1] - sa * w2[2], sa * w2[1] + ca * w2[2]};
        const float w3[3] = {adot * u3[2], 0.0f, -(adot * u3[0])};
        const float z3[3] = {cb * w3[0] - sb * w3[1], sb * w3[0] + cb * w3[1], w3[2]};
        const float t3[3] = {z3[0], ca * z3[1] - sa * z3[2], sa * z3[1] + ca * z3[2]};
        for (int i = 0; i < 3; ++i) lv[i] = (t1[i] + t2[i]) + t3[i];
    }
    place_vertex(v, lp, lv, body, m, pos, vel);
}

hipError_t launch_mesh_geometry(const MeshBuffers& b, const MeshParams& m, hipStream_t s) {
    if (m.g.nx > 65535 || m.g.ny > 65535) return hipErrorInvalidValue;   // raster packs x | y << 16
    if (m.raster_batch < 1 || m.raster_batch > 64 || m.raster_blocks < 1 || m.raster_split < 1 || m.raster_split > 64)
        return hipErrorInvalidValue;
    if (b.kinematics)
        hipLaunchKernelGGL(articulated_prepare, dim3((m.vertex_count + 255) / 256), dim3(256), 0, s, b.rest,
                           b.vertex_joint, b.span_fraction, b.kinematics, b.pos, b.vel, b.body, m);
    else
        hipLaunchKernelGGL(mesh_prepare, dim3((m.vertex_count + 255) / 256), dim3(256), 0, s, b.points, b.times,
                           b.pos, b.vel, b.body, m);
    hipLaunchKernelGGL(fill_u8, dim3((m.g.tiles + 255) / 256), dim3(256), 0, s, b.tiles_new, (uint8_t)0, m.g.tiles);
    hipLaunchKernelGGL(fill_u8, dim3((m.g.tiles + 255) / 256), dim3(256), 0, s, b.dyn, (uint8_t)0, m.g.tiles);
    hipLaunchKernelGGL(fill_u32, dim3(1), dim3(64), 0, s, b.list_count, 0u, 1u);
    if (b.seed_count > 0)
        hipLaunchKernelGGL(seed_keys, dim3((b.seed_count + 255) / 256), dim3(256), 0, s, b.seed_cells, b.seed_keys,
                           b.seed_count, b.keys, b.tiles_new, m.g);
    if (m.raster_count > 0)
        hipLaunchKernelGGL(mesh_raster, dim3(std::min((m.raster_count + m.raster_batch - 1) / m.raster_batch, m.raster_blocks), m.raster_split), dim3(64), 0, s, b.pos,
                           b.tri, b.raster_list, b.keys, b.tiles_new, m);
    hipLaunchKernelGGL(geometry_tile_list, dim3((m.g.tiles + 255) / 256), dim3(256), 0, s, b.tiles_new,
                       b.tiles_mask_old, b.tiles_wall_old, m.g.tiles, b.list, b.list_count);
    hipLaunchKernelGGL(geometry_tile_pass, dim3(m.tile_pass_blocks), dim3(256), 0, s, b.pos, b.vel, b.tri, b.tri_part, b.keys, b.list,
                       b.list_count, b.tiles_new, b.tiles_mask_old, b.tiles_wall_old, b.mask, b.wall, b.dyn, b.tri_out,
                       m);
    return hipGetLastError();
}

hipError_t launch_sphere_geometry(uint8_t* mask, float4* wall, uint8_t* dyn, const Grid& g, float4 cr, float3 gv,
                                  float4 wv, float t, const int periodic[3], hipStream_t s) {
    hipLaunchKernelGGL(fill_u8, dim3((g.tiles + 255) / 256), dim3(256), 0, s, dyn, (uint8_t)0, g.tiles);
    hipLaunchKernelGGL(sphere_geometry, dim3((g.n + 255) / 256), dim3(256), 0, s, mask, wall, dyn, g, cr, gv, wv, t,
                       periodic[0], periodic[1], periodic[2]);
    return hipGetLastError();
}

hipError_t launch_mark_tiles(const uint8_t* mask, uint8_t* dyn, const Grid& g, const int periodic[3], bool clear,
                             hipStream_t s) {
    if (clear) hipLaunchKernelGGL(fill_u8, dim3((g.tiles + 255) / 256), dim3(256), 0, s, dyn, (uint8_t)0, g.tiles);
    hipLaunchKernelGGL(mark_tiles_from_mask, dim3((g.n + 255) / 256), dim3(256), 0, s, mask, dyn, g, periodic[0],
                       periodic[1], periodic[2]);
    return hipGetLastError();
}

hipError_t launch_fill_u8(uint8_t* p, uint8_t v, uint32_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(fill_u8, dim3((n + 255) / 256), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, uint32_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(fill_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_fill_keys(unsigned long long* p, uint32_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(fill_u64, dim3((n + 255) / 256), dim3(256), 0, s, (uint64_t*)p, EMPTY_KEY, n);
    return hipGetLastError();
}

hipError_t launch_fill_wall(float4* p, uint32_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(fill_wall, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_link_census(const uint8_t* prev, const uint8_t* cur, const Grid& g, const int bc[3][2],
                              unsigned long long* counts, hipStream_t s) {
    hipLaunchKernelGGL(link_census, dim3((g.n + 255) / 256), dim3(256), 0, s, prev, cur, g, bc[0][0], bc[0][1],
                       bc[1][0], bc[1][1], bc[2][0], bc[2][1], counts);
    return hipGetLastError();
}

}  // namespace bf
