// Exact link fractions against a meshed shell: the wall rule
// BF_WALL_INTERPOLATED_RAY.
//
// The reference's production q for meshes is the linear root of the voxel
// signed-distance field along a link (BirdFlow.metal:2044-2076); its exact
// roots exist only as CPU audits that scan and bisect the distance to the
// indexed surface along each link (MetalIndexedBirdSurfaceValidation.swift:
// 14245-14444, archived in deetjen-dove-moving-wall-link-ray-root.json).
// This pass computes that root in closed form.  The solid is the union over
// triangles of the shell {p : dist(p, triangle) <= h}; each shell is the
// triangle swept by a ball of radius h, bounded by two offset faces, three
// edge cylinders and three vertex spheres.  Along the link from the fluid
// node F toward its solid source S (F = S + c dx, D = -c dx), the first
// entry into one shell is the least t >= 0 over
//   offset faces:  (n.(F - A) + t n.D) = +-h, entry point inside the face;
//   cylinders:     |(P - U) x e|^2 = h^2 |e|^2 (a quadratic in t), foot of
//                  the entry point on the edge;
//   spheres:       |P - V|^2 = h^2 (a quadratic in t);
// and F inside any of them enters at 0.  The link fraction is the least
// entry over the mesh when it lies within the link (t <= 1); otherwise the
// reference's signed-distance ratio; clamped to [1e-4, 1] like the table the
// step reads (BirdFlow.metal:2738-2745).  FP32 throughout, no contraction,
// expression for expression the oracle's bfo_ray_shell / bfo_ray_links:
// bit-identical tables.
//
// Three passes per step, after the mesh tile pass (mask, wall and the
// solid-proximity tile flags of the step are final):
//  1. init: every cell of a tile flagged as pulling from a solid cell writes
//     its 19 table entries: 2 (no entry yet) for a fluid cell's link whose
//     source is solid, 0 otherwise, and sets bit q of the source's link
//     directions;
//  2. triangles: one wavefront per triangle (grid-stride; a mesh of few
//     large triangles splits each triangle's cells over several
//     wavefronts).  Lanes take the cells S of the triangle's shell box grown
//     by one cell; each link direction of S whose segment box meets the
//     shell box (the oracle's rejection test, the same floats) is queued in
//     LDS, and the queue is spread over the lanes: the entry t is computed
//     and, when t <= 1, merged into every fluid cell that pulls from S along
//     that direction (periodic wrap and outlet clamps included) with a
//     32-bit atomicMin on the bits of max(t, 1e-4) (non-negative floats order
//     as their bits).  The minimum over the mesh does not depend on the
//     order of the atomics;
//  3. finalize: each link entry becomes the fraction (the entry when <= 1,
//     else the signed-distance ratio), clamped; the link directions are
//     cleared for the next step.
// For the adjoint (RayBuffers::winners) init also clears a winner per entry
// and a second triangles pass, before finalize, stores the triangle whose
// entry equals the finished minimum: the adjoint differentiates that
// triangle's root.
// Rejection by the shell boxes never drops a triangle whose entry lies in
// the link: an entry point lies inside its triangle's shell box, so the
// segment box and the shell box share it.
#include "bf_device.h"
#include "bf_launch.h"
#include "bf_rays.h"

namespace bf {

namespace {

__device__ __forceinline__ float rdot(const float a[3], const float b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ void rcross(const float a[3], const float b[3], float o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ void rsub(const float a[3], const float b[3], float o[3]) {
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}
// P on the inner side of edge (U, V) of a triangle with normal n
__device__ __forceinline__ bool edge_in(const float U[3], const float V[3], const float P[3], const float n[3]) {
    float e[3], m[3], c[3];
    rsub(V, U, e);
    rsub(P, U, m);
    rcross(e, m, c);
    return rdot(c, n) >= 0.0f;
}

// First entry t >= 0 of the ray F + t D into the shell of radius h around
// triangle ABC; 0 when F is inside it; 2 when no entry lies below 2.
__device__ float ray_shell(const float F[3], const float D[3], const float A[3], const float B[3], const float C[3],
                           float h) {
    float best = 2.0f;
    float e1[3], e2[3], n[3];
    rsub(B, A, e1);
    rsub(C, A, e2);
    rcross(e1, e2, n);
    const float nl = sqrtf(rdot(n, n));
    if (nl > 0.0f) {
        for (int i = 0; i < 3; ++i) n[i] = n[i] / nl;
        float fa[3];
        rsub(F, A, fa);
        const float fn = rdot(fa, n);
        if (fn <= h && fn >= -h && edge_in(A, B, F, n) && edge_in(B, C, F, n) && edge_in(C, A, F, n)) return 0.0f;
        const float dn = rdot(D, n);
        if (dn != 0.0f) {
            for (int side = 0; side < 2; ++side) {
                const float off = side == 0 ? h : -h;
                const float t = (off - fn) / dn;
                if (t >= 0.0f && t < best) {
                    const float P[3] = {F[0] + D[0] * t, F[1] + D[1] * t, F[2] + D[2] * t};
                    if (edge_in(A, B, P, n) && edge_in(B, C, P, n) && edge_in(C, A, P, n)) best = t;
                }
            }
        }
    }
    const float dd = rdot(D, D);
    for (int k = 0; k < 3; ++k) {
        const float* u = k == 0 ? A : (k == 1 ? B : C);
        const float* w = k == 0 ? B : (k == 1 ? C : A);
        float e[3], m[3];
        rsub(w, u, e);
        rsub(F, u, m);
        const float ee = rdot(e, e), de = rdot(D, e), me = rdot(m, e), md = rdot(m, D), mm = rdot(m, m);
        const float a = ee * dd - de * de;
        const float hb = ee * md - me * de;
        const float c = (ee * mm - me * me) - (h * h) * ee;
        if (c <= 0.0f && me >= 0.0f && me <= ee) return 0.0f;
        if (a > 0.0f) {
            const float disc = hb * hb - a * c;
            if (disc >= 0.0f) {
                const float t = (-hb - sqrtf(disc)) / a;
                if (t >= 0.0f && t < best) {
                    const float uu = (me + t * de) / ee;
                    if (uu >= 0.0f && uu <= 1.0f) best = t;
                }
            }
        }
    }
    for (int k = 0; k < 3; ++k) {
        const float* v = k == 0 ? A : (k == 1 ? B : C);
        float m[3];
        rsub(F, v, m);
        const float md = rdot(m, D);
        const float c = rdot(m, m) - h * h;
        if (c <= 0.0f) return 0.0f;
        const float disc = md * md - dd * c;
        if (disc >= 0.0f) {
            const float t = (-md - sqrtf(disc)) / dd;
            if (t >= 0.0f && t < best) best = t;
        }
    }
    return best;
}

__device__ __forceinline__ uint32_t lane_rank64(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

constexpr uint32_t NO_ENTRY = 0x40000000u;   // 2.0f

}  // namespace

__global__ void ray_clear_count(uint32_t* __restrict__ count) {
    if (threadIdx.x == 0) *count = 0u;
}

// Tiles whose cells pull from a solid cell this step.
__global__ __launch_bounds__(256) void ray_tile_list(const uint8_t* __restrict__ dyn, uint32_t tiles,
                                                     uint32_t* __restrict__ list, uint32_t* __restrict__ count) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool flagged = t < tiles && dyn[t] != 0;
    const unsigned long long b = __ballot(flagged);
    if (b == 0) return;
    uint32_t base = 0;
    if ((threadIdx.x & 63) == 0) base = atomicAdd(count, (uint32_t)__popcll(b));
    base = __shfl(base, 0, 64);
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
    if (flagged) list[base + rank] = t;
}

// Source of direction q of cell (x, y, z) under the face treatment; false
// beyond a far-field face.
__device__ __forceinline__ bool ray_source_cell(const RayParams& p, int x, int y, int z, int q, uint32_t& src) {
    int sx, sy, sz;
    if (!ray_source(p.bc[0], x - cx(q), p.g.nx, sx) || !ray_source(p.bc[1], y - cy(q), p.g.ny, sy) ||
        !ray_source(p.bc[2], z - cz(q), p.g.nz, sz))
        return false;
    src = (uint32_t)sx + (uint32_t)p.g.nx * ((uint32_t)sy + (uint32_t)p.g.ny * (uint32_t)sz);
    return true;
}

// 1. the table entries of the listed tiles: 2 on a fluid cell's links whose
// source is solid, 0 elsewhere; each link sets bit q of its source's link
// directions (`dirs`), which the triangle pass reads to skip solid cells no
// fluid cell pulls from.  The 18 sources' part ids are loaded together.
__global__ __launch_bounds__(256) void ray_init(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count,
                                                const uint8_t* __restrict__ mask, float* __restrict__ links,
                                                uint32_t* __restrict__ dirs, uint32_t* __restrict__ winners,
                                                RayParams p) {
    const uint32_t n_tiles = *count;
    for (uint32_t i = blockIdx.x; i < n_tiles; i += gridDim.x) {
        const uint32_t tile = list[i];
        int x0, y0, z;
        tile_origin(p.g, tile, x0, y0, z);
        const int x = x0 + (int)(threadIdx.x & (TILE_X - 1));
        const int y = y0 + (int)(threadIdx.x / TILE_X);
        if (x >= p.g.nx || y >= p.g.ny) continue;
        const uint32_t gid = (uint32_t)x + (uint32_t)p.g.nx * ((uint32_t)y + (uint32_t)p.g.ny * (uint32_t)z);
        uint32_t src[19];
        uint32_t okbits = 0;
#pragma unroll
        for (int q = 1; q < 19; ++q) {
            uint32_t sidx = gid;
            if (ray_source_cell(p, x, y, z, q, sidx)) okbits |= 1u << q;
            src[q] = sidx;
        }
        const uint8_t own = mask[gid];
        uint8_t part[19];
#pragma unroll
        for (int q = 1; q < 19; ++q) part[q] = mask[src[q]];
        uint32_t* out = (uint32_t*)links + (size_t)tile * 19 * TILE_CELLS + threadIdx.x;
        out[0] = 0u;
#pragma unroll
        for (int q = 1; q < 19; ++q) {
            const bool link = own == 0 && ((okbits >> q) & 1u) && part[q] != 0;
            out[(size_t)q * TILE_CELLS] = link ? NO_ENTRY : 0u;
            if (link) (void)atomicOr(&dirs[src[q]], 1u << q);
        }
        if (winners) {
            uint32_t* w = winners + (size_t)tile * 19 * TILE_CELLS + threadIdx.x;
#pragma unroll
            for (int q = 0; q < 19; ++q) w[(size_t)q * TILE_CELLS] = NO_WINNER;
        }
    }
}

// 2. entries of every triangle's shell, merged by atomicMin.  One
// single-wavefront workgroup per triangle (grid-stride); grid y splits a
// large triangle's cells.  Lanes take the cells of the triangle's shell box
// grown by one cell; a solid cell with link directions queues each
// direction whose segment box meets the shell box as a (cell, q) pair in
// LDS; the queued pairs are then spread over the lanes for the entry
// computation, so that every lane has work.
constexpr uint32_t RAY_CAP = 256;

// Modes: RAY_MIN merges entries into the step's table; RAY_WINNERS, a second
// pass over the finished minima, records for each entry a triangle whose
// shell gives it (entries equal in bits; any of several equal ones);
// RAY_STATIC (once per mesh, over the static triangles) merges every
// direction of every source cell, linked or not, into a per-source table
// [19][N] (the static cache).  list: the triangles to walk (NULL: all).
enum { RAY_MIN = 0, RAY_WINNERS = 1, RAY_STATIC = 2 };
template <int MODE>
__global__ __launch_bounds__(64) void ray_triangles(const float4* __restrict__ pos, const uint32_t* __restrict__ tri,
                                                    const uint32_t* __restrict__ list, uint32_t list_count,
                                                    const uint8_t* __restrict__ mask,
                                                    const uint32_t* __restrict__ dirs, float* __restrict__ links,
                                                    uint32_t* __restrict__ winners, RayParams p) {
    __shared__ uint2 cand[RAY_CAP];
    const uint32_t lane = threadIdx.x;
    const float dx = p.dx;
    const float h = p.half_thickness;
    const int n[3] = {p.g.nx, p.g.ny, p.g.nz};
    uint32_t* table = (uint32_t*)links;
    for (uint32_t i = blockIdx.x; i < list_count; i += gridDim.x) {
        const uint32_t t = list ? list[i] : i;
        const float4 a4 = pos[tri[3 * (size_t)t]], b4 = pos[tri[3 * (size_t)t + 1]], c4 = pos[tri[3 * (size_t)t + 2]];
        const float A[3] = {a4.x, a4.y, a4.z}, B[3] = {b4.x, b4.y, b4.z}, C[3] = {c4.x, c4.y, c4.z};
        // the shell's box (the oracle's rejection box)
        float bl[3], bh[3];
        int lo[3], ext[3];
        for (int a = 0; a < 3; ++a) {
            float l = A[a], u = A[a];
            l = B[a] < l ? B[a] : l;
            u = B[a] > u ? B[a] : u;
            l = C[a] < l ? C[a] : l;
            u = C[a] > u ? C[a] : u;
            bl[a] = l - h - 0.01f * dx;
            bh[a] = u + h + 0.01f * dx;
            // source cells whose segment box (within one cell of the source
            // centre) can meet it: centres in [bl - dx, bh + dx], widened by
            // 0.01 cell for the rounding of the centres
            const float sl = ((bl[a] - dx) - p.origin[a]) / dx - 0.5f;
            const float su = ((bh[a] + dx) - p.origin[a]) / dx - 0.5f;
            const float top = (float)n[a] + 2.0f;
            const int il = (int)floorf(fminf(fmaxf(sl - 0.01f, -2.0f), top));
            const int iu = (int)ceilf(fminf(fmaxf(su + 0.01f, -2.0f), top));
            lo[a] = max(il, 0);
            ext[a] = max(min(iu, n[a] - 1) - lo[a] + 1, 0);
        }
        const uint32_t cells = (uint32_t)ext[0] * (uint32_t)ext[1] * (uint32_t)ext[2];
        uint32_t count = 0;   // wave-uniform
        auto flush = [&]() {
            __syncthreads();
            for (uint32_t k = lane; k < count; k += 64) {
                const uint2 e = cand[k];
                const int sx = (int)(e.x & 0xFFFFu), sy = (int)(e.x >> 16), sz = (int)(e.y >> 5);
                const int q = (int)(e.y & 31u);
                const int c[3] = {cx(q), cy(q), cz(q)};
                const float S[3] = {p.origin[0] + ((float)sx + 0.5f) * dx, p.origin[1] + ((float)sy + 0.5f) * dx,
                                    p.origin[2] + ((float)sz + 0.5f) * dx};
                const float F[3] = {S[0] + (float)c[0] * dx, S[1] + (float)c[1] * dx, S[2] + (float)c[2] * dx};
                const float D[3] = {(float)(-c[0]) * dx, (float)(-c[1]) * dx, (float)(-c[2]) * dx};
                const float v = ray_shell(F, D, A, B, C, h);
                if (!(v <= 1.0f)) continue;
                // entries below the table's floor 1e-4 are stored as the
                // floor (the minimum of the floored entries is the floored
                // minimum; a link entry is then never 0, the init value of
                // the entries that are not links)
                const uint32_t bits = __float_as_uint(fmaxf(v, 1.0e-4f));
                if constexpr (MODE == RAY_STATIC) {
                    const uint32_t src =
                        (uint32_t)sx + (uint32_t)n[0] * ((uint32_t)sy + (uint32_t)n[1] * (uint32_t)sz);
                    (void)atomicMin(&table[(size_t)q * p.g.n + src], bits);
                    continue;
                }
                // the fluid cells pulling from S along c (up to 8 with wrap
                // and clamps)
                int fx[2], fy[2], fz[2];
                const int kx = pullers(p.bc[0], sx, c[0], n[0], fx);
                const int ky = pullers(p.bc[1], sy, c[1], n[1], fy);
                const int kz = pullers(p.bc[2], sz, c[2], n[2], fz);
                for (int iz = 0; iz < kz; ++iz)
                    for (int iy = 0; iy < ky; ++iy)
                        for (int ix = 0; ix < kx; ++ix) {
                            const uint32_t gid =
                                (uint32_t)fx[ix] + (uint32_t)n[0] * ((uint32_t)fy[iy] + (uint32_t)n[1] * (uint32_t)fz[iz]);
                            if (mask[gid] == 0) {
                                const size_t slot = table_slot(p.g, fx[ix], fy[iy], fz[iz], q);
                                if constexpr (MODE == RAY_WINNERS) {
                                    if (table[slot] == bits) winners[slot] = t;
                                } else {
                                    (void)atomicMin(&table[slot], bits);
                                }
                            }
                        }
            }
            __syncthreads();
            count = 0;
        };
        for (uint32_t k0 = blockIdx.y * 64u; k0 < cells; k0 += gridDim.y * 64u) {
            const uint32_t k = k0 + lane;
            uint32_t bits = 0u;
            int sx = 0, sy = 0, sz = 0;
            if (k < cells) {
                sx = lo[0] + (int)(k % (uint32_t)ext[0]);
                const uint32_t r = k / (uint32_t)ext[0];
                sy = lo[1] + (int)(r % (uint32_t)ext[1]);
                sz = lo[2] + (int)(r / (uint32_t)ext[1]);
                bits = MODE == RAY_STATIC
                           ? 0x7FFFEu
                           : dirs[(uint32_t)sx + (uint32_t)n[0] * ((uint32_t)sy + (uint32_t)n[1] * (uint32_t)sz)];
            }
            float S[3] = {0.0f, 0.0f, 0.0f};
            if (bits) {
                S[0] = p.origin[0] + ((float)sx + 0.5f) * dx;
                S[1] = p.origin[1] + ((float)sy + 0.5f) * dx;
                S[2] = p.origin[2] + ((float)sz + 0.5f) * dx;
            }
            // one direction per lane per round, lowest bit first
            while (__ballot(bits != 0u)) {
                bool pass = false;
                int q = 0;
                if (bits) {
                    q = __ffs(bits) - 1;
                    bits &= bits - 1u;
                    const float F[3] = {S[0] + (float)cx(q) * dx, S[1] + (float)cy(q) * dx, S[2] + (float)cz(q) * dx};
                    pass = true;
                    for (int a = 0; a < 3; ++a) {
                        const float l = F[a] < S[a] ? F[a] : S[a];
                        const float u = F[a] < S[a] ? S[a] : F[a];
                        pass = pass && !(u < bl[a] || l > bh[a]);
                    }
                }
                const unsigned long long ballot = __ballot(pass);
                if (pass)
                    cand[count + lane_rank64(ballot)] =
                        make_uint2((uint32_t)sx | ((uint32_t)sy << 16), ((uint32_t)sz << 5) | (uint32_t)q);
                count += (uint32_t)__popcll(ballot);
                if (count > RAY_CAP - 64) flush();
            }
        }
        if (count > 0) flush();
    }
}

// Static cache, built once per mesh: the per-source table's entries (a
// triangle's shell entered within the link) as (source cell, q, bits);
// order of the list is free (the step merges them by minimum).
__global__ __launch_bounds__(256) void ray_static_compact(const uint32_t* __restrict__ table, uint64_t n,
                                                          uint4* __restrict__ out, uint32_t capacity,
                                                          uint32_t* __restrict__ count) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 19ull * n) return;
    const uint32_t v = table[i];
    if (v == NO_ENTRY) return;
    const uint32_t k = atomicAdd(count, 1u);
    if (k < capacity) out[k] = make_uint4((uint32_t)(i % n), (uint32_t)(i / n), v, 0u);
}

// Each step: the static triangles' entries of the links this step has (the
// source's link direction set by init), merged into every fluid puller's
// entry like the triangles pass merges them.
__global__ __launch_bounds__(256) void ray_static_seed(const uint4* __restrict__ entries, uint32_t count,
                                                       const uint8_t* __restrict__ mask,
                                                       const uint32_t* __restrict__ dirs, float* __restrict__ links,
                                                       RayParams p) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint4 e = entries[i];
    const uint32_t src = e.x;
    const int q = (int)e.y;
    if (!((dirs[src] >> q) & 1u)) return;
    const int n[3] = {p.g.nx, p.g.ny, p.g.nz};
    const int sx = (int)(src % (uint32_t)n[0]), sy = (int)((src / (uint32_t)n[0]) % (uint32_t)n[1]),
              sz = (int)(src / ((uint32_t)n[0] * (uint32_t)n[1]));
    const int c[3] = {cx(q), cy(q), cz(q)};
    int fx[2], fy[2], fz[2];
    const int kx = pullers(p.bc[0], sx, c[0], n[0], fx);
    const int ky = pullers(p.bc[1], sy, c[1], n[1], fy);
    const int kz = pullers(p.bc[2], sz, c[2], n[2], fz);
    uint32_t* table = (uint32_t*)links;
    for (int iz = 0; iz < kz; ++iz)
        for (int iy = 0; iy < ky; ++iy)
            for (int ix = 0; ix < kx; ++ix) {
                const uint32_t gid =
                    (uint32_t)fx[ix] + (uint32_t)n[0] * ((uint32_t)fy[iy] + (uint32_t)n[1] * (uint32_t)fz[iz]);
                if (mask[gid] == 0) (void)atomicMin(&table[table_slot(p.g, fx[ix], fy[iy], fz[iz], q)], e.z);
            }
}

// 3. fractions: the entry when it lies in the link, else the signed-distance
// ratio; clamped to [1e-4, 1].  Clears the cell's link directions for the
// next step.
__global__ __launch_bounds__(256) void ray_finalize(const uint32_t* __restrict__ list,
                                                    const uint32_t* __restrict__ count,
                                                    const uint8_t* __restrict__ mask, const float4* __restrict__ wall,
                                                    float* __restrict__ links, uint32_t* __restrict__ dirs,
                                                    RayParams p) {
    const uint32_t n_tiles = *count;
    for (uint32_t i = blockIdx.x; i < n_tiles; i += gridDim.x) {
        const uint32_t tile = list[i];
        int x0, y0, z;
        tile_origin(p.g, tile, x0, y0, z);
        const int x = x0 + (int)(threadIdx.x & (TILE_X - 1));
        const int y = y0 + (int)(threadIdx.x / TILE_X);
        if (x >= p.g.nx || y >= p.g.ny) continue;
        const uint32_t gid = (uint32_t)x + (uint32_t)p.g.nx * ((uint32_t)y + (uint32_t)p.g.ny * (uint32_t)z);
        dirs[gid] = 0u;
        if (mask[gid] != 0) continue;
        float* out = links + (size_t)tile * 19 * TILE_CELLS + threadIdx.x;
        float best[19];
#pragma unroll
        for (int q = 1; q < 19; ++q) best[q] = out[(size_t)q * TILE_CELLS];
        const float wf = wall[gid].w;
#pragma unroll
        for (int q = 1; q < 19; ++q) {
            // init left 0 on every entry that is not a link
            if (best[q] == 0.0f) continue;
            float qf;
            if (best[q] <= 1.0f) {
                qf = best[q];
            } else {
                uint32_t src = gid;
                (void)ray_source_cell(p, x, y, z, q, src);
                const float ws = wall[src].w;
                const float df = wf > 0.0f ? wf : 0.0f;
                const float ds = ws < 0.0f ? ws : 0.0f;
                const float den = df - ds > 1.0e-6f ? df - ds : 1.0e-6f;
                qf = df / den;
            }
            qf = qf < 1.0e-4f ? 1.0e-4f : qf;
            qf = qf > 1.0f ? 1.0f : qf;
            out[(size_t)q * TILE_CELLS] = qf;
        }
    }
}

hipError_t launch_ray_links(const RayBuffers& b, const RayParams& p, hipStream_t s) {
    if (p.split < 1 || p.split > 64 || p.triangle_blocks < 1 || p.tile_pass_blocks < 1) return hipErrorInvalidValue;
    // queued pairs pack x | y << 16 and z << 5 | q
    if (p.g.nx > 65535 || p.g.ny > 65535 || p.g.nz >= (1 << 27)) return hipErrorInvalidValue;
    const unsigned tile_blocks = (p.g.tiles + 255) / 256;
    hipLaunchKernelGGL(ray_clear_count, dim3(1), dim3(64), 0, s, b.list_count);
    hipLaunchKernelGGL(ray_tile_list, dim3(tile_blocks), dim3(256), 0, s, (const uint8_t*)b.dyn, p.g.tiles, b.list,
                       b.list_count);
    hipLaunchKernelGGL(ray_init, dim3(p.tile_pass_blocks), dim3(256), 0, s, (const uint32_t*)b.list,
                       (const uint32_t*)b.list_count, b.mask, b.links, b.dirs, b.winners, p);
    // the static cache serves the plain pass; the adjoint's pass with
    // winners walks every triangle
    const bool cached = b.static_entries && !b.winners;
    if (cached && b.static_count > 0)
        hipLaunchKernelGGL(ray_static_seed, dim3((b.static_count + 255) / 256), dim3(256), 0, s, b.static_entries,
                           b.static_count, b.mask, (const uint32_t*)b.dirs, b.links, p);
    const uint32_t* walk = cached ? b.dynamic_list : nullptr;
    const uint32_t walk_count = cached ? b.dynamic_count : p.triangle_count;
    if (walk_count > 0) {
        hipLaunchKernelGGL(ray_triangles<RAY_MIN>, dim3(p.triangle_blocks, p.split), dim3(64), 0, s, b.pos, b.tri,
                           walk, walk_count, b.mask, (const uint32_t*)b.dirs, b.links, (uint32_t*)nullptr, p);
        if (b.winners)
            hipLaunchKernelGGL(ray_triangles<RAY_WINNERS>, dim3(p.triangle_blocks, p.split), dim3(64), 0, s, b.pos,
                               b.tri, walk, walk_count, b.mask, (const uint32_t*)b.dirs, b.links, b.winners, p);
    }
    hipLaunchKernelGGL(ray_finalize, dim3(p.tile_pass_blocks), dim3(256), 0, s, (const uint32_t*)b.list,
                       (const uint32_t*)b.list_count, b.mask, b.wall, b.links, b.dirs, p);
    return hipGetLastError();
}

}  // namespace bf

namespace bf {

hipError_t launch_ray_static(const float4* pos, const uint32_t* tri, const uint32_t* static_list,
                             uint32_t static_count, uint32_t* scratch_table, uint4* out, uint32_t capacity,
                             uint32_t* count, const RayParams& p, hipStream_t s) {
    if (p.split < 1 || p.split > 64 || p.triangle_blocks < 1) return hipErrorInvalidValue;
    if (p.g.nx > 65535 || p.g.ny > 65535 || p.g.nz >= (1 << 27)) return hipErrorInvalidValue;
    const uint64_t entries = 19ull * p.g.n;
    hipLaunchKernelGGL(ray_clear_count, dim3(1), dim3(64), 0, s, count);
    if (static_count > 0)
        hipLaunchKernelGGL(ray_triangles<RAY_STATIC>, dim3(p.triangle_blocks, p.split), dim3(64), 0, s, pos, tri,
                           static_list, static_count, (const uint8_t*)nullptr, (const uint32_t*)nullptr,
                           (float*)scratch_table, (uint32_t*)nullptr, p);
    hipLaunchKernelGGL(ray_static_compact, dim3((unsigned)((entries + 255) / 256)), dim3(256), 0, s,
                       (const uint32_t*)scratch_table, p.g.n, out, capacity, count);
    return hipGetLastError();
}

}  // namespace bf
