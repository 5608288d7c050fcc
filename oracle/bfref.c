/*
 * bfref — CPU oracle for the D3Q19 moving-boundary step.
 * TEST INFRASTRUCTURE ONLY (see bfref.h for the scope and the reference
 * files each function restates).  Build without floating-point contraction:
 * the FP32 functions are operation-for-operation restatements.
 */
#include "bfref.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* D3Q19 ordering: 0; +-x; +-y; +-z; (+-1,+-1,0); (+-1,0,+-1); (0,+-1,+-1).
 * BirdFlow.metal:7-44; D3Q19.swift:8-45. */
const int bfo_c[19][3] = {
    {0, 0, 0},
    {1, 0, 0}, {-1, 0, 0},
    {0, 1, 0}, {0, -1, 0},
    {0, 0, 1}, {0, 0, -1},
    {1, 1, 0}, {-1, -1, 0},
    {1, -1, 0}, {-1, 1, 0},
    {1, 0, 1}, {-1, 0, -1},
    {1, 0, -1}, {-1, 0, 1},
    {0, 1, 1}, {0, -1, -1},
    {0, 1, -1}, {0, -1, 1},
};
const int bfo_opp[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
const double bfo_w[19] = {
    1.0 / 3.0,
    1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
};

/* link fractions of BFO_WALL_INTERPOLATED_LINKS steps ([19][N], fluid side),
 * FP32 or (finite differences of FP64 steps) FP64 */
static const float* bfo_link_table = NULL;
static const double* bfo_link_table64 = NULL;
void bfo_set_link_table(const float* links) { bfo_link_table = links; }
void bfo_set_link_table_f64(const double* links) { bfo_link_table64 = links; }

#define R float
#define WT float
#define FN(name) name##_f32
#include "bfref_fluid.inc"
#undef R
#undef WT
#undef FN

#define R double
#define WT float
#define FN(name) name##_f64
#include "bfref_fluid.inc"
#undef R
#undef WT
#undef FN

/* FP64 twin with an FP64 wall field: the reference gradient of the adjoint
 * tests differentiates through wall values, which FP32 would round. */
#define R double
#define WT double
#define FN(name) name##_f64w
#include "bfref_fluid.inc"
#undef R
#undef WT
#undef FN

/* InterpolatedBounceBackReference.swift:28-42 (FP64, for known-answer tests). */
double bfo_bouzidi_population_f64(double q, double reflected, double farther,
                                  double previous, double correction)
{
    if (q <= 0.5) return 2 * q * reflected + (1 - 2 * q) * farther + correction;
    return (reflected + correction) / (2 * q) + (2 * q - 1) * previous / (2 * q);
}

/* ------------------------------------------------------------------------
 * Geometry (FP32).  Operation order follows the CPU twin in
 * MetalIndexedBirdSurfaceValidation.swift:18005-18110 and 20302-20359,
 * whose output the reference reports bit-identical to its GPU kernels.
 * ---------------------------------------------------------------------- */

static float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static float len3(const float a[3]) { return sqrtf(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

/* BirdFlow.metal:1529-1583 (Voronoi-region closest point and barycentrics). */
static void closest_point(const float p[3], const float a[3], const float b[3], const float c[3],
                          float out[3], float bary[3])
{
    float ab[3], ac[3], ap[3];
    for (int i = 0; i < 3; ++i) {
        ab[i] = b[i] - a[i];
        ac[i] = c[i] - a[i];
        ap[i] = p[i] - a[i];
    }
    float d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) {
        memcpy(out, a, sizeof(float) * 3);
        bary[0] = 1; bary[1] = 0; bary[2] = 0;
        return;
    }
    float bp[3];
    for (int i = 0; i < 3; ++i) bp[i] = p[i] - b[i];
    float d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) {
        memcpy(out, b, sizeof(float) * 3);
        bary[0] = 0; bary[1] = 1; bary[2] = 0;
        return;
    }
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        float v = d1 / (d1 - d3);
        for (int i = 0; i < 3; ++i) out[i] = a[i] + v * ab[i];
        bary[0] = 1.0f - v; bary[1] = v; bary[2] = 0;
        return;
    }
    float cp[3];
    for (int i = 0; i < 3; ++i) cp[i] = p[i] - c[i];
    float d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) {
        memcpy(out, c, sizeof(float) * 3);
        bary[0] = 0; bary[1] = 0; bary[2] = 1;
        return;
    }
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        float w = d2 / (d2 - d6);
        for (int i = 0; i < 3; ++i) out[i] = a[i] + w * ac[i];
        bary[0] = 1.0f - w; bary[1] = 0; bary[2] = w;
        return;
    }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int i = 0; i < 3; ++i) out[i] = b[i] + w * (c[i] - b[i]);
        bary[0] = 0; bary[1] = 1.0f - w; bary[2] = w;
        return;
    }
    float den = va + vb + vc;
    if (den < 1.0e-20f) den = 1.0e-20f;
    float inv = 1.0f / den;
    float v = vb * inv, w = vc * inv;
    for (int i = 0; i < 3; ++i) out[i] = a[i] + ab[i] * v + ac[i] * w;
    bary[0] = 1.0f - v - w; bary[1] = v; bary[2] = w;
}

/* MeasuredBirdSurfaceSequence.swift:174-204 */
int bfo_frame_interval(const float* times, int frame_count, float t, float* blend, float* duration)
{
    if (t <= times[0]) {
        if (blend) *blend = 0.0f;
        if (duration) *duration = times[1] - times[0];
        return 0;
    }
    int last = frame_count - 1;
    if (t >= times[last]) {
        if (blend) *blend = 1.0f;
        if (duration) *duration = times[last] - times[last - 1];
        return last - 1;
    }
    int lower = 0, upper = last;
    while (lower + 1 < upper) {
        int middle = lower + (upper - lower) / 2;
        if (t < times[middle]) upper = middle;
        else lower = middle;
    }
    float d = times[upper] - times[lower];
    if (blend) *blend = (t - times[lower]) / d;
    if (duration) *duration = d;
    return lower;
}

/* BirdFlow.metal:1589-1622: position = translation + mix(P_f, P_f+1, blend),
 * velocity = (P_f+1 - P_f) / duration * velocity_scale. */
void bfo_prepare_vertices(const float* points, const float* times, int frame_count,
                          int vertex_count, int first, float query_time,
                          const float translation[3], float velocity_scale,
                          float* position3, float* velocity3)
{
    if (first > frame_count - 2) first = frame_count - 2;
    int second = first + 1;
    float duration = times[second] - times[first];
    float blend = (query_time - times[first]) / duration;
    if (blend < 0.0f) blend = 0.0f;
    if (blend > 1.0f) blend = 1.0f;
    for (int v = 0; v < vertex_count; ++v) {
        const float* a = points + ((size_t)first * vertex_count + v) * 3;
        const float* b = points + ((size_t)second * vertex_count + v) * 3;
        for (int i = 0; i < 3; ++i) {
            float d = b[i] - a[i];
            position3[3 * v + i] = translation[i] + (a[i] + blend * d);
            velocity3[3 * v + i] = d / duration * velocity_scale;
        }
    }
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* BirdFlow.metal:1638-1686; CPU twin MetalIndexedBirdSurfaceValidation.swift:18005-18064. */
void bfo_raster(const bfo_params* p, const float* position3, const uint32_t* tri,
                int triangle_count, float half_thickness, int key_bits, void* keys)
{
    const int n[3] = {p->nx, p->ny, p->nz};
    const float dx = p->dx;
    for (int t = 0; t < triangle_count; ++t) {
        if (key_bits == 32 && t >= 4096) break;
        const float* a = position3 + 3 * (size_t)tri[3 * t + 0];
        const float* b = position3 + 3 * (size_t)tri[3 * t + 1];
        const float* c = position3 + 3 * (size_t)tri[3 * t + 2];
        float guard = half_thickness + 2.0f * dx;
        int lo[3], hi[3];
        for (int i = 0; i < 3; ++i) {
            float mn = fminf(a[i], fminf(b[i], c[i]));
            float mx = fmaxf(a[i], fmaxf(b[i], c[i]));
            float lw = mn - guard, uw = mx + guard;
            float ls = (lw - p->origin[i]) / dx - 0.5f;
            float us = (uw - p->origin[i]) / dx - 0.5f;
            lo[i] = clampi((int)floorf(ls), 0, n[i] - 1);
            hi[i] = clampi((int)ceilf(us), 0, n[i] - 1);
        }
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x) {
                    float w[3] = {
                        p->origin[0] + ((float)x + 0.5f) * dx,
                        p->origin[1] + ((float)y + 0.5f) * dx,
                        p->origin[2] + ((float)z + 0.5f) * dx,
                    };
                    float cp[3], bary[3];
                    closest_point(w, a, b, c, cp, bary);
                    float d[3] = {w[0] - cp[0], w[1] - cp[1], w[2] - cp[2]};
                    float dist_cells = len3(d) / dx;
                    float scaled = roundf(dist_cells * 65536.0f);
                    uint32_t bin = scaled >= (float)0xFFFFF ? 0xFFFFFu : (uint32_t)scaled;
                    size_t gid = (size_t)x + (size_t)p->nx * ((size_t)y + (size_t)p->ny * (size_t)z);
                    if (key_bits == 32) {
                        uint32_t key = (bin << 12) | (uint32_t)t;
                        uint32_t* k = (uint32_t*)keys;
                        if (key < k[gid]) k[gid] = key;
                    } else {
                        uint64_t key = ((uint64_t)bin << 32) | (uint64_t)(uint32_t)t;
                        uint64_t* k = (uint64_t*)keys;
                        if (key < k[gid]) k[gid] = key;
                    }
                }
    }
}

void bfo_clear_geometry(const bfo_params* p, uint8_t* mask, float* wall4)
{
    const size_t N = (size_t)p->nx * p->ny * p->nz;
    memset(mask, 0, N);
    for (size_t i = 0; i < N; ++i) {
        wall4[4 * i + 0] = 0.0f;
        wall4[4 * i + 1] = 0.0f;
        wall4[4 * i + 2] = 0.0f;
        wall4[4 * i + 3] = 16.0f;
    }
}

/* BirdFlow.metal:1697-1738; CPU twin MetalIndexedBirdSurfaceValidation.swift:18066-18106. */
void bfo_resolve(const bfo_params* p, const float* position3, const float* velocity3,
                 const uint32_t* tri, const uint8_t* tri_part, int triangle_count,
                 float half_thickness, int key_bits, const void* keys,
                 uint8_t* mask, float* wall4)
{
    const size_t N = (size_t)p->nx * p->ny * p->nz;
    const float dx = p->dx;
    for (size_t gid = 0; gid < N; ++gid) {
        uint32_t t;
        if (key_bits == 32) {
            uint32_t key = ((const uint32_t*)keys)[gid];
            if (key == UINT32_MAX) continue;
            t = key & 0xFFFu;
        } else {
            uint64_t key = ((const uint64_t*)keys)[gid];
            if (key == UINT64_MAX) continue;
            t = (uint32_t)(key & 0xFFFFFFFFu);
        }
        if ((int)t >= triangle_count) continue;
        int x = (int)(gid % (size_t)p->nx);
        size_t yz = gid / (size_t)p->nx;
        int y = (int)(yz % (size_t)p->ny);
        int z = (int)(yz / (size_t)p->ny);
        float w[3] = {
            p->origin[0] + ((float)x + 0.5f) * dx,
            p->origin[1] + ((float)y + 0.5f) * dx,
            p->origin[2] + ((float)z + 0.5f) * dx,
        };
        uint32_t ia = tri[3 * t + 0], ib = tri[3 * t + 1], ic = tri[3 * t + 2];
        float cp[3], bary[3];
        closest_point(w, position3 + 3 * ia, position3 + 3 * ib, position3 + 3 * ic, cp, bary);
        float vel[3];
        for (int i = 0; i < 3; ++i)
            vel[i] = bary[0] * velocity3[3 * ia + i] + bary[1] * velocity3[3 * ib + i]
                + bary[2] * velocity3[3 * ic + i];
        float d[3] = {w[0] - cp[0], w[1] - cp[1], w[2] - cp[2]};
        float sd = len3(d) - half_thickness;
        mask[gid] = sd <= 0.0f ? tri_part[t] : 0;
        wall4[4 * gid + 0] = vel[0];
        wall4[4 * gid + 1] = vel[1];
        wall4[4 * gid + 2] = vel[2];
        wall4[4 * gid + 3] = sd / dx;
    }
}

/* BirdFlow.metal:183-195 and 2307-2350 (occupancy and wall velocity only). */
void bfo_translating_sphere(const bfo_params* p, const float center_radius[4],
                            const float geometry_velocity[3], const float wall_velocity[4],
                            float t, uint8_t* mask, float* wall4)
{
    float center[3];
    for (int i = 0; i < 3; ++i) center[i] = center_radius[i] + geometry_velocity[i] * t;
    float r = center_radius[3];
    int mode = (int)((wall_velocity[3] > 0.0f ? wall_velocity[3] : 0.0f) + 0.5f);
    for (int z = 0; z < p->nz; ++z)
        for (int y = 0; y < p->ny; ++y)
            for (int x = 0; x < p->nx; ++x) {
                size_t gid = (size_t)x + (size_t)p->nx * ((size_t)y + (size_t)p->ny * (size_t)z);
                float rel[3] = {(float)x + 0.5f - center[0], (float)y + 0.5f - center[1],
                                (float)z + 0.5f - center[2]};
                int inside = dot3(rel, rel) <= r * r;
                float wv[3] = {wall_velocity[0], wall_velocity[1], wall_velocity[2]};
                if (mode != 0) {
                    float l = len3(rel);
                    if (l < 1.0e-6f) l = 1.0e-6f;
                    float nrm[3] = {rel[0] / l, rel[1] / l, rel[2] / l};
                    float dn = dot3(wv, nrm);
                    float nw[3] = {dn * nrm[0], dn * nrm[1], dn * nrm[2]};
                    for (int i = 0; i < 3; ++i) wv[i] = mode == 1 ? wv[i] - nw[i] : nw[i];
                }
                mask[gid] = inside ? 1 : 0;
                wall4[4 * gid + 0] = wv[0];
                wall4[4 * gid + 1] = wv[1];
                wall4[4 * gid + 2] = wv[2];
                wall4[4 * gid + 3] = 0.0f;
            }
}

void bfo_link_census(const bfo_params* p, const uint8_t* solid_prev, const uint8_t* solid_cur,
                     uint64_t link_counts[19], uint64_t* covered, uint64_t* uncovered)
{
    for (int q = 0; q < 19; ++q) link_counts[q] = 0;
    uint64_t cov = 0, unc = 0;
    for (int z = 0; z < p->nz; ++z)
        for (int y = 0; y < p->ny; ++y)
            for (int x = 0; x < p->nx; ++x) {
                size_t gid = (size_t)x + (size_t)p->nx * ((size_t)y + (size_t)p->ny * (size_t)z);
                int s = solid_cur[gid] != 0, w = solid_prev[gid] != 0;
                if (s && !w) ++cov;
                if (!s && w) ++unc;
                if (s || w) continue;
                for (int q = 1; q < 19; ++q) {
                    size_t src;
                    if (pull_source_f32(p, x, y, z, q, &src)) continue;
                    if (solid_cur[src] != 0) ++link_counts[q];
                }
            }
    *covered = cov;
    *uncovered = unc;
}

/* ------------------------------------------------------------------------
 * Rigid body (FP64), BirdFlow.metal:6574-6627 / RigidBody.swift:85-127.
 * ---------------------------------------------------------------------- */
static void qnormalize(double q[4])
{
    double n2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    if (n2 > 1.0e-20) {
        double s = 1.0 / sqrt(n2);
        for (int i = 0; i < 4; ++i) q[i] *= s;
    } else {
        q[0] = q[1] = q[2] = 0.0;
        q[3] = 1.0;
    }
}

static void cross3d(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

static void qrotate(const double q[4], const double v[3], double o[3])
{
    double t[3], c[3];
    cross3d(q, v, t);
    for (int i = 0; i < 3; ++i) t[i] *= 2.0;
    cross3d(q, t, c);
    for (int i = 0; i < 3; ++i) o[i] = v[i] + q[3] * t[i] + c[i];
}

void bfo_body_integrate(bfo_body* b, const double force[3], const double torque_world[3],
                        double mass, const double inertia[3], const double gravity[3],
                        double dt_total, int substeps)
{
    if (substeps < 1) substeps = 1;
    double dt = dt_total / (double)substeps;
    double q[4];
    memcpy(q, b->orientation, sizeof q);
    qnormalize(q);
    for (int s = 0; s < substeps; ++s) {
        for (int i = 0; i < 3; ++i) {
            b->linear_velocity[i] += (force[i] / mass + gravity[i]) * dt;
            b->position[i] += b->linear_velocity[i] * dt;
        }
        double qc[4] = {-q[0], -q[1], -q[2], q[3]};
        double tb[3];
        qrotate(qc, torque_world, tb);
        double* w = b->angular_velocity_body;
        double L[3] = {inertia[0] * w[0], inertia[1] * w[1], inertia[2] * w[2]};
        double g[3];
        cross3d(w, L, g);
        for (int i = 0; i < 3; ++i) w[i] += (tb[i] - g[i]) / inertia[i] * dt;
        /* derivative = q (x) (omega, 0) */
        double d[4];
        double cq[3];
        cross3d(q, w, cq);
        for (int i = 0; i < 3; ++i) d[i] = q[3] * w[i] + cq[i];
        d[3] = -(q[0] * w[0] + q[1] * w[1] + q[2] * w[2]);
        for (int i = 0; i < 4; ++i) q[i] = q[i] + 0.5 * dt * d[i];
        qnormalize(q);
    }
    memcpy(b->orientation, q, sizeof q);
}

/* BirdFlow.metal:6629-6707, evaluated in FP64. */
uint32_t bfo_body_safety(const bfo_body* b, const double extent[3], double rotation_radius,
                         double wing_speed, double mach_limit, const double domain_min[3],
                         const double domain_max[3], const double far_velocity_world[3],
                         double velocity_to_lattice, uint64_t step, double metrics[4], uint32_t event[3])
{
    double q[4];
    memcpy(q, b->orientation, sizeof q);
    qnormalize(q);
    double world_extent[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < 3; ++a) {
        double e[3] = {0.0, 0.0, 0.0}, r[3];
        e[a] = extent[a];
        qrotate(q, e, r);
        for (int i = 0; i < 3; ++i) world_extent[i] += fabs(r[i]);
    }
    double clearance = INFINITY;
    for (int i = 0; i < 3; ++i) {
        double lo = b->position[i] - world_extent[i] - domain_min[i];
        double hi = domain_max[i] - b->position[i] - world_extent[i];
        clearance = fmin(clearance, fmin(lo, hi));
    }
    double rel[3];
    for (int i = 0; i < 3; ++i) rel[i] = b->linear_velocity[i] - far_velocity_world[i];
    const double* w = b->angular_velocity_body;
    double translation = sqrt(rel[0] * rel[0] + rel[1] * rel[1] + rel[2] * rel[2]);
    double rotation = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]) * rotation_radius;
    double mach = (translation + rotation + wing_speed) * velocity_to_lattice / sqrt(1.0 / 3.0);
    int finite = isfinite(mach) && isfinite(clearance);
    for (int i = 0; i < 3; ++i)
        finite = finite && isfinite(b->position[i]) && isfinite(b->linear_velocity[i]) && isfinite(w[i]);
    for (int i = 0; i < 4; ++i) finite = finite && isfinite(b->orientation[i]);
    if (isfinite(mach)) metrics[0] = fmax(metrics[0], mach);
    if (isfinite(clearance)) metrics[1] = fmin(metrics[1], clearance);
    uint32_t flags = 0;
    if (!finite) flags |= 4u;
    if (isfinite(mach) && mach > mach_limit) flags |= 1u;
    if (isfinite(clearance) && clearance < 0.0) flags |= 2u;
    if (flags != 0u && event[2] == 0u) {
        event[0] = (uint32_t)step;
        event[1] = (uint32_t)(step >> 32);
        event[2] = flags;
        metrics[2] = mach;
        metrics[3] = clearance;
    }
    return flags;
}

static void qrotatef(const float q[4], const float v[3], float o[3])
{
    float t[3] = {2.0f * (q[1] * v[2] - q[2] * v[1]), 2.0f * (q[2] * v[0] - q[0] * v[2]),
                  2.0f * (q[0] * v[1] - q[1] * v[0])};
    float c[3] = {q[1] * t[2] - q[2] * t[1], q[2] * t[0] - q[0] * t[2], q[0] * t[1] - q[1] * t[0]};
    for (int i = 0; i < 3; ++i) o[i] = v[i] + q[3] * t[i] + c[i];
}

void bfo_place_vertices(const bfo_body* b, int vertex_count, float velocity_scale,
                        const float* local_position3, const float* local_velocity3,
                        float* position3, float* velocity3)
{
    double qd[4];
    memcpy(qd, b->orientation, sizeof qd);
    qnormalize(qd);
    const float q[4] = {(float)qd[0], (float)qd[1], (float)qd[2], (float)qd[3]};
    const float x[3] = {(float)b->position[0], (float)b->position[1], (float)b->position[2]};
    const float vb[3] = {(float)b->linear_velocity[0], (float)b->linear_velocity[1], (float)b->linear_velocity[2]};
    const float w[3] = {(float)b->angular_velocity_body[0], (float)b->angular_velocity_body[1],
                        (float)b->angular_velocity_body[2]};
    for (int v = 0; v < vertex_count; ++v) {
        const float* p = local_position3 + 3 * v;
        const float* u = local_velocity3 + 3 * v;
        float wxp[3] = {w[1] * p[2] - w[2] * p[1], w[2] * p[0] - w[0] * p[2], w[0] * p[1] - w[1] * p[0]};
        float r[3], rw[3], ru[3];
        qrotatef(q, p, r);
        qrotatef(q, wxp, rw);
        qrotatef(q, u, ru);
        for (int i = 0; i < 3; ++i) {
            position3[3 * v + i] = x[i] + r[i];
            velocity3[3 * v + i] = (vb[i] + (rw[i] + ru[i])) * velocity_scale;
        }
    }
}

#include "bfref_bird.inc"
#include "bfref_wing.inc"
#include "bfref_kin.inc"

/* ---- exact link fractions against a meshed shell ---- */
static float rdot(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void rcross(const float a[3], const float b[3], float o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static void rsub(const float a[3], const float b[3], float o[3])
{
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}
/* point P on the plane side of edge (U, V) of a triangle with normal n */
static int redge_in(const float U[3], const float V[3], const float P[3], const float n[3])
{
    float e[3], m[3], c[3];
    rsub(V, U, e);
    rsub(P, U, m);
    rcross(e, m, c);
    return rdot(c, n) >= 0.0f;
}

float bfo_ray_shell(const float F[3], const float D[3], const float A[3], const float B[3], const float C[3],
                    float h)
{
    /* F within the shell (inside a vertex sphere, an edge cylinder over the
     * edge, or the slab over the face) enters at 0 */
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
        if (fn <= h && fn >= -h && redge_in(A, B, F, n) && redge_in(B, C, F, n) && redge_in(C, A, F, n))
            return 0.0f;
        const float dn = rdot(D, n);
        if (dn != 0.0f) {
            for (int side = 0; side < 2; ++side) {
                const float off = side == 0 ? h : -h;
                const float t = (off - fn) / dn;
                if (t >= 0.0f && t < best) {
                    const float P[3] = {F[0] + D[0] * t, F[1] + D[1] * t, F[2] + D[2] * t};
                    if (redge_in(A, B, P, n) && redge_in(B, C, P, n) && redge_in(C, A, P, n)) best = t;
                }
            }
        }
    }
    const float* ends[3][2] = {{A, B}, {B, C}, {C, A}};
    const float dd = rdot(D, D);
    for (int k = 0; k < 3; ++k) {
        float e[3], m[3];
        rsub(ends[k][1], ends[k][0], e);
        rsub(F, ends[k][0], m);
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
                    const float u = (me + t * de) / ee;
                    if (u >= 0.0f && u <= 1.0f) best = t;
                }
            }
        }
    }
    const float* verts[3] = {A, B, C};
    for (int k = 0; k < 3; ++k) {
        float m[3];
        rsub(F, verts[k], m);
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

#include "bfref_ray64.inc"

/* the source coordinate of a link along one axis (the step's face
 * treatment); 0: beyond a far-field face */
static int ray_source(const int32_t bc[2], int s, int n, int* out)
{
    if (s < 0) {
        if (bc[0] == BFO_BC_FARFIELD) return 0;
        s = bc[0] == BFO_BC_PERIODIC ? n - 1 : 0;
    } else if (s >= n) {
        if (bc[1] == BFO_BC_FARFIELD) return 0;
        s = bc[1] == BFO_BC_PERIODIC ? 0 : n - 1;
    }
    *out = s;
    return 1;
}

void bfo_ray_links_winners(const bfo_params* p, const float* position3, const uint32_t* tri, int triangle_count,
                           float half_thickness, const uint8_t* mask, const float* wall4, float* links,
                           uint32_t* winners);

void bfo_ray_links(const bfo_params* p, const float* position3, const uint32_t* tri, int triangle_count,
                   float half_thickness, const uint8_t* mask, const float* wall4, float* links)
{
    bfo_ray_links_winners(p, position3, tri, triangle_count, half_thickness, mask, wall4, links, NULL);
}

void bfo_ray_links_winners(const bfo_params* p, const float* position3, const uint32_t* tri, int triangle_count,
                           float half_thickness, const uint8_t* mask, const float* wall4, float* links,
                           uint32_t* winners)
{
    const size_t n = (size_t)p->nx * p->ny * p->nz;
    const float dx = p->dx;
    /* per-triangle boxes of the shell, for a quick rejection */
    float* box = (float*)malloc(sizeof(float) * 6 * (size_t)(triangle_count > 0 ? triangle_count : 1));
    for (int t = 0; t < triangle_count; ++t)
        for (int a = 0; a < 3; ++a) {
            float lo = position3[3 * tri[3 * t] + a], hi = lo;
            for (int k = 1; k < 3; ++k) {
                const float v = position3[3 * tri[3 * t + k] + a];
                lo = v < lo ? v : lo;
                hi = v > hi ? v : hi;
            }
            box[6 * t + a] = lo - half_thickness - 0.01f * dx;
            box[6 * t + 3 + a] = hi + half_thickness + 0.01f * dx;
        }
    for (size_t i = 0; i < 19 * n; ++i) links[i] = 0.0f;
    if (winners)
        for (size_t i = 0; i < 19 * n; ++i) winners[i] = 0xFFFFFFFFu;
#pragma omp parallel for schedule(dynamic, 1)
    for (int z = 0; z < p->nz; ++z)
        for (int y = 0; y < p->ny; ++y)
            for (int x = 0; x < p->nx; ++x) {
                const size_t gid = (size_t)x + (size_t)p->nx * ((size_t)y + (size_t)p->ny * (size_t)z);
                if (mask[gid] != 0) continue;
                for (int q = 1; q < 19; ++q) {
                    int sx, sy, sz;
                    if (!ray_source(p->bc[0], x - bfo_c[q][0], p->nx, &sx) ||
                        !ray_source(p->bc[1], y - bfo_c[q][1], p->ny, &sy) ||
                        !ray_source(p->bc[2], z - bfo_c[q][2], p->nz, &sz))
                        continue;
                    const size_t src = (size_t)sx + (size_t)p->nx * ((size_t)sy + (size_t)p->ny * (size_t)sz);
                    if (mask[src] == 0) continue;
                    const float S[3] = {p->origin[0] + ((float)sx + 0.5f) * dx, p->origin[1] + ((float)sy + 0.5f) * dx,
                                        p->origin[2] + ((float)sz + 0.5f) * dx};
                    const float F[3] = {S[0] + (float)bfo_c[q][0] * dx, S[1] + (float)bfo_c[q][1] * dx,
                                        S[2] + (float)bfo_c[q][2] * dx};
                    const float D[3] = {(float)(-bfo_c[q][0]) * dx, (float)(-bfo_c[q][1]) * dx,
                                        (float)(-bfo_c[q][2]) * dx};
                    float lo[3], hi[3];
                    for (int a = 0; a < 3; ++a) {
                        lo[a] = F[a] < S[a] ? F[a] : S[a];
                        hi[a] = F[a] < S[a] ? S[a] : F[a];
                    }
                    float best = 2.0f;
                    uint32_t win = 0xFFFFFFFFu;   /* the first triangle with the least entry */
                    for (int t = 0; t < triangle_count; ++t) {
                        const float* bx = box + 6 * t;
                        if (hi[0] < bx[0] || lo[0] > bx[3] || hi[1] < bx[1] || lo[1] > bx[4] || hi[2] < bx[2] ||
                            lo[2] > bx[5])
                            continue;
                        const float r = bfo_ray_shell(F, D, position3 + 3 * tri[3 * t], position3 + 3 * tri[3 * t + 1],
                                                      position3 + 3 * tri[3 * t + 2], half_thickness);
                        if (r < best) win = (uint32_t)t;
                        best = r < best ? r : best;
                    }
                    if (winners && best <= 1.0f) winners[(size_t)q * n + gid] = win;
                    float qf;
                    if (best <= 1.0f) {
                        qf = best;
                    } else {
                        const float df = wall4[4 * gid + 3] > 0.0f ? wall4[4 * gid + 3] : 0.0f;
                        const float ds = wall4[4 * src + 3] < 0.0f ? wall4[4 * src + 3] : 0.0f;
                        const float den = df - ds > 1.0e-6f ? df - ds : 1.0e-6f;
                        qf = df / den;
                    }
                    qf = qf < 1.0e-4f ? 1.0e-4f : qf;
                    qf = qf > 1.0f ? 1.0f : qf;
                    links[(size_t)q * n + gid] = qf;
                }
            }
    free(box);
}
