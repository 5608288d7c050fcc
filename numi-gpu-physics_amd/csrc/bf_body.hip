// Six-degree-of-freedom rigid body on the device, FP64, one thread
// (integrateBirdBody, BirdFlow.metal:6574-6627; RigidBody.swift:85-126), and
// the runtime safety monitor (monitorBirdRuntimeSafety,
// BirdFlow.metal:6629-6707).  Runs on the load-reduction stream right after
// a step's loads are final, so the state never leaves the device between
// steps; the fluid kernels read the torque reference point from it and the
// mesh placement reads the pose from it.
#include "bf_launch.h"

namespace bf {

__device__ void qnormalize_d(double q[4]) {
    const double n2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    if (n2 > 1.0e-20) {
        const double s = 1.0 / sqrt(n2);
        for (int i = 0; i < 4; ++i) q[i] *= s;
    } else {
        q[0] = q[1] = q[2] = 0.0;
        q[3] = 1.0;
    }
}

__device__ void cross_d(const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// v' = v + w t + q x t with t = 2 q x v (q = x, y, z, w)
__device__ void qrotate_d(const double q[4], const double v[3], double o[3]) {
    double t[3], c[3];
    cross_d(q, v, t);
    for (int i = 0; i < 3; ++i) t[i] *= 2.0;
    cross_d(q, t, c);
    for (int i = 0; i < 3; ++i) o[i] = v[i] + q[3] * t[i] + c[i];
}

// state: position[3], orientation[4], linear_velocity[3], angular_velocity_body[3]
__global__ void body_step(double* __restrict__ state, const double* __restrict__ load, double* __restrict__ history,
                          uint32_t history_slot, double* __restrict__ metrics, uint32_t* __restrict__ event,
                          uint64_t step, BodyParams bp, const float* __restrict__ wing_reaction) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double x[3] = {state[0], state[1], state[2]};
    double q[4] = {state[3], state[4], state[5], state[6]};
    double v[3] = {state[7], state[8], state[9]};
    double w[3] = {state[10], state[11], state[12]};
    double force[3], torque[3];
    for (int i = 0; i < 3; ++i) {
        force[i] = load[i] + bp.external_force[i];
        torque[i] = load[3 + i] + bp.external_torque[i];
        if (wing_reaction) {
            // the prescribed wings' hinge reactions (BirdFlow.metal:6586-6591)
            force[i] = force[i] + (double)wing_reaction[i] + (double)wing_reaction[8 + i];
            torque[i] = torque[i] + (double)wing_reaction[4 + i] + (double)wing_reaction[12 + i];
        }
    }
    const int substeps = bp.substeps < 1 ? 1 : bp.substeps;
    const double dt = bp.dt / (double)substeps;
    qnormalize_d(q);
    for (int s = 0; s < substeps; ++s) {
        for (int i = 0; i < 3; ++i) {
            v[i] += (force[i] / bp.mass + bp.gravity[i]) * dt;
            x[i] += v[i] * dt;
        }
        const double qc[4] = {-q[0], -q[1], -q[2], q[3]};
        double tb[3];
        qrotate_d(qc, torque, tb);
        const double L[3] = {bp.inertia[0] * w[0], bp.inertia[1] * w[1], bp.inertia[2] * w[2]};
        double g[3];
        cross_d(w, L, g);
        for (int i = 0; i < 3; ++i) w[i] += (tb[i] - g[i]) / bp.inertia[i] * dt;
        // q <- normalize(q + dt/2 q (x) (w, 0))
        double cq[3], d[4];
        cross_d(q, w, cq);
        for (int i = 0; i < 3; ++i) d[i] = q[3] * w[i] + cq[i];
        d[3] = -(q[0] * w[0] + q[1] * w[1] + q[2] * w[2]);
        for (int i = 0; i < 4; ++i) q[i] = q[i] + 0.5 * dt * d[i];
        qnormalize_d(q);
    }
    const double out[13] = {x[0], x[1], x[2], q[0], q[1], q[2], q[3], v[0], v[1], v[2], w[0], w[1], w[2]};
    for (int i = 0; i < 13; ++i) state[i] = out[i];
    if (history)
        for (int i = 0; i < 13; ++i) history[(size_t)history_slot * 16 + i] = out[i];

    if (!bp.monitor) return;
    // safety: clearance of the rotated extent box to the sponge-padded
    // domain, and a Mach estimate of the fastest surface point
    double ext[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < 3; ++a) {
        double e[3] = {0.0, 0.0, 0.0}, r[3];
        e[a] = bp.extent[a];
        qrotate_d(q, e, r);
        for (int i = 0; i < 3; ++i) ext[i] += fabs(r[i]);
    }
    double clearance = INFINITY;
    for (int i = 0; i < 3; ++i) {
        const double lo = x[i] - ext[i] - bp.domain_min[i];
        const double hi = bp.domain_max[i] - x[i] - ext[i];
        clearance = fmin(clearance, fmin(lo, hi));
    }
    double rel[3];
    for (int i = 0; i < 3; ++i) rel[i] = v[i] - bp.far_velocity_world[i];
    const double translation = sqrt(rel[0] * rel[0] + rel[1] * rel[1] + rel[2] * rel[2]);
    const double rotation = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]) * bp.rotation_radius;
    const double mach = (translation + rotation + bp.wing_speed) * bp.velocity_to_lattice / sqrt(1.0 / 3.0);
    bool finite = isfinite(mach) && isfinite(clearance);
    for (int i = 0; i < 13; ++i) finite = finite && isfinite(out[i]);
    if (isfinite(mach)) metrics[0] = fmax(metrics[0], mach);
    if (isfinite(clearance)) metrics[1] = fmin(metrics[1], clearance);
    uint32_t flags = 0;
    if (!finite) flags |= 4u;
    if (isfinite(mach) && mach > bp.mach_limit) flags |= 1u;
    if (isfinite(clearance) && clearance < 0.0) flags |= 2u;
    if (flags != 0u && event[2] == 0u) {
        event[0] = (uint32_t)step;
        event[1] = (uint32_t)(step >> 32);
        event[2] = flags;
        metrics[2] = mach;
        metrics[3] = clearance;
    }
}

hipError_t launch_body_step(double* state, const double* load, double* history, uint32_t history_slot,
                            double* metrics, uint32_t* event, uint64_t step, const BodyParams& bp, hipStream_t s,
                            const float* wing_reaction) {
    hipLaunchKernelGGL(body_step, dim3(1), dim3(64), 0, s, state, load, history, history_slot, metrics, event, step,
                       bp, wing_reaction);
    return hipGetLastError();
}

}  // namespace bf
