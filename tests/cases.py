"""Canonical cases shared by the oracle tests and the GPU parity tests."""
from __future__ import annotations

import math

import numpy as np

import bfref

# Shear wave (MetalShearWaveValidation.swift:666-700; Reference/shear_wave_reference.py:144-178)
SHEAR_AMPLITUDE = 0.01


def shear_wave_omegas(viscosity=0.1, magic=0.25):
    tau_p = 0.5 + 3.0 * viscosity
    tau_m = 0.5 + magic / (tau_p - 0.5)
    return 1.0 / tau_p, 1.0 / tau_m


def shear_wave_params(n, omega_plus=None, omega_minus=None, **kw):
    if omega_plus is None:
        omega_plus, omega_minus = shear_wave_omegas()
    return bfref.make_params((n, n, n), omega_plus=omega_plus, omega_minus=omega_minus, **kw)


def mode_amplitude(p, f):
    """Sine projection of u_x over y (MetalShearWaveValidation.swift:357-369)."""
    _, u = bfref.moments(p, f)
    ux = u[:, 0].astype(np.float64).reshape(p.nz, p.ny, p.nx)
    y = np.arange(p.ny)
    s = np.sin(2 * np.pi * y / p.ny)
    return float(2.0 * (ux * s[None, :, None]).sum() / ux.size)


# Translating sphere (MetalTranslatingBodyTopologyValidation.swift:894-902, 4829-4860)
SPHERE_CENTER_RADIUS = (8.0, 12.0, 12.0, 3.25)
SPHERE_GEOMETRY_VELOCITY = (0.05, 0.0, 0.0)
SPHERE_WALL_VELOCITY = (0.05, 0.0, 0.0, 0.0)


def translating_sphere_omegas():
    """LatticeScaling for characteristic length 8 cells, U = 0.05, nu = 0.1, magic 3/16 (FP32)."""
    f32 = np.float32
    re = f32(0.05) * f32(8) / f32(0.1)
    nu = f32(0.05) * f32(8) / re
    tau_p = f32(0.5) + f32(3) * nu
    tau_m = f32(0.5) + f32(3.0 / 16.0) / (tau_p - f32(0.5))
    return float(f32(1) / tau_p), float(f32(1) / tau_m)


def translating_sphere_params(force_mode=6):
    op, om = translating_sphere_omegas()
    return bfref.make_params((24, 24, 24), omega_plus=op, omega_minus=om,
                             wall_rule=bfref.WALL_INTERPOLATED_HALF, force_mode=force_mode,
                             load_component=0, selected_part=1, accumulate_loads=True,
                             body_position=SPHERE_CENTER_RADIUS[:3])


def run_translating_sphere_oracle(force_mode=6, steps=40):
    p = translating_sphere_params(force_mode)
    m0, w0 = bfref.translating_sphere(p, SPHERE_CENTER_RADIUS, SPHERE_GEOMETRY_VELOCITY, SPHERE_WALL_VELOCITY, 0.0)
    f = bfref.init_equilibrium(p, m0, w0)
    prev = m0
    fx, cov, unc, transitions = [], 0, 0, 0
    for s in range(1, steps + 1):
        m, w = bfref.translating_sphere(p, SPHERE_CENTER_RADIUS, SPHERE_GEOMETRY_VELOCITY, SPHERE_WALL_VELOCITY, float(s))
        c = int(((m != 0) & (prev == 0)).sum())
        u = int(((m == 0) & (prev != 0)).sum())
        cov, unc = cov + c, unc + u
        transitions += int(c > 0 or u > 0)
        f, loads, _, _ = bfref.step(p, f, prev, m, w)
        fx.append(loads[0])
        prev = m
    return np.array(fx), cov, unc, transitions


# Deetjen dove indexed mesh (ValidationInputs/deetjen-ob-f03-surface-v1)
def deetjen_points(d):
    return {int(fr): d["points"][i] for i, fr in enumerate(d["frames"])}


def deetjen_frame_points(d, frame, points_by_frame):
    """Two-frame point array, first frame index and query time reproducing the
    reference's prepare at the exact time of `frame`.  Interior and first
    frames interpolate with blend 0 (position = that frame); the final frame
    uses the preceding segment with blend 1."""
    times = d["times"]
    last = len(times) - 1
    t = float(times[frame])
    if frame == last:
        first = last - 1
        pts = np.stack([points_by_frame[first], points_by_frame[last]])
    else:
        first = frame
        pts = np.stack([points_by_frame[frame], points_by_frame[frame]])
    return pts, first, t
