"""Coupled linear-momentum ledger of free flight.

Mirrors BirdFlowSimulation.advanceWithCoupledMomentumLedger
(Sources/BirdFlowMetal/BirdFlowSimulation.swift:805-1085): steps one at a
time and closes, per step,

  boundary:  aerodynamic impulse + (fluid change - open boundary - sponge) = 0
  system:    fluid change + body momentum change
             - (open boundary + sponge + gravity + external force impulses) = 0

with the fluid sums from the device ledger (FP64, bf_ledger_*), lattice
momentum scaled by force_to_physical * time_step, the body's translational
momentum mass * v, the aerodynamic impulse as the step's reduced force times
time_step.  The persistent-link exchange is reported beside the boundary
impulse; their difference is the cover / uncover conversion.  Passes when
the RMS residuals relative to the RMS aerodynamic (and external) impulse are
at most `maximum_relative_residual` (the reference's 0.005).  Prescribed
wing momentum is not modelled here (the reference's massless-wing mode).
"""
from __future__ import annotations

import math

import numpy as np

from . import _abi
from .simulation import BirdFlowError, BirdFlowSimulation


def _rms(vectors):
    v = np.asarray(vectors, np.float64).reshape(-1, 3)
    return float(math.sqrt((v * v).sum(axis=1).mean())) if len(v) else 0.0


def advance_with_momentum_ledger(sim: BirdFlowSimulation, steps: int, maximum_relative_residual: float = 0.005):
    params = sim.body_parameters
    if params is None or not params.free_flight:
        raise BirdFlowError(_abi.ERROR_INVALID_ARGUMENT, "the total-system ledger requires free flight")
    if not (steps > 0 and math.isfinite(maximum_relative_residual) and maximum_relative_residual > 0):
        raise BirdFlowError(_abi.ERROR_INVALID_ADVANCE_REQUEST, "steps and the residual bound must be positive")
    cfg = sim.config
    if not cfg.sponge_strength < 1.0 - 1.0e-6:
        raise BirdFlowError(_abi.ERROR_INVALID_ARGUMENT,
                            "sponge strength must be below 0.999999 so the pre-sponge population can be rebuilt")
    dt = float(np.float32(cfg.time_step))
    scale = float(np.float32(cfg.force_to_physical)) * dt
    mass = float(params.mass)
    gravity = np.asarray(params.gravity, np.float64)
    external = np.asarray(params.external_force, np.float64)
    samples = []
    fluid_before = sim.fluid_momentum()
    for _ in range(int(steps)):
        body_before = sim.body_state()
        sim.advance(1)
        src = sim.step_momentum_sources()
        fluid_after = sim.fluid_momentum()
        body_after = sim.body_state()
        load = sim.read_loads()
        change = src["change"][1:] * scale
        open_b = src["open_boundary"][1:] * scale
        sponge = src["sponge"][1:] * scale
        link = src["persistent_link"][1:] * scale
        aero = load[:3] * dt
        grav = gravity * (mass * dt)
        ext = external * dt
        boundary = change - open_b - sponge
        body_change = (body_after[7:10] - body_before[7:10]) * mass
        total = change + body_change
        external_impulse = open_b + sponge + grav + ext
        samples.append({
            "step": sim.step_index,
            "fluid_momentum_before": (fluid_before[1:] * scale).tolist(),
            "fluid_momentum_after": (fluid_after[1:] * scale).tolist(),
            "fluid_momentum_change": change.tolist(),
            "body_momentum_before": (body_before[7:10] * mass).tolist(),
            "body_momentum_after": (body_after[7:10] * mass).tolist(),
            "aerodynamic_impulse": aero.tolist(),
            "gravity_impulse": grav.tolist(),
            "external_force_impulse": ext.tolist(),
            "open_boundary_impulse": open_b.tolist(),
            "sponge_impulse": sponge.tolist(),
            "persistent_link_impulse": link.tolist(),
            "topology_impulse": (boundary - link).tolist(),
            "fluid_boundary_impulse": boundary.tolist(),
            "boundary_residual": (aero + boundary).tolist(),
            "total_momentum_change": total.tolist(),
            "external_impulse": external_impulse.tolist(),
            "system_residual": (total - external_impulse).tolist(),
            "open_boundary_links": src["open_boundary_links"],
            "sponge_cells": src["sponge_cells"],
            "persistent_links": src["persistent_links"],
            "topology_transitions": src["topology_transitions"],
        })
        fluid_before = fluid_after
    aero_rms = _rms([s["aerodynamic_impulse"] for s in samples])
    ext_rms = _rms([s["external_impulse"] for s in samples])
    boundary_rms = _rms([s["boundary_residual"] for s in samples])
    system_rms = _rms([s["system_residual"] for s in samples])
    rel_boundary = boundary_rms / max(aero_rms, 1e-30)
    rel_system = system_rms / max(max(aero_rms, ext_rms), 1e-30)
    finite = all(np.all(np.isfinite(np.asarray(v, np.float64))) for s in samples for v in s.values())
    passed = bool(finite and rel_boundary <= maximum_relative_residual and rel_system <= maximum_relative_residual)
    return {
        "schema": 1,
        "steps": int(steps),
        "time_step": dt,
        "samples": samples,
        "rms_aerodynamic_impulse": aero_rms,
        "rms_external_impulse": ext_rms,
        "rms_boundary_residual": boundary_rms,
        "rms_system_residual": system_rms,
        "relative_boundary_residual": rel_boundary,
        "relative_system_residual": rel_system,
        "maximum_relative_residual": float(maximum_relative_residual),
        "finite": finite,
        "passed": passed,
    }
