#!/usr/bin/env python3
"""Writes the committed golden fixtures under tests/golden/.

Run in the build container only (it reads the read-only reference checkout at
/root/reference); the fixtures it writes are data and travel with the tests.

Sources:
  * Reference/shear_wave_reference.py — the reference's FP64 NumPy D3Q19 TRT
    shear-wave program.  Its equilibrium/stream/moments/collide_trt functions
    produce a full population trajectory at 16^3; run_case produces the
    64^3, 1,000-step decay summary at tau+ = 0.8, magic parameter 1/4.
  * ValidationInputs/deetjen-ob-f03-surface-v1/ — the reference's indexed
    dove mesh (a frame subset is kept; origin and grid are derived from all
    144 frames here).
  * ValidationArtifacts/deetjen-dove-indexed-metal-geometry.json — per-frame
    occupancy SHA-256 and per-part solid-cell counts.
  * ValidationArtifacts/translating-body-topology-summary.json — cover /
    uncover events and mean conservative force of the translating sphere.
  * ValidationArtifacts/deetjen-dove-indexed-production-coupling.json — per
    step newly covered / uncovered cells, persistent wall links and impulses
    of the 8-step indexed-dove replay (frames 0 and 1 of the mesh).
  * ValidationArtifacts/deetjen-dove-moving-wall-link-ray-root.json and
    deetjen-dove-moving-wall-link-intersection.json — the 15 links of the
    exact link-ray-root audit (cells, directions, exact fractions).

Usage: make_golden.py [deetjen] [coupling] [translating] [shear] [rayroots]  (default: all)
"""
from __future__ import annotations

import sys

# Importing the reference's Python program must leave nothing behind in the
# read-only checkout: no __pycache__ is written (or read) for it.
sys.dont_write_bytecode = True

import hashlib  # noqa: E402
import json  # noqa: E402
import os  # noqa: E402

import numpy as np  # noqa: E402

REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def shear_wave():
    sys.path.insert(0, os.path.join(REF, "Reference"))
    import shear_wave_reference as swr  # noqa: E402  (reference FP64 NumPy program)

    n, steps, amplitude = 16, 8, 0.01
    viscosity = 0.1
    magic = 0.25
    tau_plus = 0.5 + 3.0 * viscosity
    tau_minus = 0.5 + magic / (tau_plus - 0.5)
    omega_plus, omega_minus = 1.0 / tau_plus, 1.0 / tau_minus
    rho = np.ones((n, n, n))
    velocity = np.zeros((3, n, n, n))
    y = np.arange(n, dtype=np.float64)
    velocity[0] = amplitude * np.sin(2.0 * np.pi * y / n)[None, :, None]
    f = swr.equilibrium(rho, velocity)
    f0 = f.copy()
    history = []
    for _ in range(steps):
        streamed = swr.stream(f)
        rho, velocity = swr.moments(streamed)
        f = swr.collide_trt(streamed, rho, velocity, omega_plus, omega_minus)
        history.append(f.copy())
    # Reference arrays are indexed [q, x, y, z]; store as direction-major
    # SoA with x fastest: f[q, z, y, x].
    to_soa = lambda a: np.ascontiguousarray(a.transpose(0, 3, 2, 1)).reshape(19, -1)
    np.savez_compressed(
        os.path.join(OUT, "shear_wave_16_fp64.npz"),
        initial=to_soa(f0),
        step1=to_soa(history[0]),
        step8=to_soa(history[-1]),
        omega_plus=omega_plus, omega_minus=omega_minus, amplitude=amplitude, n=n,
    )
    r64 = swr.run_case(resolution=64, steps=1000, viscosity=viscosity,
                       initial_amplitude=amplitude, magic_parameter=magic)
    r32 = swr.run_case()  # the reference program's own default case
    with open(os.path.join(OUT, "shear_wave_decay.json"), "w") as fh:
        json.dump({"n64_tau0.8_magic0.25_1000steps": r64.__dict__,
                   "reference_default_case": r32.__dict__}, fh, indent=1, sort_keys=True)


def deetjen():
    d = os.path.join(REF, "ValidationInputs", "deetjen-ob-f03-surface-v1")
    manifest = json.load(open(os.path.join(d, "manifest.json")))
    art = json.load(open(os.path.join(REF, "ValidationArtifacts", "deetjen-dove-indexed-metal-geometry.json")))
    frames = manifest["frames"]["count"]
    verts = manifest["topology"]["vertexCount"]
    tris = manifest["topology"]["triangleCount"]
    raw = open(os.path.join(d, "positions.f32le"), "rb").read()
    assert hashlib.sha256(raw).hexdigest() == manifest["binary"]["positions"]["sha256"]
    pts = np.frombuffer(raw, "<f4").reshape(frames, verts, 3)
    traw = open(os.path.join(d, "triangles.u16le"), "rb").read()
    assert hashlib.sha256(traw).hexdigest() == manifest["binary"]["triangles"]["sha256"]
    tri = np.frombuffer(traw, "<u2").reshape(tris, 3)
    part = np.zeros(tris, np.uint8)
    for c in manifest["topology"]["components"]:
        part[c["triangleOffset"]:c["triangleOffset"] + c["triangleCount"]] = c["partIdentifier"]
    times = np.asarray(manifest["frames"]["timesSeconds"], np.float32)
    # Domain of MetalIndexedBirdSurfaceValidation.swift:17766-17777 in FP32:
    # origin = min - 4 dx, grid = ceil(extent / dx) + 1 per axis (>= 16).
    dx = np.float32(art["cellSizeMeters"])
    pad = np.float32(4) * dx
    lo = (pts.reshape(-1, 3).min(axis=0) - pad).astype(np.float32)
    hi = (pts.reshape(-1, 3).max(axis=0) + pad).astype(np.float32)
    grid = [max(16, int(np.ceil(np.float32(e / dx))) + 1) for e in (hi - lo).astype(np.float32)]
    assert grid == [art["gridX"], art["gridY"], art["gridZ"]], grid
    # 26 and 27 bracket the ray-root audit's 26.5 ms phase
    keep = sorted(set(list(range(0, frames, 12)) + [1, 26, 27, 33, 89, 126, 134, 142, 143]))
    np.savez_compressed(
        os.path.join(OUT, "deetjen_subset.npz"),
        frames=np.asarray(keep, np.int32), points=pts[keep], times=times, triangles=tri.astype(np.uint32),
        triangle_parts=part, origin=lo, grid=np.asarray(grid, np.int32), dx=dx,
        half_thickness_cells=np.float32(art["halfThicknessCells"]),
    )
    audits = [{"frame": a["frameIndex"], "sha256": a["occupancySHA256"],
               "counts": a["componentSolidCellCounts"], "cpu_compared": a["cpuRasterCompared"]}
              for a in art["frameAudits"]]
    with open(os.path.join(OUT, "deetjen_geometry.json"), "w") as fh:
        json.dump({"grid": grid, "cell_size": float(dx), "half_thickness_cells": art["halfThicknessCells"],
                   "frames": audits}, fh, indent=1)


def translating_body():
    s = json.load(open(os.path.join(REF, "ValidationArtifacts", "translating-body-topology-summary.json")))
    keys = ["gridResolution", "sphereRadiusCells", "translationSpeedLattice", "steps", "newlyCoveredCellEvents",
            "newlyUncoveredCellEvents", "topologyTransitionSteps", "conservativeMeanForceX",
            "rawBudgetMeanForceX", "legacyMeanForceX"]
    with open(os.path.join(OUT, "translating_body.json"), "w") as fh:
        json.dump({k: s[k] for k in keys}, fh, indent=1)


def coupling():
    s = json.load(open(os.path.join(REF, "ValidationArtifacts", "deetjen-dove-indexed-production-coupling.json")))
    keys = ["gridX", "gridY", "gridZ", "cellSizeMeters", "timeStepSeconds", "steps", "periodicBoundaries",
            "spongeStrength", "forceEstimator", "fluidKernel", "linkKernel", "maximumWallSpeedLattice",
            "newlyCoveredCellEvents", "newlyUncoveredCellEvents", "persistentBoundaryLinkEvents",
            "componentSolidCellCounts", "relativeRMSBoundaryClosureResidual", "deviceName"]
    sample_keys = ["step", "sourceTimeSeconds", "newlyCoveredCellCount", "newlyUncoveredCellCount",
                   "persistentBoundaryLinkCount", "sourceLedgerTransitionCellCount", "aerodynamicImpulse",
                   "fluidMomentumBefore", "fluidMomentumAfter", "diagnosticPersistentLinkImpulseToFluid"]
    out = {k: s[k] for k in keys}
    out["samples"] = [{k: smp[k] for k in sample_keys} for smp in s["samples"]]
    with open(os.path.join(OUT, "deetjen_coupling.json"), "w") as fh:
        json.dump(out, fh, indent=1)


def ray_roots():
    """The 15 links of the reference's exact link-ray-root audit
    (deetjen-dove-moving-wall-link-ray-root.json, with the solid / fluid
    cells and production fractions of deetjen-dove-moving-wall-link-
    intersection.json) on its D12 and D16 grids at 26.5 ms: each link's
    solid cell, direction toward the fluid cell, and the fluid-to-
    intersection fraction of the global union's first root (256-interval
    scan + 48 bisections in the reference).  The grids' origin and spacing
    are recovered from the archived intersection points (least squares over
    every record; residual printed)."""
    inter = json.load(open(os.path.join(REF, "ValidationArtifacts", "deetjen-dove-moving-wall-link-intersection.json")))
    roots = json.load(open(os.path.join(REF, "ValidationArtifacts", "deetjen-dove-moving-wall-link-ray-root.json")))
    c = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, -1, 0),
         (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1), (-1, 0, 1), (0, 1, 1), (0, -1, -1), (0, 1, -1),
         (0, -1, 1)]
    out = {"timeSeconds": None, "cases": {}}
    for case in ("d12", "d16"):
        outliers = inter[case]["outliers"]
        rows, rhs = [], []
        for o in outliers:
            cell = np.asarray(o["cellCoordinate"], float)
            d = np.asarray(c[o["directionIndex"]], float)
            for a in range(3):
                r = np.zeros(4)
                r[a] = 1.0
                r[3] = cell[a] + 0.5 + d[a] * (1.0 - o["fluidToIntersectionFraction"])
                rows.append(r)
                rhs.append(o["intersectionMeters"][a])
        sol = np.linalg.lstsq(np.asarray(rows), np.asarray(rhs), rcond=None)[0]
        residual = float(np.abs(np.asarray(rows) @ sol - np.asarray(rhs)).max())
        print(f"{case}: origin {sol[:3]}, dx {sol[3]}, max residual {residual:.2e} m")
        links = []
        for smp in roots[case]["samples"]:
            o = outliers[smp["sourceOutlierIndex"]]
            assert o["cellCoordinate"] == smp["cellCoordinate"] and o["directionIndex"] == smp["directionIndex"]
            links.append({"solidCell": o["cellCoordinate"], "fluidCell": o["neighborCellCoordinate"],
                          "directionIndex": o["directionIndex"],
                          "productionFluidToIntersectionFraction": o["fluidToIntersectionFraction"],
                          "exactGlobalFluidToIntersectionFraction": smp["exactGlobalFluidToIntersectionFraction"],
                          "exactGlobalRootTriangleIndex": smp["exactGlobalRootTriangleIndex"],
                          "exactGlobalRootClosureResidualCells": smp["exactGlobalRootClosureResidualCells"]})
        out["cases"][case] = {"grid": [inter[case]["gridX"], inter[case]["gridY"], inter[case]["gridZ"]],
                              "origin": [float(v) for v in sol[:3]], "cellSize": float(sol[3]),
                              "originResidualMeters": residual, "links": links}
        out["timeSeconds"] = inter[case]["frozenSourceTimeSeconds"]
    out["halfThicknessMeters"] = 0.0075
    with open(os.path.join(OUT, "deetjen_ray_roots.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    parts = sys.argv[1:] or ["deetjen", "coupling", "translating", "shear", "rayroots"]
    if "rayroots" in parts:
        ray_roots()
    if "deetjen" in parts:
        deetjen()
    if "coupling" in parts:
        coupling()
    if "translating" in parts:
        translating_body()
    if "shear" in parts:
        shear_wave()
