#!/usr/bin/env python3
"""Throughput of the D3Q19 moving-boundary step on one MI355X.

Workload (default): the synthetic articulated bird, 512 x 256 x 256 cells
(33.5M), about 40k triangles, RR3 collision, interpolated moving bounce-back
with signed-distance link fractions, conservative momentum-exchange loads
reduced in FP64 every step, link rebuild every step, inflow u_x = 0.04 at -x,
zero-gradient outlet at +x, periodic y/z, tau+ = 0.52.  Synthetic data: the
mesh comes from birdflow_hip.workloads (seed 7), uploaded as an articulated
mesh whose 3-harmonic Fourier wing kinematics the device evaluates every
step (--mesh frames uploads one recorded frame per step instead).

A step is one full pass of the path: vertex prepare, raster, resolve, fluid
(fast + general tiles), load reduction.  value = cells x steps / wall time of
K steps bracketed by stream synchronizations, in MLUPS; the roofline entry is
the fast fluid kernel (populations only: 152 B per cell update) timed with HIP
events recorded on the library's stream inside the timed region.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "numi-gpu-physics_amd")]

HBM_PEAK_GBS = 8000.0
METRIC = "MLUPS + %HBM roofline, moving-bird 512×256×256 FP32 vs FP16-store; adjoint/fwd"
STORAGES = {"fp32": 0, "fp16": 1, "bf16": 2}
ARITHMETIC = {"default": 0, "strict": 1, "fused": 2}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fluid_fast_traffic.json")
POP_BYTES = {"fp32": 152, "fp16": 76, "bf16": 76}
BREAKDOWN_STEPS = 10


def bird_config(shape, storage, arithmetic="default"):
    import birdflow_hip as bh
    tau = 0.52
    return bh.SolverConfig(arithmetic=ARITHMETIC[arithmetic],
        shape=shape, omega_plus=1.0 / tau, omega_minus=1.0 / (0.5 + (3.0 / 16.0) / (tau - 0.5)),
        storage=STORAGES[storage], collision=bh.COLLISION_RR3, wall_rule=bh.WALL_INTERPOLATED_SDF, force_mode=6,
        accumulate_loads=True, boundary=((bh.BC_FARFIELD, bh.BC_OUTLET), (bh.BC_PERIODIC, bh.BC_PERIODIC),
                                         (bh.BC_PERIODIC, bh.BC_PERIODIC)),
        far_velocity=(0.04, 0.0, 0.0), body_position=(160.0 * shape[0] / 512.0, 128.0 * shape[0] / 512.0,
                                                     128.0 * shape[0] / 512.0))


def bird_mesh(steps, shape, form="articulated"):
    """The synthetic bird placed in `shape`, as an articulated mesh (on-device
    kinematics) or as one frame per step; for other than 512 cells along x
    the bird (and its stroke period) scale with the domain, so that
    1024x512x512 is the 512x256x256 case scaled 2x per axis."""
    from birdflow_hip import workloads
    s = shape[0] / 512.0
    geometry = {} if s == 1.0 else dict(body_radii=(40.0 * s, 14.0 * s, 14.0 * s), half_span=100.0 * s,
                                        root_chord=40.0 * s, tip_chord=12.0 * s, tail=(30.0 * s, 20.0 * s),
                                        period=1200.0 * s)
    center = (160.0 * s, 128.0 * s, 128.0 * s)
    if form == "articulated":
        return workloads.synthetic_bird_articulated(seed=7, center=center, **geometry)
    return workloads.synthetic_bird(steps, seed=7, center=center, **geometry)


def upload_bird(sim, mesh):
    if hasattr(mesh, "joints"):
        sim.upload_articulated(mesh)
    else:
        sim.upload_mesh(mesh.points, mesh.times, mesh.triangles, mesh.parts, mesh.half_thickness)


def run_gpu(shape, steps, warmup, storage, arithmetic="default", form="articulated", wall_rule=None):
    import birdflow_hip as bh
    import dataclasses
    mesh = bird_mesh(warmup + steps + BREAKDOWN_STEPS, shape, form)
    cfg = bird_config(shape, storage, arithmetic)
    if wall_rule is not None:
        cfg = dataclasses.replace(cfg, wall_rule=wall_rule)
    sim = bh.BirdFlowSimulation(cfg)
    try:
        upload_bird(sim, mesh)
        sim.init_equilibrium()
        sim.advance(warmup)
        sim.synchronize()
        # timed steps: HIP events around the fast kernel only (on its stream)
        sim.set_profiling(True, level=1)
        t0 = time.perf_counter()
        sim.advance(steps, synchronize=False)
        sim.synchronize()
        t1 = time.perf_counter()
        prof = sim.read_profile()
        loads = sim.load_history(steps)
        census, cov, unc = sim.link_census()
        # untimed: general kernel and geometry durations (more events per step)
        sim.set_profiling(True, level=2)
        sim.advance(BREAKDOWN_STEPS)
        brk = sim.read_profile()
        sim.set_profiling(False)
    finally:
        sim.close()
    n = cfg.cells
    elapsed = t1 - t0
    fast_tiles = prof["tiles"] - prof["general_tiles"]
    fast_cells = fast_tiles * 256
    fast_ms = prof["fast_ms"] / max(prof["steps"], 1)
    achieved = POP_BYTES[storage] * fast_cells / (fast_ms * 1e-3) / 1e9
    step_bytes = (POP_BYTES[storage] + 2) * n
    return {
        "mlups": n * steps / elapsed / 1e6,
        "ms_per_step": elapsed * 1e3 / steps,
        "fast_kernel_ms": fast_ms,
        "general_ms": brk["general_ms"] / max(brk["steps"], 1),
        "geometry_ms": brk["geometry_ms"] / max(brk["steps"], 1),
        "fast_cells": fast_cells,
        "general_tiles": prof["general_tiles"],
        "achieved_gbs": achieved,
        "step_gbs": step_bytes * steps / elapsed / 1e9,
        "links": int(census.sum()), "covered": cov, "uncovered": unc,
        "force_last": loads[-1, :3].tolist() if len(loads) else None,
        "finite": bool(np.isfinite(loads).all()),
    }


def run_sdf_bird(shape, steps, warmup):
    """The reference's product path on its demonstration bird
    (BirdParameters.demonstration, analytic signed-distance geometry every
    step): free flight under gravity, far field -10 m/s in x, TRT, halfway
    walls, force mode 6, sponge 0.04 over 8 cells, FP32; dx = 4 mm (root chord
    35 cells), Re 500 on the chord (tau+ 0.5105).  MLUPS of whole steps
    (geometry, fluid, loads, body update)."""
    import birdflow_hip as bh
    sc = bh.LatticeScaling.make(0.14, 35, 10.0, 500.0, 1.225, lattice_reference_speed=0.05)
    dx = sc.cell_size
    cfg = bh.SolverConfig(shape=shape, omega_plus=sc.omega_plus, omega_minus=sc.omega_minus,
                          storage=bh.STORAGE_FP32, collision=bh.COLLISION_TRT, wall_rule=bh.WALL_HALFWAY,
                          force_mode=6, accumulate_loads=True, boundary=((bh.BC_FARFIELD, bh.BC_FARFIELD),) * 3,
                          cell_size=dx, time_step=sc.time_step, velocity_to_lattice=sc.velocity_to_lattice,
                          force_to_physical=sc.force_to_physical, sponge_strength=0.04, sponge_width=8.0,
                          far_velocity=(-10.0 * sc.velocity_to_lattice, 0.0, 0.0),
                          body_position=(0.3 * shape[0] * dx, 0.5 * shape[1] * dx, 0.5 * shape[2] * dx))
    sim = bh.BirdFlowSimulation(cfg)
    try:
        state0 = list(cfg.body_position) + [0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0]
        sim.set_body(state0, bh.BodyParameters(mass=0.42, inertia=(0.0024, 0.0068, 0.0074), gravity=(0.0, 0.0, -9.81)))
        sim.set_bird(bh.BirdGeometry.demonstration())
        sim.init_equilibrium()
        mask, _ = sim.read_geometry()
        solid = int((mask != 0).sum())
        sim.advance(warmup)
        sim.synchronize()
        t0 = time.perf_counter()
        sim.advance(steps, synchronize=False)
        sim.synchronize()
        t1 = time.perf_counter()
        loads = sim.load_history(steps)
        sim.set_profiling(True, level=2)
        sim.advance(BREAKDOWN_STEPS)
        brk = sim.read_profile()
        sim.set_profiling(False)
        body = sim.body_state()
        last = sim.read_loads()
    finally:
        sim.close()
    n = cfg.cells
    k = max(brk["steps"], 1)
    return {"value": n * steps / (t1 - t0) / 1e6, "unit": "MLUPS", "ms_per_step": (t1 - t0) * 1e3 / steps,
            "steps": steps, "solid_cells": solid, "general_tiles": int(brk["general_tiles"]), "tiles": int(brk["tiles"]),
            "step_breakdown_ms": {"geometry": brk["geometry_ms"] / k, "fluid_fast": brk["fast_ms"] / k,
                                  "fluid_general": brk["general_ms"] / k},
            "finite": bool(np.isfinite(loads).all() and np.isfinite(body).all() and np.isfinite(last).all()),
            "body_velocity_end": body[7:10].tolist(),
            "workload": f"analytic demonstration bird {shape[0]}x{shape[1]}x{shape[2]}, free flight, TRT, "
                        "halfway walls, dx 4 mm, Re 500, FP32"}


def run_wing(shape, steps, warmup):
    """The prescribed Li & Nabawy beta-planform wing (BirdFlow.metal:940-1297)
    with the link-table wall rule: 32-cell mean chord, span 96 cells rooted
    at the grid centre, cycle 8192 steps (tip speed at most 0.08 lattice),
    far field at rest, sponge 0.04 over 8 cells, TRT at tau+ 0.577 (Re 100
    on the chord at 0.08), force mode 6, FP32.  MLUPS of whole steps
    (frame, mask, wall, link bisection, fluid, loads)."""
    import birdflow_hip as bh
    root = (0.5 * shape[0], 0.5 * shape[1], 0.5 * shape[2])
    tau = 3.0 * 0.08 * 32.0 / 100.0 + 0.5
    omega_plus = 1.0 / tau
    omega_minus = 1.0 / (0.1875 / (tau - 0.5) + 0.5)
    cfg = bh.SolverConfig(shape=shape, omega_plus=omega_plus, omega_minus=omega_minus, storage=bh.STORAGE_FP32,
                          collision=bh.COLLISION_TRT, wall_rule=bh.WALL_INTERPOLATED_LINKS, force_mode=6,
                          accumulate_loads=True, boundary=((bh.BC_FARFIELD, bh.BC_FARFIELD),) * 3, cell_size=1.0,
                          time_step=1.0, velocity_to_lattice=1.0, sponge_strength=0.04, sponge_width=8.0,
                          body_position=root)
    sim = bh.BirdFlowSimulation(cfg)
    try:
        sim.set_flapping_wing(bh.FlappingWing.li_nabawy(32.0, root, 8192.0))
        sim.init_equilibrium()
        sim.advance(warmup)
        sim.synchronize()
        t0 = time.perf_counter()
        sim.advance(steps, synchronize=False)
        sim.synchronize()
        t1 = time.perf_counter()
        loads = sim.load_history(steps)
        mask, _ = sim.read_geometry()
        sim.set_profiling(True, level=2)
        sim.advance(BREAKDOWN_STEPS)
        brk = sim.read_profile()
        sim.set_profiling(False)
    finally:
        sim.close()
    n = cfg.cells
    k = max(brk["steps"], 1)
    return {"value": n * steps / (t1 - t0) / 1e6, "unit": "MLUPS", "ms_per_step": (t1 - t0) * 1e3 / steps,
            "steps": steps, "solid_cells": int((mask != 0).sum()), "general_tiles": int(brk["general_tiles"]),
            "tiles": int(brk["tiles"]),
            "step_breakdown_ms": {"geometry": brk["geometry_ms"] / k, "fluid_fast": brk["fast_ms"] / k,
                                  "fluid_general": brk["general_ms"] / k},
            "finite": bool(np.isfinite(loads).all()),
            "workload": f"Li & Nabawy beta-planform wing {shape[0]}x{shape[1]}x{shape[2]}, chord 32 cells, "
                        "link-table walls, TRT, FP32"}


def omp_threads(n):
    """Sets the oracle's OpenMP thread count (libgomp, loaded with it)."""
    import ctypes
    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))


def run_cpu_baseline(shape, steps, threads=None, form="articulated"):
    """The oracle (C, OpenMP over z) on a bounded sample of the same workload,
    on `threads` OpenMP threads (default: OMP_NUM_THREADS or all host cores)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import bfref
    all_threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    omp_threads(threads or all_threads)
    mesh = bird_mesh(steps, shape, form)
    cfg = bird_config(shape, "fp32")
    p = bfref.make_params(shape, omega_plus=cfg.omega_plus, omega_minus=cfg.omega_minus,
                          bc=((bfref.BC_FARFIELD, bfref.BC_OUTLET), (0, 0), (0, 0)), collision=bfref.COLLISION_RR3,
                          wall_rule=bfref.WALL_INTERPOLATED_SDF, force_mode=6, accumulate_loads=True,
                          far_u=(0.04, 0.0, 0.0), body_position=cfg.body_position)

    if hasattr(mesh, "joints"):
        k = bfref.make_kinematics(mesh.period, mesh.amplitude_scale, mesh.center,
                                  [dict(side=j.side, root=j.root, stroke=j.stroke, deviation=j.deviation,
                                        pitch=j.pitch, twist=j.twist) for j in mesh.joints],
                                  harmonics=mesh.harmonics)

        def vertices(t):
            return bfref.articulated_vertices(mesh.rest, mesh.vertex_joint, mesh.span_fraction, k, t)
    else:
        def vertices(t):
            first, _, _ = bfref.frame_interval(mesh.times, t)
            return bfref.prepare_vertices(mesh.points, mesh.times, first, t)

    def geometry(t):
        pos, vel = vertices(t)
        mask, wall, _ = bfref.raster_resolve(p, pos, vel, mesh.triangles, mesh.parts, mesh.half_thickness, 64)
        return mask, wall

    mask0, wall0 = geometry(0.0)
    f = bfref.init_equilibrium(p, mask0, wall0, np.float32)
    prev = mask0
    t0 = time.perf_counter()
    for s in range(1, steps + 1):
        mask, wall = geometry(float(s))
        f, _, _, _ = bfref.step(p, f, prev, mask, wall)
        prev = mask
    elapsed = time.perf_counter() - t0
    omp_threads(all_threads)
    n = shape[0] * shape[1] * shape[2]
    return n * steps / elapsed / 1e6, threads or all_threads, elapsed


def plate_theta(seed=7):
    """Kinematic parameters of the differentiation workload: the flapping
    plate's heave 8 sin(2 pi t / 800) and pitch 30 deg sin(2 pi t / 800 +
    pi / 2) as first harmonics, second harmonics of a quarter of that
    amplitude, each of the 8 values drawn +-20% with the seed."""
    base = np.array([8.0, 0.0, 2.0, 0.0, np.radians(30.0), np.pi / 2, np.radians(7.5), np.pi / 2])
    rng = np.random.default_rng(seed)
    return base * (1.0 + 0.2 * rng.uniform(-1.0, 1.0, 8))


def run_adjoint(shape, steps, interval):
    """Adjoint cost on the differentiation workload: the two-harmonic flapping
    plate (span 48, chord 16, thickness 2 cells at x = 64; 8 kinematic
    parameters), RR3, interpolated walls, far field in x, periodic y / z,
    u_x = 0.05; J = lift averaged over the second half of the steps.  Returns
    forward and reverse times per step."""
    import birdflow_hip as bh
    from birdflow_hip import workloads
    s = shape[0] / 256.0
    tau = 0.6
    cfg = bh.SolverConfig(
        shape=shape, omega_plus=1.0 / tau, omega_minus=1.0 / (0.5 + (3.0 / 16.0) / (tau - 0.5)),
        collision=bh.COLLISION_RR3, wall_rule=bh.WALL_INTERPOLATED_SDF, force_mode=6, accumulate_loads=True,
        boundary=((bh.BC_FARFIELD, bh.BC_FARFIELD), (0, 0), (0, 0)), far_velocity=(0.05, 0.0, 0.0),
        body_position=(64.0 * s, 64.0 * s, 64.0 * s))
    theta = plate_theta()
    seq, _, tan = workloads.parametric_plate(theta, steps, center=(64.0 * s, 64.0 * s, 64.0 * s), span=48.0 * s,
                                             chord=16.0 * s, thickness=2.0, period=800.0, half_thickness=0.5)
    weights = np.zeros((steps, 6))
    weights[steps // 2:, 2] = 1.0 / (steps - steps // 2)
    with bh.BirdFlowSimulation(cfg) as sim:
        sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
        sim.adjoint_set_tangents(tan)
        sim.init_equilibrium()
        J, grad, t = sim.adjoint_gradient(steps, weights, interval)
    return {"forward_ms_per_step": t["forward_ms"] / steps, "reverse_ms_per_step": t["reverse_ms"] / steps,
            "ratio_reverse_vs_forward": t["reverse_ms"] / t["forward_ms"],
            "ratio_gradient_vs_forward": (t["forward_ms"] + t["reverse_ms"]) / t["forward_ms"],
            "objective": J, "gradient": grad.tolist(), "finite": bool(np.isfinite(grad).all()),
            "config": {"workload": f"parametric flapping plate {shape[0]}x{shape[1]}x{shape[2]}, RR3, "
                                   "interpolated walls, 8 kinematic parameters (seed 7), lift averaged over "
                                   "the second half", "steps": steps, "checkpoint_interval": interval}}


def measured_traffic(storage, arithmetic, cells):
    """HBM bytes of one fast-kernel launch over `cells` cells, from the
    committed rocprofv3 PMC passes (FETCH_SIZE and WRITE_SIZE per dispatch,
    calibrated on a copy kernel) for this storage and arithmetic: the
    measured bytes per cell times the cells of this launch (the number of
    wall-free tiles moves by a few per step as the bird flaps).  None when no
    such measurement exists."""
    try:
        rec = json.load(open(TRAFFIC_FILE))[f"{storage}/{arithmetic}"]
        per_cell = rec["bytes_per_launch"] / rec["cells_per_launch"]
    except (OSError, KeyError, ValueError, ZeroDivisionError):
        return None
    return per_cell * cells


def cpu_model():
    """The host CPU's model name (the CPU baseline's hardware)."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--storage", choices=list(STORAGES), default="fp32")
    ap.add_argument("--arithmetic", choices=list(ARITHMETIC), default="default",
                    help="collision arithmetic; default: strict for fp32, fused for fp16/bf16")
    ap.add_argument("--shape", type=int, nargs=3, default=(512, 256, 256))
    ap.add_argument("--mesh", choices=("articulated", "frames"), default="articulated",
                    help="the bird's upload form: on-device kinematics, or one recorded frame per step")
    ap.add_argument("--ray-steps", type=int, default=30,
                    help="timed steps of the same bird with exact ray-root link fractions (0: skip)")
    ap.add_argument("--cpu-steps", type=int, default=2, help="oracle steps timed for the CPU baseline (0: skip)")
    ap.add_argument("--cpu-single-steps", type=int, default=1,
                    help="oracle steps timed on one thread beside it (0: skip)")
    ap.add_argument("--skip-fp16", action="store_true",
                    help="skip the FP16-offset storage form (reported beside FP32 by default)")
    ap.add_argument("--adjoint-steps", type=int, default=4000,
                    help="steps of the adjoint cost measurement on the 256x128x128 plate (0: skip)")
    ap.add_argument("--adjoint-interval", type=int, default=100, help="checkpoint interval of that sweep")
    ap.add_argument("--sdf-bird-steps", type=int, default=100,
                    help="timed steps of the analytic demonstration bird on the same grid (0: skip)")
    ap.add_argument("--wing-steps", type=int, default=100,
                    help="timed steps of the prescribed Li & Nabawy wing on a 384x384x192 grid (0: skip)")
    ap.add_argument("--wing-only", action="store_true",
                    help="run only the prescribed wing and print its result line (profiling)")
    ap.add_argument("--sdf-bird-only", action="store_true",
                    help="run only the analytic demonstration bird and print its result line (profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "requested_gpus": args.gpus}), flush=True)
        sys.exit(2)
    shape = tuple(args.shape)
    if args.wing_only:
        wr = run_wing((384, 384, 192), args.wing_steps, args.warmup)
        print(json.dumps(dict(wr, metric="prescribed flapping wing step", n_gpus=1, warmup=args.warmup,
                              dtype="f32", data="synthetic")), flush=True)
        return
    if args.sdf_bird_only:
        sdf = run_sdf_bird(shape, args.sdf_bird_steps, args.warmup)
        print(json.dumps(dict(sdf, metric="analytic demonstration bird step", n_gpus=1, warmup=args.warmup,
                              dtype="f32", data="synthetic")), flush=True)
        return
    print(f"# bench: bird {shape}, storage {args.storage}, warmup {args.warmup}, steps {args.steps}",
          file=sys.stderr, flush=True)
    import birdflow_hip as bh
    copy_gbs = bh.copy_bandwidth(1 << 30, 20)
    print(f"# copy kernel: {copy_gbs:.1f} GB/s (1 GiB buffers)", file=sys.stderr, flush=True)
    r = run_gpu(shape, args.steps, args.warmup, args.storage, args.arithmetic, args.mesh)
    print(f"# gpu: {r}", file=sys.stderr, flush=True)
    extra = {"copy_kernel_gbs": copy_gbs}
    if not r["finite"]:
        extra["error"] = "non-finite loads: the run diverged"
    if args.storage == "fp32" and not args.skip_fp16:
        r16 = run_gpu(shape, args.steps, args.warmup, "fp16", args.arithmetic, args.mesh)
        print(f"# gpu fp16: {r16}", file=sys.stderr, flush=True)
        extra["fp16_store"] = {"value": r16["mlups"], "unit": "MLUPS", "ratio_vs_fp32": r16["mlups"] / r["mlups"],
                               "fast_kernel_ms": r16["fast_kernel_ms"], "achieved_gbs": r16["achieved_gbs"]}
    if args.ray_steps > 0:
        rr = run_gpu(shape, args.ray_steps, args.warmup, args.storage, args.arithmetic, args.mesh,
                     wall_rule=bh.WALL_INTERPOLATED_RAY)
        print(f"# gpu, exact ray roots: {rr}", file=sys.stderr, flush=True)
        extra["exact_ray_roots"] = {"value": rr["mlups"], "unit": "MLUPS", "ms_per_step": rr["ms_per_step"],
                                    "geometry_ms": rr["geometry_ms"], "fast_kernel_ms": rr["fast_kernel_ms"],
                                    "steps": args.ray_steps,
                                    "workload": "the same bird with BF_WALL_INTERPOLATED_RAY link fractions"}
    if args.sdf_bird_steps > 0:
        sdf = run_sdf_bird(shape, args.sdf_bird_steps, args.warmup)
        print(f"# analytic bird: {sdf}", file=sys.stderr, flush=True)
        extra["analytic_bird"] = sdf
    if args.wing_steps > 0:
        wr = run_wing((384, 384, 192), args.wing_steps, args.warmup)
        print(f"# prescribed wing: {wr}", file=sys.stderr, flush=True)
        extra["prescribed_wing"] = wr
    if args.adjoint_steps > 0:
        adj = run_adjoint((256, 128, 128), args.adjoint_steps, args.adjoint_interval)
        print(f"# adjoint: {adj}", file=sys.stderr, flush=True)
        extra["adjoint"] = adj
    cpu = None
    if args.cpu_steps > 0:
        mlups, threads, elapsed = run_cpu_baseline(shape, args.cpu_steps, form=args.mesh)
        print(f"# cpu: {mlups:.2f} MLUPS over {args.cpu_steps} steps ({elapsed:.1f} s, {threads} threads)",
              file=sys.stderr, flush=True)
        cpu = {"value": mlups, "unit": "MLUPS", "cores": threads, "kind": "port", "cpu_model": cpu_model(),
               "host_cpus": os.cpu_count(),
               "sample": f"oracle (C, OpenMP over z) on the same bird workload, {args.cpu_steps} full steps "
                         f"(geometry + fluid) at {shape[0]}x{shape[1]}x{shape[2]}"}
        if args.cpu_single_steps > 0:
            m1, _, e1 = run_cpu_baseline(shape, args.cpu_single_steps, threads=1, form=args.mesh)
            print(f"# cpu, 1 thread: {m1:.2f} MLUPS over {args.cpu_single_steps} steps ({e1:.1f} s)",
                  file=sys.stderr, flush=True)
            cpu["single_thread"] = {"value": m1, "unit": "MLUPS", "cores": 1,
                                    "sample": f"the same, {args.cpu_single_steps} full step(s), one thread"}
    line = {
        "metric": METRIC,
        "value": r["mlups"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.storage == "fp32" else f"{args.storage}-offset-store/f32-compute",
        "data": "synthetic",
        "config": {"workload": f"moving bird {shape[0]}x{shape[1]}x{shape[2]}, ~40k-triangle synthetic "
                               "articulated mesh, RR3, interpolated moving bounce-back, FP64 loads, "
                               "link rebuild every step",
                   "cells": shape[0] * shape[1] * shape[2], "storage": args.storage,
                   "arithmetic": args.arithmetic,
                   "kinematics": "on-device Fourier (3 harmonics)" if args.mesh == "articulated"
                   else "one uploaded frame per step"},
        "roofline": {"bound": "hbm", "achieved": r["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": r["achieved_gbs"] / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.storage, args.arithmetic, r["fast_cells"]),
                     "kernel": "fluid_fast", "bytes_per_cell": POP_BYTES[args.storage],
                     "cells_per_launch": r["fast_cells"], "avg_ms": r["fast_kernel_ms"]},
        "step_breakdown_ms": {"geometry": r["geometry_ms"], "fluid_fast": r["fast_kernel_ms"],
                              "fluid_general": r["general_ms"]},
        "step_roofline_frac": r["step_gbs"] / HBM_PEAK_GBS,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
