"""GPU checks of the hand-written adjoint.

* dJ/df at the start state (the whole reverse sweep through stream, collide,
  wall links, sponge, far-field and outlet faces, checkpoint / recompute)
  against central finite differences of the FP64 oracle; also for TRT with
  the symmetric limiter binding in many cells.
* dJ/dtheta of the plate with FP16 / BF16 offset storage against the FP64
  reference gradient (the storage's error stated).
* dJ/dtheta of the plate with exact-root link fractions (each fraction
  differentiated through its entry triangle's shell root) against the FP64
  oracle with the FP64 roots of the frozen entry triangles, bar 1e-3.
* dJ/dtheta of a flapping plate's time-averaged lift with respect to its 8
  kinematic parameters against an FP64 oracle reference gradient: central
  differences of the FP64 step with FP64 wall fields on the same frozen link
  sets (masks and nearest triangles of each step), the quantity the adjoint
  differentiates.  Bar: 1e-3 relative.
* the same gradient against the solver's own FP32 central differences at
  two step sizes with the unperturbed run's masks replayed
  (bf_mesh_freeze_masks); differences with re-rastered masks are printed
  beside.
* the differentiation workload's inlet / outlet plate on its 64 x 32 x 32
  twin over one full flapping period against the FP64 oracle's central
  differences, bar 1e-3.
* an articulated bird (on-device Fourier kinematics; 96 x 48 x 48, RR3,
  inlet / outlet) with 6 of its Fourier coefficients as parameters, placed
  in the world and by a rotated, spinning body pose (move_mesh): dJ/dtheta
  against central differences of the FP64 oracle objective, the vertices
  from the FP64 kinematics (articulated_vertices_f64), on the frozen link
  sets.  Bar: 1e-3 relative.
* the torch.autograd wrapper returns the adjoint gradient (frame tangents
  and kinematic parameters).
"""
import dataclasses
import sys

import numpy as np
import pytest

import bfref
import birdflow_hip as bh
import cases
from birdflow_hip import workloads

pytestmark = pytest.mark.gpu

THETA0 = np.array([2.0, 0.1, 0.5, 0.3, 0.3, 1.2, 0.1, 0.2])
PLATE = dict(center=(32.0, 24.0, 24.0), span=20.0, chord=8.0, thickness=2.0, period=200.0, half_thickness=0.5)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def objective_oracle(p, f0, steps, weights, geometry, dtype=np.float64):
    """Oracle objective (FP64 unless dtype says otherwise); geometry(s) ->
    (mask_prev, mask, wall)."""
    f = np.asarray(f0, dtype)
    J = 0.0
    for s in range(1, steps + 1):
        prev, cur, wall = geometry(s)
        f, loads, _, _ = bfref.step(p, f, prev, cur, wall)
        J += float(np.dot(weights[s - 1], loads[:6]))
    return J


FACES = {"far field": ((1, 1), (1, 1), (1, 1)),
         "inlet / outlet": ((bfref.BC_FARFIELD, bfref.BC_OUTLET), (0, 0), (1, 1)),
         "outlets on both x faces": ((bfref.BC_OUTLET, bfref.BC_OUTLET), (1, 1), (0, 0))}


@pytest.mark.parametrize("faces", list(FACES))
@pytest.mark.parametrize("collision", [bh.COLLISION_TRT, bh.COLLISION_REGULARIZED, bh.COLLISION_RR3])
def test_adjoint_sensitivity_to_start_state_matches_fp64_differences(collision, faces):
    say(f"case: adjoint dJ/df0, static walls, {faces} faces + sponge, collision {collision}, 6 steps, 3 segments")
    rng = np.random.default_rng(40 + collision)
    # whole 64-cell tiles along x (the wall-free adjoint kernel at the x faces)
    # and a block 3 cells from each x face: the faces are within the 6 steps'
    # reach of the walls
    shape = (64, 16, 14)
    n = shape[0] * shape[1] * shape[2]
    p = bfref.make_params(shape, omega_plus=1.0 / 0.56, omega_minus=1.0 / 1.1, bc=FACES[faces],
                          collision=collision, wall_rule=bfref.WALL_INTERPOLATED_SDF, force_mode=6,
                          accumulate_loads=True, sponge_strength=0.3, sponge_width=4.0, far_u=(0.05, 0.01, 0.0),
                          body_position=(32.0, 8.0, 7.0))
    solid = np.zeros((shape[2], shape[1], shape[0]), np.uint8)
    solid[5:9, 5:11, 3:9] = 2
    solid[5:9, 5:11, 55:61] = 3
    wall = np.zeros((n, 4), np.float32)
    wall[:, :3] = rng.uniform(-0.04, 0.04, (n, 3))
    wall[:, 3] = rng.uniform(-1.0, 2.0, n)
    f0 = (bfref.init_equilibrium(p, dtype=np.float32).reshape(19, n)
          * rng.uniform(0.97, 1.03, (19, n)).astype(np.float32)).ravel()
    steps = 6
    weights = rng.normal(size=(steps, 6))
    with bh.BirdFlowSimulation(cases.solver_config(p)) as sim:
        sim.set_geometry(solid.ravel(), wall)
        sim.write_populations(f0)
        J, grad, timings = sim.adjoint_gradient(steps, weights, checkpoint_interval=2)
        lam = sim.adjoint_sensitivity().astype(np.float64)
    geo = lambda s: (solid.ravel(), solid.ravel(), wall)
    J64 = objective_oracle(p, f0, steps, weights, geo)
    J32 = objective_oracle(p, f0, steps, weights, geo, dtype=np.float32)
    face = np.abs(lam.reshape(19, shape[2], shape[1], shape[0])[:, :, :, [0, -1]]).sum()
    say(f"  J gpu {J:.9e} vs FP32 oracle {J32:.9e}, FP64 oracle {J64:.9e}; |lambda| {np.linalg.norm(lam):.3e}, "
        f"on the x faces {face:.3e}; timings {timings}")
    assert face > 0.0                    # the faces' populations reach J
    # the forward pass is the FP32 step: the FP32 oracle's objective to the
    # order of its FP64 load sums; FP64 differs by FP32 rounding over 6 steps
    assert abs(J - J32) <= 1e-9 * max(1.0, abs(J32))
    assert abs(J - J64) <= 5e-5 * max(1.0, abs(J64))
    for trial in range(2):
        delta = rng.normal(size=19 * n) * f0
        h = 1e-6
        fd = (objective_oracle(p, f0 + h * delta, steps, weights, geo)
              - objective_oracle(p, f0 - h * delta, steps, weights, geo)) / (2 * h)
        ad = float(np.dot(lam, delta))
        say(f"  direction {trial}: adjoint {ad:.9e}, FP64 central difference {fd:.9e}, rel {abs(ad - fd) / abs(fd):.2e}")
        assert abs(ad - fd) <= 2e-4 * abs(fd)


def test_adjoint_of_the_symmetric_limiter_where_it_binds():
    # TRT with the symmetric limiter (far-field faces, static walls) from a
    # start state whose 6 x 6 x 6 interior block is far from equilibrium, so
    # that the limiter scales the symmetric increment of many cells in the
    # first steps: dJ/df0 against FP64 central differences of the oracle.
    say("case: adjoint dJ/df0, TRT with the symmetric limiter binding, static walls, far field, 4 steps")
    rng = np.random.default_rng(77)
    shape = (64, 16, 14)
    n = shape[0] * shape[1] * shape[2]
    p = bfref.make_params(shape, omega_plus=1.0 / 0.52, omega_minus=1.0 / 1.1, bc=FACES["far field"],
                          collision=bfref.COLLISION_TRT_SYMMETRIC_LIMITER, wall_rule=bfref.WALL_INTERPOLATED_SDF,
                          force_mode=6, accumulate_loads=True, far_u=(0.05, 0.01, 0.0), body_position=(32.0, 8.0, 7.0))
    solid = np.zeros((shape[2], shape[1], shape[0]), np.uint8)
    solid[5:9, 5:11, 40:46] = 2
    wall = np.zeros((n, 4), np.float32)
    wall[:, :3] = rng.uniform(-0.04, 0.04, (n, 3))
    wall[:, 3] = rng.uniform(-1.0, 2.0, n)
    mult = rng.uniform(0.97, 1.03, (19, shape[2], shape[1], shape[0]))
    mult[:, 4:10, 5:11, 12:18] = rng.uniform(0.02, 1.98, (19, 6, 6, 6))
    f0 = (bfref.init_equilibrium(p, dtype=np.float32).reshape(19, n) * mult.reshape(19, n).astype(np.float32)).ravel()
    steps = 4
    weights = rng.normal(size=(steps, 6))
    with bh.BirdFlowSimulation(cases.solver_config(p)) as sim:
        sim.set_geometry(solid.ravel(), wall)
        sim.write_populations(f0)
        J, grad, timings = sim.adjoint_gradient(steps, weights, checkpoint_interval=2)
        lam = sim.adjoint_sensitivity().astype(np.float64)
    geo = lambda s: (solid.ravel(), solid.ravel(), wall)
    J32 = objective_oracle(p, f0, steps, weights, geo, dtype=np.float32)
    _, loads, _, _ = bfref.step(p, np.asarray(f0, np.float64), solid.ravel(), solid.ravel(), wall)
    say(f"  J gpu {J:.9e} vs FP32 oracle {J32:.9e}; cells limited in step 1: {int(loads[6])}")
    assert loads[6] >= 30                # 44 cells limited in step 1, 79 over the 4 steps
    assert abs(J - J32) <= 1e-9 * max(1.0, abs(J32))
    for trial in range(2):
        delta = rng.normal(size=19 * n) * f0
        h = 1e-7
        fd = (objective_oracle(p, f0 + h * delta, steps, weights, geo)
              - objective_oracle(p, f0 - h * delta, steps, weights, geo)) / (2 * h)
        ad = float(np.dot(lam, delta))
        say(f"  direction {trial}: adjoint {ad:.9e}, FP64 central difference {fd:.9e}, rel {abs(ad - fd) / abs(fd):.2e}")
        assert abs(ad - fd) <= 2e-4 * abs(fd)


def plate_case(steps):
    p = cases.plate_params(shape=(96, 48, 48))
    p.bc[0][1] = bfref.BC_FARFIELD          # far field on both x faces (the inlet / outlet plate is below)
    weights = np.zeros((steps, 6))
    weights[steps // 3:, 2] = 1.0 / (steps - steps // 3)      # time-averaged lift (z force)
    return p, weights


def plate_frames(theta, steps):
    seq, pts64, tan = workloads.parametric_plate(theta, steps, **PLATE)
    return seq, pts64, tan


def test_adjoint_gradient_of_plate_lift_matches_fp64_reference():
    steps = 30
    p, weights = plate_case(steps)
    seq, _, tan = plate_frames(THETA0, steps)
    say(f"case: adjoint d(mean lift)/d theta, parametric plate {p.nx}x{p.ny}x{p.nz} RR3, {steps} steps, 8 parameters")
    with bh.BirdFlowSimulation(cases.solver_config(p)) as sim:
        sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
        sim.adjoint_set_tangents(tan)
        sim.init_equilibrium()
        f0 = sim.read_populations()
        J, grad, timings = sim.adjoint_gradient(steps, weights, checkpoint_interval=10)
    # frozen link sets: masks and nearest triangles of each step at theta0
    masks, tris = [], []
    for s in range(steps + 1):
        t = float(seq.times[0]) + s
        first, _, _ = bfref.frame_interval(seq.times, t)
        pos, vel = bfref.prepare_vertices(seq.points, seq.times, first, t)
        mask, _, keys = bfref.raster_resolve(p, pos, vel, seq.triangles, seq.parts, seq.half_thickness, 64)
        masks.append(mask)
        tris.append(np.where(keys != np.iinfo(np.uint64).max, (keys & 0xFFFFFFFF).astype(np.int64), -1))

    def J64(theta):
        _, pts64, _ = plate_frames(theta, steps)

        def geo(s):
            pos, vel = bfref.interpolate_frames_f64(pts64, seq.times, float(seq.times[0]) + s)
            return masks[s - 1], masks[s], bfref.wall_field_f64(p, pos, vel, seq.triangles, tris[s],
                                                                 seq.half_thickness)
        return objective_oracle(p, f0, steps, weights, geo)

    J0 = J64(THETA0)
    ref = np.zeros(8)
    for k in range(8):
        h = 1e-5
        e = np.zeros(8)
        e[k] = h
        ref[k] = (J64(THETA0 + e) - J64(THETA0 - e)) / (2 * h)
    err = np.abs(grad - ref).max() / np.abs(ref).max()
    say(f"  J gpu {J:.9e} vs FP64 reference {J0:.9e}")
    for k, name in enumerate(workloads.PLATE_PARAMETERS):
        say(f"  {name:14s} adjoint {grad[k]: .6e}  FP64 reference {ref[k]: .6e}")
    say(f"  max |adjoint - reference| / max |reference| = {err:.2e}; timings {timings}")
    assert abs(J - J0) <= 1e-4 * abs(J0)
    assert err <= 1e-3


# bars for (|J - J64| / |J64|, gradient error): measured 1.7e-5 / 3.9e-5 (FP16)
# and 7.9e-4 / 8.5e-4 (BF16), FP32 storage from the same start 9e-6
OFFSET_BARS = {bh.STORAGE_FP16_OFFSET: (2e-4, 5e-4), bh.STORAGE_BF16_OFFSET: (3e-3, 5e-3)}


@pytest.mark.parametrize("storage", [bh.STORAGE_FP16_OFFSET, bh.STORAGE_BF16_OFFSET], ids=["fp16", "bf16"])
def test_adjoint_with_offset_storage_against_fp64_reference(storage):
    # The same plate with populations stored as FP16 / BF16 offsets: the
    # forward steps and the recompute run in that storage, lambda in FP32,
    # the storage rounding taken as the identity.  Against the FP64
    # reference gradient (no storage rounding) the error is the storage's;
    # the FP32-storage gradient from the same start state is printed beside.
    steps = 30
    p, weights = plate_case(steps)
    seq, _, tan = plate_frames(THETA0, steps)
    name = {bh.STORAGE_FP16_OFFSET: "FP16", bh.STORAGE_BF16_OFFSET: "BF16"}[storage]
    say(f"case: adjoint d(mean lift)/d theta, parametric plate {p.nx}x{p.ny}x{p.nz} RR3, {steps} steps, "
        f"{name} offset storage")
    grads = {}
    for st in (storage, bh.STORAGE_FP32):
        with bh.BirdFlowSimulation(cases.solver_config(p, storage=st)) as sim:
            sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
            sim.adjoint_set_tangents(tan)
            if st == storage:
                sim.init_equilibrium()
                f0 = sim.read_populations()          # decoded from the stored offsets
            else:
                sim.write_populations(f0)            # FP32 holds those values exactly
            J, grad, _ = sim.adjoint_gradient(steps, weights, checkpoint_interval=10)
            grads[st] = (J, grad)
    masks, tris = frozen_links(p, seq, steps)

    def J64(theta):
        _, pts64, _ = plate_frames(theta, steps)

        def geo(s):
            pos, vel = bfref.interpolate_frames_f64(pts64, seq.times, float(seq.times[0]) + s)
            return masks[s - 1], masks[s], bfref.wall_field_f64(p, pos, vel, seq.triangles, tris[s],
                                                                 seq.half_thickness)
        return objective_oracle(p, f0, steps, weights, geo)

    J0 = J64(THETA0)
    ref = np.zeros(8)
    for k in range(8):
        h = 1e-5
        e = np.zeros(8)
        e[k] = h
        ref[k] = (J64(THETA0 + e) - J64(THETA0 - e)) / (2 * h)
    J, grad = grads[storage]
    err = np.abs(grad - ref).max() / np.abs(ref).max()
    err32 = np.abs(grads[bh.STORAGE_FP32][1] - ref).max() / np.abs(ref).max()
    say(f"  J {name} {J:.9e}, FP32 storage {grads[bh.STORAGE_FP32][0]:.9e}, FP64 reference {J0:.9e}")
    for k, pname in enumerate(workloads.PLATE_PARAMETERS):
        say(f"  {pname:14s} adjoint {name} {grad[k]: .6e}  FP32 storage {grads[bh.STORAGE_FP32][1][k]: .6e}  "
            f"FP64 reference {ref[k]: .6e}")
    say(f"  max |adjoint - reference| / max |reference|: {name} {err:.2e}, FP32 storage {err32:.2e}; "
        f"|J - J64| / |J64| {abs(J - J0) / abs(J0):.2e}")
    bar_J, bar_g = OFFSET_BARS[storage]
    assert abs(J - J0) <= bar_J * abs(J0)
    assert err <= bar_g and err32 <= 1e-4


NO_WINNER = 0xFFFFFFFF


def ray_table_f64(p, pos64, tri, h, wall64, links32, win):
    """The exact-root link table in FP64 with the entry triangles frozen:
    each link with an entry triangle takes that triangle's shell root along
    the link, each other link the signed-distance ratio of the FP64 walls
    (the rule's fallback).  Links as the table of the unperturbed run holds
    them; an entry held at the table's floor 1e-4 (the fluid node in or at
    the shell) keeps it: its fraction does not move with theta, which the
    adjoint mirrors."""
    nx, ny, nz = p.nx, p.ny, p.nz
    n = nx * ny * nz
    L = links32.reshape(19, n)
    W = win.reshape(19, n)
    out = np.zeros((19, n), np.float64)
    q, g = np.nonzero(L != 0)
    c = bfref.C[q]
    xyz = np.stack([g % nx, (g // nx) % ny, g // (nx * ny)], 1)
    S = xyz - c
    for a, m in enumerate((nx, ny, nz)):
        for side, sel in ((0, S[:, a] < 0), (1, S[:, a] >= m)):
            bc = p.bc[a][side]
            assert bc != bfref.BC_FARFIELD or not sel.any()
            S[sel, a] = (m - 1 if side == 0 else 0) if bc == bfref.BC_PERIODIC else (0 if side == 0 else m - 1)
    src = S[:, 0] + nx * (S[:, 1] + ny * S[:, 2])
    origin = np.array([p.origin[0], p.origin[1], p.origin[2]], np.float64)
    dx = float(p.dx)
    F = origin + (S + 0.5) * dx + c * dx
    D = (-c * dx).astype(np.float64)
    hit = W[q, g] != NO_WINNER
    floor = L[q, g] <= np.float32(1e-4)
    root = hit & ~floor
    t = np.asarray(tri, np.int64)[W[q[root], g[root]]]
    r = bfref.ray_shell_f64(F[root], D[root], pos64[t[:, 0]], pos64[t[:, 1]], pos64[t[:, 2]], h)
    val = np.empty(q.size)
    val[hit & floor] = L[q, g][hit & floor]
    val[root] = np.minimum(np.maximum(r, 1e-4), 1.0)
    df = np.maximum(wall64[g, 3], 0.0)
    ds = np.minimum(wall64[src, 3], 0.0)
    sdf = np.minimum(np.maximum(df / np.maximum(df - ds, 1e-6), 1e-4), 1.0)
    val[~hit] = sdf[~hit]
    out[q, g] = val
    return out.ravel(), int(hit.sum()), int((~hit).sum())


def test_adjoint_of_exact_root_fractions_matches_fp64_reference():
    # The plate with the exact-root wall rule: the link fractions are the
    # roots of the links in the plate's shell; the adjoint differentiates
    # each through its entry triangle (dual numbers through the shell's
    # face / cylinder / sphere root).  Against central differences of the
    # FP64 oracle on the frozen link sets (masks, nearest triangles, entry
    # triangles of the unperturbed run).  Bar 1e-3.
    steps = 30
    p, weights = plate_case(steps)
    # the plate off the half-cell grid: with its edges at half-integers the
    # links through the cell centres graze its edge cylinders (distance h
    # exactly), where the root has no derivative
    plate = dict(PLATE, center=(32.0, 24.3, 24.15))
    frames = lambda theta: workloads.parametric_plate(theta, steps, **plate)
    seq, _, tan = frames(THETA0)
    cfg = dataclasses.replace(cases.solver_config(p), wall_rule=bh.WALL_INTERPOLATED_RAY)
    say(f"case: adjoint d(mean lift)/d theta, parametric plate {p.nx}x{p.ny}x{p.nz} RR3, {steps} steps, "
        f"exact-root link fractions")
    with bh.BirdFlowSimulation(cfg) as sim:
        sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
        sim.adjoint_set_tangents(tan)
        sim.init_equilibrium()
        f0 = sim.read_populations()
        J, grad, timings = sim.adjoint_gradient(steps, weights, checkpoint_interval=10)
    masks, tris, tables, wins = [], [], [], []
    for s in range(steps + 1):
        t = float(seq.times[0]) + s
        first, _, _ = bfref.frame_interval(seq.times, t)
        pos, vel = bfref.prepare_vertices(seq.points, seq.times, first, t)
        mask, wall, keys = bfref.raster_resolve(p, pos, vel, seq.triangles, seq.parts, seq.half_thickness, 64)
        links, win = bfref.ray_links(p, pos, seq.triangles, seq.half_thickness, mask, wall, winners=True)
        masks.append(mask)
        tris.append(np.where(keys != np.iinfo(np.uint64).max, (keys & 0xFFFFFFFF).astype(np.int64), -1))
        tables.append(links)
        wins.append(win)
    pl = type(p).from_buffer_copy(p)
    pl.wall_rule = bfref.WALL_INTERPOLATED_LINKS
    counts = []

    def J64(theta):
        _, pts64, _ = frames(theta)
        f = np.asarray(f0, np.float64)
        total = 0.0
        for s in range(1, steps + 1):
            pos, vel = bfref.interpolate_frames_f64(pts64, seq.times, float(seq.times[0]) + s)
            wall = bfref.wall_field_f64(p, pos, vel, seq.triangles, tris[s], seq.half_thickness)
            table, hits, falls = ray_table_f64(p, pos, seq.triangles, seq.half_thickness, wall, tables[s], wins[s])
            counts.append((hits, falls))
            f, loads, _, _ = bfref.step_with_links(pl, f, masks[s - 1], masks[s], wall, table)
            total += float(np.dot(weights[s - 1], loads[:6]))
        return total

    J0 = J64(THETA0)
    ref = np.zeros(8)
    for k in range(8):
        h = 1e-5
        e = np.zeros(8)
        e[k] = h
        ref[k] = (J64(THETA0 + e) - J64(THETA0 - e)) / (2 * h)
    err = np.abs(grad - ref).max() / np.abs(ref).max()
    hits = sum(h for h, _ in counts[:steps])
    falls = sum(f for _, f in counts[:steps])
    say(f"  J gpu {J:.9e} vs FP64 reference {J0:.9e}; links over the run: {hits} with an entry triangle, "
        f"{falls} signed-distance fallbacks")
    for k, name in enumerate(workloads.PLATE_PARAMETERS):
        say(f"  {name:14s} adjoint {grad[k]: .6e}  FP64 reference {ref[k]: .6e}")
    say(f"  max |adjoint - reference| / max |reference| = {err:.2e}; timings {timings}")
    assert hits > 1000
    assert abs(J - J0) <= 1e-4 * abs(J0)
    assert err <= 1e-3


def frozen_links(p, seq, steps):
    """Masks and nearest triangles of each step at the unperturbed
    parameters: the link sets the adjoint holds fixed."""
    masks, tris = [], []
    for s in range(steps + 1):
        t = float(seq.times[0]) + s
        first, _, _ = bfref.frame_interval(seq.times, t)
        pos, vel = bfref.prepare_vertices(seq.points, seq.times, first, t)
        mask, _, keys = bfref.raster_resolve(p, pos, vel, seq.triangles, seq.parts, seq.half_thickness, 64)
        masks.append(mask)
        tris.append(np.where(keys != np.iinfo(np.uint64).max, (keys & 0xFFFFFFFF).astype(np.int64), -1))
    return masks, tris


def test_adjoint_of_the_inlet_outlet_plate_twin_matches_fp64_reference():
    """The differentiation workload's plate (inflow u_x = 0.05 through a
    far-field inlet at -x, zero-gradient outlet at +x, periodic y / z, RR3,
    tau+ 0.6) on its 64 x 32 x 32 twin: plate 12 x 4 x 2 cells at the
    quarter-scale position, period 200 steps, 8 kinematic parameters drawn
    +-20 % from the workload's values with seed 7 (heave scaled to the twin),
    J = lift averaged over the second of two periods (one full flapping
    period), checkpoints every 100 steps.  dJ/dtheta against central
    differences (h = 1e-5) of the FP64 oracle objective on the frozen link
    sets.  Bar: 1e-3 relative."""
    steps, period = 400, 200.0
    shape = (64, 32, 32)
    p = cases.plate_params(shape=shape, collision=bfref.COLLISION_RR3, omega_plus=1.0 / 0.6, inflow=0.05)
    p.body_position[0], p.body_position[1], p.body_position[2] = 16.0, 16.0, 16.0
    base = np.array([2.0, 0.0, 0.5, 0.0, np.radians(30.0), np.pi / 2, np.radians(7.5), np.pi / 2])
    theta0 = base * (1.0 + 0.2 * np.random.default_rng(7).uniform(-1.0, 1.0, 8))
    plate = dict(center=(16.0, 16.0, 16.0), span=12.0, chord=4.0, thickness=2.0, period=period, half_thickness=0.5)
    weights = np.zeros((steps, 6))
    weights[steps // 2:, 2] = 1.0 / (steps - steps // 2)
    seq, _, tan = workloads.parametric_plate(theta0, steps, **plate)
    say(f"case: adjoint of mean lift, inlet / outlet plate twin {shape}, RR3, {steps} steps (window {steps // 2}), "
        f"interval 100, theta {np.round(theta0, 4).tolist()}")
    with bh.BirdFlowSimulation(cases.solver_config(p)) as sim:
        sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
        sim.adjoint_set_tangents(tan)
        sim.init_equilibrium()
        f0 = sim.read_populations()
        J, grad, timings = sim.adjoint_gradient(steps, weights, checkpoint_interval=100)
    masks, tris = frozen_links(p, seq, steps)

    def J64(theta):
        _, pts64, _ = workloads.parametric_plate(theta, steps, **plate)

        def geo(s):
            pos, vel = bfref.interpolate_frames_f64(pts64, seq.times, float(seq.times[0]) + s)
            return masks[s - 1], masks[s], bfref.wall_field_f64(p, pos, vel, seq.triangles, tris[s],
                                                                 seq.half_thickness)
        return objective_oracle(p, f0, steps, weights, geo)

    # h = 1e-5: over 400 steps some link fractions of the frozen links reach
    # their clamps (1e-4, 1) within +-1e-3 of theta0, so wider differences
    # straddle kinks (at h = 1e-3 the heave and pitch amplitude derivatives
    # are off by 2x and 6x; at 1e-4 and 1e-5 they agree to 3e-4)
    J0 = J64(theta0)
    ref = np.zeros(8)
    for k in range(8):
        h = 1e-5
        e = np.zeros(8)
        e[k] = h
        ref[k] = (J64(theta0 + e) - J64(theta0 - e)) / (2 * h)
    err = np.abs(grad - ref).max() / np.abs(ref).max()
    say(f"  J gpu {J:.9e} vs FP64 reference {J0:.9e}")
    for k, name in enumerate(workloads.PLATE_PARAMETERS):
        say(f"  {name:14s} adjoint {grad[k]: .6e}  FP64 reference {ref[k]: .6e}")
    say(f"  max |adjoint - reference| / max |reference| = {err:.2e}; forward {timings['forward_ms']:.1f} ms, "
        f"reverse {timings['reverse_ms']:.1f} ms")
    assert abs(J - J0) <= 1e-4 * abs(J0)
    assert err <= 1e-3


def test_adjoint_gradient_against_solver_finite_differences():
    # The solver's own FP32 central differences from the same start state,
    # with the masks of the unperturbed run replayed (the adjoint holds each
    # step's link set fixed), at two step sizes h and h/4.  With fixed masks
    # the objective is still only piecewise smooth (interpolation fractions of
    # the replayed links clamp as the wall moves), and FP32 rounding limits
    # the small step: the two differences bracket the derivative to about
    # 1 %.  Bar: the adjoint lies within the interval the two differences span,
    # widened by half its width plus 1e-3 of the largest component, and within
    # 1e-2 relative L2 of the h/4 differences.  Differences with re-rastered
    # masks are printed beside (cover / uncover events dominate them).
    steps = 30
    p, weights = plate_case(steps)
    cfg = cases.solver_config(p)
    base_masks, f0 = [], None
    seq, _, tan = plate_frames(THETA0, steps)
    with bh.BirdFlowSimulation(cfg) as sim:
        sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
        sim.init_equilibrium()
        f0 = sim.read_populations()
        base_masks.append(sim.read_geometry()[0])
        for _ in range(steps):
            sim.advance(1)
            base_masks.append(sim.read_geometry()[0])
    with bh.BirdFlowSimulation(cfg) as sim:
        sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
        sim.adjoint_set_tangents(tan)
        sim.write_populations(f0)
        J, grad, timings = sim.adjoint_gradient(steps, weights, checkpoint_interval=15)

    def J_gpu(theta, frozen):
        sq, _, _ = plate_frames(theta, steps)
        with bh.BirdFlowSimulation(cfg) as sim:
            sim.upload_mesh(sq.points, sq.times, sq.triangles, sq.parts, sq.half_thickness)
            if frozen:
                sim.freeze_masks(np.array(base_masks))
            sim.write_populations(f0)
            sim.advance(steps)
            hist = sim.load_history()
        return float(np.sum(weights * hist[:, :6]))

    fd, fd_small, fd_free = np.zeros(8), np.zeros(8), np.zeros(8)
    for k in range(8):
        h = 4e-3 * max(1.0, abs(THETA0[k]))
        e = np.zeros(8)
        e[k] = 1.0
        fd[k] = (J_gpu(THETA0 + h * e, True) - J_gpu(THETA0 - h * e, True)) / (2 * h)
        fd_small[k] = (J_gpu(THETA0 + h / 4 * e, True) - J_gpu(THETA0 - h / 4 * e, True)) / (h / 2)
        fd_free[k] = (J_gpu(THETA0 + h / 4 * e, False) - J_gpu(THETA0 - h / 4 * e, False)) / (h / 2)
    lo, hi = np.minimum(fd, fd_small), np.maximum(fd, fd_small)
    outside = np.maximum(0.0, np.maximum(lo - grad, grad - hi))
    allowed = 1e-3 * np.abs(fd).max() + 0.5 * (hi - lo)
    for k, name in enumerate(workloads.PLATE_PARAMETERS):
        say(f"  {name:14s} adjoint {grad[k]: .6e}  FP32 central differences (fixed links) h {fd[k]: .6e}, "
            f"h/4 {fd_small[k]: .6e}  outside {outside[k]:.2e} (allowed {allowed[k]:.2e});  "
            f"re-rastered masks, h/4: {fd_free[k]: .6e}")
    l2 = np.linalg.norm(grad - fd_small) / np.linalg.norm(fd_small)
    say(f"  relative L2 against h/4 differences {l2:.2e}; adjoint sweep: forward {timings['forward_ms']:.1f} ms, "
        f"reverse {timings['reverse_ms']:.1f} ms")
    assert np.all(outside <= allowed)
    assert l2 <= 1e-2


def test_torch_autograd_returns_adjoint_gradient():
    torch = pytest.importorskip("torch")
    from birdflow_hip.autograd import LoadFunctional, load_functional
    steps = 12
    p, weights = plate_case(steps)

    def frames(theta):
        seq, _, tan = plate_frames(theta, steps)
        return seq, tan

    fn = LoadFunctional(cases.solver_config(p), frames, weights, checkpoint_interval=6)
    theta = torch.tensor(THETA0, dtype=torch.float64, requires_grad=True)
    J = load_functional(theta, fn)
    J.backward()
    say(f"case: torch autograd, J {J.item():.6e}, grad {theta.grad.numpy()}")
    assert np.allclose(theta.grad.numpy(), fn.last["gradient"], rtol=0, atol=0)
    assert J.item() == fn.last["objective"]


def test_torch_autograd_over_fourier_coefficients():
    torch = pytest.importorskip("torch")
    from birdflow_hip.autograd import KinematicsFunctional, load_functional
    steps = 8
    p, mesh = articulated_bird_case(None)
    weights = np.zeros((steps, 6))
    weights[:, 2] = 1.0 / steps
    fn = KinematicsFunctional(cases.solver_config(p), mesh, BIRD_PARAMETERS, weights, checkpoint_interval=4)
    theta = torch.zeros(len(BIRD_PARAMETERS), dtype=torch.float64, requires_grad=True)
    J = load_functional(theta, fn)
    (2.0 * J).backward()
    say(f"case: torch autograd over Fourier coefficients, J {J.item():.6e}, grad {theta.grad.numpy()}")
    assert np.array_equal(theta.grad.numpy(), 2.0 * fn.last["gradient"])
    assert np.abs(fn.last["gradient"]).max() > 0.0


BIRD_PARAMETERS = [bh.KinematicParameter((0, 1), "stroke", 0, "sin"), bh.KinematicParameter((0, 1), "stroke", 0, "cos"),
                   bh.KinematicParameter((0, 1), "pitch", 0, "sin"), bh.KinematicParameter((0,), "pitch", 0, "cos"),
                   bh.KinematicParameter((0, 1), "twist", 1, "sin"), bh.KinematicParameter((1,), "deviation", 0, "cos")]


def articulated_bird_case(pose):
    """The synthetic bird at 96 / 512 scale (1,168 triangles), Mach-limited
    kinematics, RR3 at tau+ 0.6, far-field inlet and outlet on x, periodic
    y / z.  pose: None (mesh centred in the world) or a body state (mesh in
    the body frame, placed by it)."""
    shape = (96, 48, 48)
    s = shape[0] / 512.0
    center = (30.0, 24.0, 24.0)
    bird = workloads.SyntheticBird(seed=7, center=(0.0, 0.0, 0.0) if pose is not None else center,
                                   body_radii=(40 * s, 14 * s, 14 * s), half_span=100 * s, root_chord=40 * s,
                                   tip_chord=12 * s, tail=(30 * s, 20 * s), period=1200 * s,
                                   body_segments=(24, 12), wing_segments=(24, 6), tail_segments=(8, 4))
    bird.limit_mach(0.04, 0.15)
    tau = 0.6
    p = bfref.make_params(shape, omega_plus=1.0 / tau, omega_minus=1.0 / (0.5 + (3.0 / 16.0) / (tau - 0.5)),
                          bc=((bfref.BC_FARFIELD, bfref.BC_OUTLET), (bfref.BC_PERIODIC, bfref.BC_PERIODIC),
                              (bfref.BC_PERIODIC, bfref.BC_PERIODIC)),
                          collision=bfref.COLLISION_RR3, wall_rule=bfref.WALL_INTERPOLATED_SDF, force_mode=6,
                          accumulate_loads=True, far_u=(0.04, 0.0, 0.0), body_position=center)
    return p, bird.articulated(half_thickness=1.0)


def place_f64(pose, pos, vel):
    """The body pose placing body-frame vertices (place_vertex) in FP64."""
    if pose is None:
        return pos, vel
    x, y, z, w = np.asarray(pose[3:7], np.float64) / np.linalg.norm(pose[3:7])
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    om = np.asarray(pose[10:13], np.float64)
    return pose[:3] + pos @ R.T, pose[7:10] + np.cross(om, pos) @ R.T + vel @ R.T


@pytest.mark.parametrize("posed", [False, True], ids=["world", "body pose"])
def test_adjoint_of_articulated_bird_matches_fp64_reference(posed):
    steps, interval = 40, 20
    pose = None
    if posed:
        # 12 degrees about (1, 2, 3), spinning slowly; free flight off, so the
        # pose stays put and only places the mesh
        ax = np.array([1.0, 2.0, 3.0]) / np.sqrt(14.0)
        half = np.radians(12.0) / 2
        pose = np.array([30.0, 24.0, 24.0, *(ax * np.sin(half)), np.cos(half), 0.001, 0.0, 0.0, 0.002, -0.001, 0.0015])
    p, mesh = articulated_bird_case(pose)
    weights = np.zeros((steps, 6))
    weights[steps // 2:, 2] = 1.0 / (steps - steps // 2)
    say(f"case: adjoint of mean lift, articulated bird {p.nx}x{p.ny}x{p.nz}, {len(mesh.triangles)} triangles, RR3, "
        f"{steps} steps, 6 Fourier coefficients, {'body pose' if posed else 'world'}")
    cfg = cases.solver_config(p)
    with bh.BirdFlowSimulation(cfg) as sim:
        if posed:
            sim.set_body(pose, bh.BodyParameters(mass=1e6, inertia=(1e7, 1e7, 1e7), free_flight=False,
                                                 move_mesh=True))
        sim.upload_articulated(mesh, 0.0)
        sim.adjoint_set_kinematic_parameters(BIRD_PARAMETERS)
        sim.init_equilibrium()
        f0 = sim.read_populations()
        J, grad, timings = sim.adjoint_gradient(steps, weights, checkpoint_interval=interval)
    # frozen link sets from the FP32 oracle geometry (the GPU's, bit for bit)
    k32 = cases.oracle_kinematics(mesh)
    body = None if pose is None else bfref.make_body(position=pose[:3], orientation=pose[3:7],
                                                     linear_velocity=pose[7:10], angular_velocity_body=pose[10:13])
    masks, tris = [], []
    for s in range(steps + 1):
        pos, vel = bfref.articulated_vertices(mesh.rest, mesh.vertex_joint, mesh.span_fraction, k32, float(s))
        if body is not None:
            pos, vel = bfref.place_vertices(body, pos, vel)
        mask, _, keys = bfref.raster_resolve(p, pos, vel, mesh.triangles, mesh.parts, mesh.half_thickness, 64)
        masks.append(mask)
        tris.append(np.where(keys != np.iinfo(np.uint64).max, (keys & 0xFFFFFFFF).astype(np.int64), -1))
    theta0 = np.zeros(len(BIRD_PARAMETERS))       # offsets of the coefficients

    def J64(theta):
        k64 = cases.kinematics_f64(mesh.perturbed(BIRD_PARAMETERS, theta))

        def geo(s):
            pos, vel = bfref.articulated_vertices_f64(mesh.rest, mesh.vertex_joint, mesh.span_fraction, k64, float(s))
            pos, vel = place_f64(pose, pos, vel)
            return masks[s - 1], masks[s], bfref.wall_field_f64(p, pos, vel, mesh.triangles, tris[s],
                                                                 mesh.half_thickness)
        return objective_oracle(p, f0, steps, weights, geo)

    J0 = J64(theta0)
    ref = np.zeros(len(theta0))
    for k in range(len(theta0)):
        h = 1e-5
        e = np.zeros(len(theta0))
        e[k] = h
        ref[k] = (J64(theta0 + e) - J64(theta0 - e)) / (2 * h)
    err = np.abs(grad - ref).max() / np.abs(ref).max()
    say(f"  J gpu {J:.9e} vs FP64 reference {J0:.9e}")
    for k, q in enumerate(BIRD_PARAMETERS):
        say(f"  {q.channel:9s} h{q.harmonic} {q.term} joints {q.joints}: adjoint {grad[k]: .6e}  "
            f"FP64 reference {ref[k]: .6e}")
    say(f"  max |adjoint - reference| / max |reference| = {err:.2e}; forward {timings['forward_ms']:.1f} ms, "
        f"reverse {timings['reverse_ms']:.1f} ms")
    assert abs(J - J0) <= 1e-4 * abs(J0)
    assert err <= 1e-3


def test_kinematic_parameter_refusals():
    p, mesh = articulated_bird_case(None)
    seq, _, _ = plate_frames(THETA0, 4)
    with bh.BirdFlowSimulation(cases.solver_config(p)) as sim:
        sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
        with pytest.raises(bh.BirdFlowError):          # frames, not an articulated mesh
            sim.adjoint_set_kinematic_parameters(BIRD_PARAMETERS[:1])
        sim.upload_articulated(mesh, 0.0)
        for bad in (bh.KinematicParameter((2,), "stroke", 0, "sin"),      # the bird has joints 0 and 1
                    bh.KinematicParameter((), "stroke", 0, "sin"),
                    bh.KinematicParameter((0,), "stroke", 3, "sin")):     # 3 harmonics
            with pytest.raises(bh.BirdFlowError):
                sim.adjoint_set_kinematic_parameters([bad])
        with pytest.raises(bh.BirdFlowError):          # frame tangents need frames
            sim.adjoint_set_tangents(np.zeros((1, 2, len(mesh.rest), 3), np.float32))
        sim.adjoint_set_kinematic_parameters(BIRD_PARAMETERS)
        sim.upload_articulated(mesh, 0.0)              # a new upload clears them
        sim.init_equilibrium()
        J, grad, _ = sim.adjoint_gradient(4, np.ones((4, 6)), checkpoint_interval=2)
        assert grad.size == 0 and np.isfinite(J)
