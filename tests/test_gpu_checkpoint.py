"""GPU checks of restart archives, after the reference's checkpoint test
(Tests/BirdFlowMetalTests/ObservationAndCheckpointTests.swift:115-178):

* a context restored from an archive and saved again writes byte-identical
  files;
* the restored context continues exactly as the saved one: stored
  populations, loads, masks and body state bit-identical after more steps
  (the reference allows 1e-6 / 1e-5; here the restart is exact, for FP32,
  FP16 and BF16 storage, a flapping-plate mesh, an articulated mesh with
  on-device kinematics, a translating sphere, a
  free-flying body carrying its mesh, a free-flying analytic bird and
  the prescribed flapping wing with its link table);
* a corrupted populations file is refused with CheckpointError.
"""
import os
import dataclasses
import sys

import numpy as np
import pytest

import bfref
import birdflow_hip as bh
import cases
from birdflow_hip import checkpoint, workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def plate_sequence(steps):
    return workloads.flapping_plate(steps, center=(32.0, 24.0, 24.0), span=20.0, chord=8.0, thickness=2.0,
                                    heave=2.0, period=40.0, half_thickness=0.5)


def files_of(directory):
    out = {}
    for name in sorted(os.listdir(directory)):
        with open(os.path.join(directory, name), "rb") as fh:
            out[name] = fh.read()
    return out


def check_restart(sim, tmp_path, before, after):
    """Advance `before` steps, save, restore, save again, advance both
    `after` steps, and compare."""
    sim.advance(before)
    a = str(tmp_path / "a")
    manifest = checkpoint.save_checkpoint(sim, a)
    restored = checkpoint.load_checkpoint(a)
    with restored:
        b = str(tmp_path / "b")
        checkpoint.save_checkpoint(restored, b)
        fa, fb = files_of(a), files_of(b)
        say(f"  saved at step {manifest['step']}: files {sorted(fa)}; "
            f"{sum(len(v) for v in fa.values())} bytes")
        assert sorted(fa) == sorted(fb)
        for name in fa:
            assert fa[name] == fb[name], f"re-saved {name} differs"
        assert restored.step_index == sim.step_index == before
        sim.advance(after)
        restored.advance(after)
        pa, pb = sim.read_population_storage(), restored.read_population_storage()
        la, lb = sim.load_history()[-after:], restored.load_history()
        ma, mb = sim.read_geometry()[0], restored.read_geometry()[0]
        ba, bb = sim.body_state(), restored.body_state()
        say(f"  after {after} more steps: populations differing {int((pa.view(np.uint8) != pb.view(np.uint8)).sum())} B, "
            f"load records equal {np.array_equal(la, lb)}, masks equal {np.array_equal(ma, mb)}, "
            f"body max |diff| {np.abs(ba - bb).max():.3e}")
        assert lb.shape[0] == after
        assert np.all(np.isfinite(la)) and np.all(np.isfinite(ba))
        assert np.array_equal(pa.view(np.uint8), pb.view(np.uint8))
        assert np.array_equal(la, lb)
        assert np.array_equal(ma, mb)
        assert np.array_equal(ba, bb)
        assert restored.restart_state() == sim.restart_state()


@pytest.mark.parametrize("storage", [bh.STORAGE_FP32, bh.STORAGE_FP16_OFFSET, bh.STORAGE_BF16_OFFSET])
def test_restart_of_flapping_plate_is_exact(storage, tmp_path):
    say(f"case: checkpoint of a flapping plate 96x48x48, storage {storage}, 5 + 4 steps")
    p = cases.plate_params()
    seq = plate_sequence(12)
    with bh.BirdFlowSimulation(cases.solver_config(p, storage=storage)) as sim:
        sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
        sim.init_equilibrium()
        check_restart(sim, tmp_path, 5, 4)


def test_restart_under_exact_root_rule_is_exact(tmp_path):
    # the exact-root wall rule is part of the saved configuration; its
    # tables are rebuilt from the geometry at every step
    say("case: checkpoint of a flapping plate 96x48x48, exact-root link fractions, 5 + 4 steps")
    p = cases.plate_params()
    seq = plate_sequence(12)
    cfg = dataclasses.replace(cases.solver_config(p), wall_rule=bh.WALL_INTERPOLATED_RAY)
    with bh.BirdFlowSimulation(cfg) as sim:
        sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
        sim.init_equilibrium()
        check_restart(sim, tmp_path, 5, 4)


def test_restart_of_articulated_bird_is_exact(tmp_path):
    say("case: checkpoint of the articulated synthetic bird (on-device kinematics) 128x64x64, 5 + 4 steps")
    p = cases.plate_params(shape=(128, 64, 64))
    mesh = workloads.synthetic_bird_articulated(seed=7, center=(40.0, 32.0, 32.0), body_radii=(10.0, 3.5, 3.5),
                                                half_span=25.0, root_chord=10.0, tip_chord=3.0, tail=(7.5, 5.0),
                                                period=300.0, body_segments=(24, 16), wing_segments=(40, 10),
                                                tail_segments=(12, 6))
    with bh.BirdFlowSimulation(cases.solver_config(p)) as sim:
        sim.upload_articulated(mesh)
        sim.init_equilibrium()
        check_restart(sim, tmp_path, 5, 4)


def test_restart_of_translating_sphere_is_exact(tmp_path):
    say("case: checkpoint of the translating sphere 24^3, 6 + 5 steps")
    p = cases.translating_sphere_params()
    with bh.BirdFlowSimulation(cases.solver_config(p)) as sim:
        sim.set_translating_sphere(cases.SPHERE_CENTER_RADIUS, cases.SPHERE_GEOMETRY_VELOCITY,
                                   cases.SPHERE_WALL_VELOCITY)
        sim.init_equilibrium()
        check_restart(sim, tmp_path, 6, 5)


def test_restart_of_free_flying_body_carrying_its_mesh_is_exact(tmp_path):
    # The mesh follows the body pose: the mask the next step starts from was
    # built at the pose before the last body update, so the archive's mask,
    # not a rebuilt one, is the restart's previous mask.
    say("case: checkpoint of a free-flying plate carried by its body, 96x48x48, 6 + 4 steps")
    p = cases.plate_params()
    seq = workloads.flapping_plate(12, center=(0.0, 0.0, 0.0), span=20.0, chord=8.0, thickness=2.0,
                                   heave=2.0, period=40.0, half_thickness=0.5)
    state0 = np.array([24.0, 24.0, 24.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0])
    with bh.BirdFlowSimulation(cases.solver_config(p)) as sim:
        sim.set_body(state0, bh.BodyParameters(mass=2.0e4, inertia=(2.0e5, 1.5e5, 2.5e5), gravity=(0.0, 0.0, -2e-4),
                                               substeps=2, move_mesh=True))
        sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
        sim.init_equilibrium()
        check_restart(sim, tmp_path, 6, 4)
        moved = np.abs(sim.body_state()[:3] - state0[:3]).max()
        say(f"  body moved {moved:.3e} cells")
        assert moved > 0.0


def test_restart_of_free_flying_analytic_bird_is_exact(tmp_path):
    say("case: checkpoint of the free-flying compact analytic bird 48^3 on measured keyframes, 5 + 4 steps")
    p, cfg, bird, _, _ = cases.compact_bird_case(free_flight=True)
    cfg.time_step = 0.002
    state0 = np.array(list(p.body_position) + [0.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0])
    with bh.BirdFlowSimulation(cfg) as sim:
        sim.set_body(state0, bh.BodyParameters(mass=0.025, inertia=(4e-6, 7e-6, 8e-6), gravity=(0.0, 0.0, -9.81),
                                               substeps=2))
        sim.set_bird(bird, keyframes=cases.compact_bird_keyframes())
        sim.init_equilibrium()
        check_restart(sim, tmp_path, 5, 4)


def test_restart_of_prescribed_wing_is_exact(tmp_path):
    say("case: checkpoint of the Li & Nabawy wing 64x64x32, link rule, 5 + 4 steps")
    _, cfg, wing, _ = cases.wing_case()
    with bh.BirdFlowSimulation(cfg) as sim:
        sim.set_flapping_wing(wing)
        sim.init_equilibrium()
        check_restart(sim, tmp_path, 5, 4)


def test_corrupted_checkpoint_is_refused(tmp_path):
    say("case: corrupted checkpoint populations file")
    p = cases.plate_params(shape=(64, 32, 32))
    seq = plate_sequence(4)
    a = str(tmp_path / "a")
    with bh.BirdFlowSimulation(cases.solver_config(p)) as sim:
        sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
        sim.init_equilibrium()
        sim.advance(2)
        checkpoint.save_checkpoint(sim, a, compression="none")
    path = os.path.join(a, checkpoint.POPULATIONS_NAME)
    with open(path, "r+b") as fh:
        fh.seek(1000)
        b = fh.read(1)
        fh.seek(1000)
        fh.write(bytes([b[0] ^ 0x10]))
    with pytest.raises(checkpoint.CheckpointError, match="checksum"):
        checkpoint.load_checkpoint(a)


def test_restored_archive_steps_like_the_oracle(tmp_path):
    # An archive restored into a fresh context, stepped on, against the FP32
    # oracle stepped from the archive's own contents: the stored populations
    # and the mask it saved, the mesh frames at the restored clock.
    # Strict FP32, RR3, inlet / outlet plate: populations bit-identical and
    # loads to 1e-12 after 4 more steps.
    say("case: plate 96x48x48 saved at step 5, restored, 4 steps against the oracle from the archive")
    p = cases.plate_params()
    seq = plate_sequence(12)
    a = str(tmp_path / "a")
    with bh.BirdFlowSimulation(cases.solver_config(p)) as sim:
        sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
        sim.init_equilibrium()
        sim.advance(5)
        checkpoint.save_checkpoint(sim, a)
    ck = checkpoint.read_checkpoint(a)
    with checkpoint.load_checkpoint(a) as restored:
        f = restored.read_populations()
        restored.advance(4)
        g = restored.read_populations()
        loads = restored.load_history()
    step0 = ck.manifest["step"]
    origin = ck.manifest["mesh_time_origin"]
    prev = ck.solid_mask
    loads_o = []
    for s in range(step0 + 1, step0 + 5):
        t = float(np.float32(origin) + np.float32(float(s) * 1.0))
        mask, wall = cases.mesh_geometry_oracle(p, seq, t)
        f, ld, _, _ = bfref.step(p, f, prev, mask, wall)
        loads_o.append(ld)
        prev = mask
    loads_o = np.array(loads_o)
    differing = int((g.view(np.uint32) != f.view(np.uint32)).sum())
    rel = float(np.abs(loads[:, :6] - loads_o[:, :6]).max() / np.abs(loads_o[:, :6]).max())
    say(f"  restored at step {step0}: populations differing {differing}, loads max rel err {rel:.2e}")
    assert differing == 0 and rel <= 1e-12
