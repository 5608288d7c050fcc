"""GPU parity of the six-degree-of-freedom body (FP64 on the device) against
the FP64 oracle integrator, the safety monitor, and free flight with the mesh
carried by the body pose."""
import math
import sys

import numpy as np
import pytest

import bfref
import birdflow_hip as bh
import cases
from birdflow_hip import workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _empty_box_sim(shape=(64, 16, 16), time_step=1.0):
    p = bfref.make_params(shape, omega_plus=1.0 / 0.8, omega_minus=1.0, accumulate_loads=True,
                          body_position=(32.0, 8.0, 8.0))
    return p, bh.BirdFlowSimulation(cases.solver_config(p, time_step=time_step))


def test_free_body_without_fluid_load_matches_oracle_integrator():
    # no walls: the fluid load is exactly zero and the body follows gravity,
    # the external force and torque; FP64 on both sides, same order
    say("case: free body, 40 steps x 3 substeps, zero fluid load")
    params = dict(mass=2.5, inertia=(1.0, 2.0, 3.0), gravity=(0.0, -0.01, 0.002), external_force=(0.3, 0.0, -0.1),
                  external_torque=(0.01, -0.02, 0.05), substeps=3)
    state0 = dict(position=(30.0, 8.0, 8.0), orientation=(0.1, 0.2, 0.3, 0.9), linear_velocity=(0.01, 0.0, 0.0),
                  angular_velocity_body=(0.2, -0.1, 0.4))
    p, sim = _empty_box_sim(time_step=0.5)
    with sim:
        sim.init_equilibrium()
        sim.set_body(np.concatenate([state0[k] for k in ("position", "orientation", "linear_velocity",
                                                         "angular_velocity_body")]),
                     bh.BodyParameters(mass=params["mass"], inertia=params["inertia"], gravity=params["gravity"],
                                       external_force=params["external_force"],
                                       external_torque=params["external_torque"], substeps=3))
        sim.advance(40)
        hist = sim.body_history()
        loads = sim.load_history()
    assert hist.shape == (40, 13) and np.all(loads[:, :6] == 0.0)
    b = bfref.make_body(**state0)
    ref = []
    for _ in range(40):
        cases.integrate_body_oracle(b, np.zeros(8), params, 0.5)
        ref.append(bfref.body_array(b))
    ref = np.array(ref)
    err = np.abs(hist - ref).max()
    say(f"  max |gpu - oracle| over the history {err:.3e}; bitwise {np.array_equal(hist, ref)}")
    assert err <= 1e-13


def test_torque_free_spin_follows_closed_form_on_device():
    say("case: torque-free spin about a principal axis, 500 steps")
    p, sim = _empty_box_sim()
    with sim:
        sim.init_equilibrium()
        sim.set_body(np.array([32, 8, 8, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0.5]),
                     bh.BodyParameters(mass=1.0, inertia=(1.0, 2.0, 3.0), substeps=2))
        sim.advance(500)
        s = sim.body_state()
    angle = 2 * math.atan2(s[5], s[6])
    exact = 1000 * 2 * math.atan(0.5 * 0.5 / 2)
    exact = math.remainder(exact, 2 * math.pi)
    say(f"  angle {angle:.15f} vs {exact:.15f}")
    assert abs(math.remainder(angle - exact, 2 * math.pi)) < 1e-11
    assert s[12] == 0.5 and s[10] == 0.0 and s[11] == 0.0


def test_safety_monitor_reports_clearance_event_once():
    say("case: body drifting into the padded domain boundary")
    p, sim = _empty_box_sim(shape=(64, 32, 32))
    params = dict(mass=1.0, inertia=(1.0, 1.0, 1.0), gravity=(0.0, 0.0, 0.0), external_force=(0.0, 0.0, 0.0),
                  external_torque=(0.0, 0.0, 0.0), substeps=1)
    with sim:
        sim.init_equilibrium()
        state = np.array([40.0, 16, 16, 0, 0, 0, 1, 1.0, 0, 0, 0, 0, 0])
        sim.set_body(state, bh.BodyParameters(mass=1.0, inertia=(1, 1, 1), safety_extent=(2, 1, 1),
                                              safety_radius=1.0, mach_limit=5.0))
        raised = []
        for s in range(1, 31):
            try:
                sim.advance(1)
            except bh.BirdFlowError as e:
                raised.append((s, e.status))
        metrics, step, flags = sim.safety()
    # oracle: x domain [0 + 7, 64 - 7] (sponge width 4 + 3 cells)
    b = bfref.make_body(position=(40, 16, 16), linear_velocity=(1.0, 0, 0))
    rec = bfref.Safety()
    for s in range(1, 31):
        cases.integrate_body_oracle(b, np.zeros(8), params, 1.0)
        bfref.body_safety(b, rec, s, (2, 1, 1), 1.0, 0.0, 5.0, (7, 7, 7), (57, 25, 25), (0, 0, 0), 1.0)
    say(f"  gpu event step {step} flags {flags} metrics {metrics}; raised {raised}; oracle {rec.event} "
        f"{rec.metrics}")
    assert int(rec.event[2]) == 2 and int(rec.event[0]) == 16
    assert raised == [(16, bh._abi.ERROR_RUNTIME_SAFETY_VIOLATION)]
    assert flags == 2 and step == 16
    assert np.array_equal(metrics, rec.metrics)


def test_free_flight_plate_carried_by_body_matches_oracle():
    steps = 30
    p = cases.plate_params()
    # plate frames in the body frame (centred on the origin)
    seq = workloads.flapping_plate(steps, center=(0.0, 0.0, 0.0), span=20.0, chord=8.0, thickness=2.0,
                                   heave=2.0, period=200.0, half_thickness=0.5)
    params = dict(mass=2.0e4, inertia=(2.0e5, 1.5e5, 2.5e5), gravity=(0.0, 0.0, -2e-4),
                  external_force=(0.0, 0.0, 0.0), external_torque=(0.0, 0.0, 0.0), substeps=2)
    state0 = np.array([24.0, 24.0, 24.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0])
    say(f"case: free-flying flapping plate carried by the body, {steps} steps, strict arithmetic")
    with bh.BirdFlowSimulation(cases.solver_config(p)) as sim:
        sim.set_body(state0, bh.BodyParameters(mass=params["mass"], inertia=params["inertia"],
                                               gravity=params["gravity"], substeps=2, move_mesh=True))
        sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
        sim.init_equilibrium()
        sim.advance(steps)
        hist = sim.body_history()
        loads = sim.load_history()
        g = sim.read_populations()
        mask, _ = sim.read_geometry()
    body = bfref.make_body(position=state0[:3])
    f_o, loads_o, states_o, mask_o = cases.run_coupled_body_oracle(p, seq, steps, body, params)
    ferr = np.abs(loads[:, :3] - loads_o[:, :3]).max() / np.abs(loads_o[:, :3]).max()
    berr = np.abs(hist - states_o).max()
    moved = np.abs(states_o[-1, :3] - state0[:3]).max()
    say(f"  body moved {moved:.4f} cells; body history max |diff| {berr:.3e}; force history rel err {ferr:.3e}; "
        f"mask cells differing {int((mask != mask_o).sum())}; populations rel L2 "
        f"{np.linalg.norm(g - f_o) / np.linalg.norm(f_o):.3e}")
    assert moved > 1e-3
    assert np.array_equal(mask, mask_o)
    assert ferr <= 1e-6 and berr <= 1e-9
