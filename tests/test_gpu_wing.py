"""GPU parity of the prescribed beta-planform flapping wing
(BirdFlow.metal:940-1297, driven by MetalFlappingWingValidation.swift) and
of the link-table wall rule (BirdFlow.metal:2738-2745) against the FP32
oracle, on the Li & Nabawy planform (chord 8 cells, span 24 cells, cycle
200 time units, 2 time units per step, so the tip moves about 1.6 cells a
step and cells change sides every step):

* every step's prepared frame and mask are bit-identical to the oracle's,
  and so is the wall wherever a step reads it (cells on the wing in this
  step or the previous one; the device evaluates only the tiles near the
  planform, the oracle its whole guard sphere);
* the link table holds the oracle's bisected fraction, bit for bit, for
  every fluid cell whose source is on the wing;
* fluid steps with the link rule: populations bit-identical, loads to 1e-12
  relative (FP64 sums of the same FP32 terms in another order);
* the fused-FP32 and FP16 / BF16 offset forms of the link rule read the
  same table and track strict FP32 (populations 2e-4, force history 2e-3
  of the largest force; measured 4e-5 and 2e-4 for BF16);
* physics: the Li & Nabawy fifth-cycle mean lift coefficient within 10 %
  of the published 1.460 at 16 cells per chord;
* the link rule without the wing, and the adjoint of a wing, are refused.
"""
import sys

import numpy as np
import pytest

import bfref
import birdflow_hip as bh
import cases

pytestmark = pytest.mark.gpu

SHAPE, ROOT, CYCLE, TIME_STEP, VELOCITY_TO_LATTICE = (cases.WING_SHAPE, cases.WING_ROOT, cases.WING_CYCLE,
                                                      cases.WING_TIME_STEP, cases.WING_VELOCITY_TO_LATTICE)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def geometry_time(s):
    return np.float32(np.float64(s) * np.float64(np.float32(TIME_STEP)))


def oracle_geometry(p, ow, s, prev):
    fr = bfref.wing_prepare(ow, geometry_time(s))
    m, w = bfref.wing_build(p, ow, fr, prev, VELOCITY_TO_LATTICE)
    return fr, m, w


def source_is_solid(mask, q, periodic=False):
    """Fluid cells whose source cell - c_q is solid, [N] bool: in-domain
    sources, or wrapped across the faces when periodic."""
    nx, ny, nz = SHAPE
    m3 = mask.reshape(nz, ny, nx)
    cx, cy, cz = bfref.C[q]
    if periodic:
        src = np.roll(m3, (cz, cy, cx), axis=(0, 1, 2))
    else:
        src = np.zeros_like(m3)
        src[max(cz, 0):nz + min(cz, 0), max(cy, 0):ny + min(cy, 0), max(cx, 0):nx + min(cx, 0)] = \
            m3[max(-cz, 0):nz + min(-cz, 0), max(-cy, 0):ny + min(-cy, 0), max(-cx, 0):nx + min(-cx, 0)]
    return (src.ravel() != 0) & (mask == 0)


# the wing rooted 3 cells inside the +x face with every face periodic: at
# the stroke reversal its tip reaches the last x plane, so cells of the first
# plane pull from wing cells across the face
CROSSING_ROOT = (61.0, 32.0, 16.0)


@pytest.mark.parametrize("root,periodic", [(ROOT, False), (CROSSING_ROOT, True)])
def test_wing_geometry_and_links_match_oracle(root, periodic):
    p, cfg, wing, ow = cases.wing_case(root, periodic)
    steps = 6
    say(f"case: Li & Nabawy wing, {SHAPE}, root {root}, periodic {periodic}, cycle {CYCLE}, {TIME_STEP} per step, "
        f"geometry and links, steps 0-{steps}")
    with bh.BirdFlowSimulation(cfg) as sim:
        sim.set_flapping_wing(wing)
        sim.init_equilibrium()
        prev = None
        changed = 0
        total_wrapped = 0
        for s in range(steps + 1):
            if s > 0:
                sim.advance(1)
            fr_o, m_o, w_o = oracle_geometry(p, ow, s, prev)
            fr = sim.wing_frame()
            mask, wall = sim.read_geometry()
            links = sim.wing_links()
            links_o = bfref.wing_links(p, ow, fr_o, m_o).reshape(19, -1)
            guard = (m_o != 0) if prev is None else ((m_o != 0) | (prev != 0))
            n_links = 0
            wrapped = 0
            link_ok = True
            for q in range(1, 19):
                want = source_is_solid(m_o, q, periodic)
                n_links += int(want.sum())
                wrapped += int((want & ~source_is_solid(m_o, q)).sum())
                link_ok &= bool(np.array_equal(links[q][want].view(np.uint32), links_o[q][want].view(np.uint32)))
                # wrapped sources carry a bisected fraction, not the 1e-4 floor
                link_ok &= bool((links_o[q][want] > 1e-4).mean() > 0.5) if want.any() else True
            if prev is not None:
                changed += int((prev != m_o).sum())
            total_wrapped += wrapped
            say(f"  step {s}: wing cells {int((m_o != 0).sum())}, links {n_links} ({wrapped} across a face), mask equal "
                f"{np.array_equal(mask, m_o)}, wall equal where read {np.array_equal(wall[guard].view(np.uint32), w_o[guard].view(np.uint32))}, "
                f"links equal {link_ok}")
            assert np.array_equal(fr.view(np.uint32), fr_o.view(np.uint32))
            assert np.array_equal(mask, m_o)
            assert np.array_equal(wall[guard].view(np.uint32), w_o[guard].view(np.uint32))
            assert link_ok and n_links > 500
            prev = m_o
        assert changed > 0
        assert (total_wrapped > 0) == periodic


@pytest.mark.parametrize("root,periodic", [(ROOT, False), (CROSSING_ROOT, True)])
def test_wing_fluid_steps_match_oracle_bit_for_bit(root, periodic):
    p, cfg, wing, ow = cases.wing_case(root, periodic)
    steps = 8
    say(f"case: Li & Nabawy wing fluid steps, {SHAPE}, root {root}, periodic {periodic}, TRT, link rule, mode 6, "
        f"strict FP32, {steps} steps")
    with bh.BirdFlowSimulation(cfg) as sim:
        sim.set_flapping_wing(wing)
        sim.init_equilibrium()
        sim.advance(steps)
        g = sim.read_populations()
        loads = sim.load_history()
    _, m, w = oracle_geometry(p, ow, 0, None)
    f = bfref.init_equilibrium(p, m, w)
    prev = m
    loads_o = []
    for s in range(1, steps + 1):
        fr, m, w = oracle_geometry(p, ow, s, prev)
        links = bfref.wing_links(p, ow, fr, m)
        f, ld, _, _ = bfref.step_with_links(p, f, prev, m, w, links)
        loads_o.append(ld)
        prev = m
    loads_o = np.array(loads_o)
    fe = np.abs(loads[:, :3] - loads_o[:, :3]).max() / max(np.abs(loads_o[:, :3]).max(), 1e-30)
    te = np.abs(loads[:, 3:6] - loads_o[:, 3:6]).max() / max(np.abs(loads_o[:, 3:6]).max(), 1e-30)
    say(f"  populations differing {int((g != f).sum())}; load rel err force {fe:.2e} torque {te:.2e}; "
        f"last force {loads[-1][:3]}")
    assert np.array_equal(g.view(np.uint32), f.view(np.uint32))
    assert fe <= 1e-12 and te <= 1e-12 and np.array_equal(loads[:, 6:], loads_o[:, 6:])
    assert np.abs(loads_o[:, :3]).max() > 0


def test_link_rule_needs_the_wing_and_the_adjoint_refuses_it():
    p, cfg, wing, _ = cases.wing_case()
    say("case: link rule without the wing; adjoint of a wing")
    with bh.BirdFlowSimulation(cfg) as sim:
        sim.init_equilibrium()
        with pytest.raises(bh.BirdFlowError):
            sim.advance(1)
    with bh.BirdFlowSimulation(cfg) as sim:
        sim.set_flapping_wing(wing)
        sim.init_equilibrium()
        with pytest.raises(bh.BirdFlowError):
            sim.adjoint_gradient(2, np.zeros((2, 6)))


@pytest.mark.parametrize("storage,arithmetic", [(bh.STORAGE_FP32, bh.ARITHMETIC_FUSED),
                                                (bh.STORAGE_FP16_OFFSET, bh.ARITHMETIC_DEFAULT),
                                                (bh.STORAGE_BF16_OFFSET, bh.ARITHMETIC_DEFAULT)])
def test_wing_link_rule_in_the_other_forms_tracks_strict_fp32(storage, arithmetic):
    """The link rule in the fused-FP32 and FP16 / BF16 offset instantiations:
    the same link table (bit-identical: the geometry does not depend on the
    storage) and populations and loads close to the strict FP32 run."""
    p, cfg, wing, _ = cases.wing_case()
    steps = 8
    say(f"case: wing link rule, storage {storage}, arithmetic {arithmetic}, {steps} steps vs strict FP32")
    out = []
    for st, ar in ((bh.STORAGE_FP32, bh.ARITHMETIC_STRICT), (storage, arithmetic)):
        cfg.storage, cfg.arithmetic = st, ar
        with bh.BirdFlowSimulation(cfg) as sim:
            sim.set_flapping_wing(wing)
            sim.init_equilibrium()
            sim.advance(steps)
            out.append((sim.read_populations(), sim.load_history(), sim.wing_links()))
    (f32, l32, k32), (f, l, k) = out
    pop = float(np.abs(f - f32).max())
    force = float(np.abs(l[:, :3] - l32[:, :3]).max() / np.abs(l32[:, :3]).max())
    say(f"  max |population diff| {pop:.2e}; force history max diff / max force {force:.2e}")
    assert np.array_equal(k.view(np.uint32), k32.view(np.uint32))
    assert np.all(np.isfinite(f)) and np.all(np.isfinite(l))
    assert pop <= 2e-4 and force <= 2e-3


def test_li_nabawy_fifth_cycle_mean_lift():
    """The reference's physics check of the prescribed wing
    (MetalFlappingWingValidation.swift:528-569, 1736-1782): Li & Nabawy
    (2022, doi 10.3390/insects13050459) Re 100 beta planform, AR 3, fifth
    cycle mean lift coefficient 1.460 (their Table 2; the reference accepts
    30 %).  Domain 10 x 10 x 8 chords, root at (0.5, 0.5, 0.65) of it, speed
    at the radius of gyration 0.035 lattice, CL = mean F_z / (U^2 AR c^2 / 2).
    At 16 cells per chord this build gives 1.424; the bar here is 10 %."""
    import math
    c = 16
    rg, ar, half = 0.5593218135819926, 3.0, math.radians(80.0)
    travel = 4 * half * rg * ar
    cycle = int(round(travel * c / 0.035))
    speed = travel * c / cycle
    tau = 3 * speed * c / 100.0 + 0.5
    shape = (10 * c, 10 * c, 8 * c)
    root = (0.5 * shape[0], 0.5 * shape[1], 0.65 * shape[2])
    cfg = bh.SolverConfig(shape=shape, omega_plus=1 / tau, omega_minus=1 / (0.1875 / (tau - 0.5) + 0.5),
                          storage=bh.STORAGE_FP32, collision=bh.COLLISION_TRT, wall_rule=bh.WALL_INTERPOLATED_LINKS,
                          force_mode=6, accumulate_loads=True, boundary=((bh.BC_FARFIELD, bh.BC_FARFIELD),) * 3,
                          cell_size=1.0, time_step=1.0, velocity_to_lattice=1.0, sponge_strength=0.04,
                          sponge_width=8.0, body_position=root)
    say(f"case: Li & Nabawy wing, chord {c} cells, {shape}, cycle {cycle} steps, 5 cycles")
    with bh.BirdFlowSimulation(cfg) as sim:
        sim.set_flapping_wing(bh.FlappingWing.li_nabawy(float(c), root, float(cycle)))
        sim.init_equilibrium()
        sim.advance(5 * cycle)
        h = sim.load_history(1 << 16)
    assert h.shape[0] == 5 * cycle and np.all(np.isfinite(h))
    lift = float(h[4 * cycle:, 2].mean() / (0.5 * speed * speed * ar * c * c))
    say(f"  fifth-cycle mean CL {lift:.3f} (Li & Nabawy 1.460)")
    assert abs(lift - 1.460) / 1.460 <= 0.10
