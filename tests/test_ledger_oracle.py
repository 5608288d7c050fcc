"""Pins the oracle's momentum ledger (BirdFlow.metal:4272-4526) against the
oracle's own fluid step, an independent code path: the step conserves mass
and momentum up to FP32 rounding except across open boundaries, in the
sponge and at walls, so

  * a periodic box without walls or sponge: change = 0, no sources;
  * far-field or outlet faces and a sponge: change = open + sponge;
  * a static moving wall, halfway rule: change - open - sponge = the
    persistent-link exchange, and that exchange is minus the step's own
    momentum-exchange force (lattice units).

A dropped term, a wrong sign or a wrong source index breaks one of them."""
import numpy as np
import pytest

import bfref


def perturbed(p, rng, u=(0.03, -0.01, 0.02), amp=0.02):
    n = bfref.cells(p)
    f = np.empty((19, n), np.float32)
    for q in range(19):
        f[q] = bfref.equilibrium_f32(q, 1.0, u)
    return (f * rng.uniform(1 - amp, 1 + amp, f.shape)).astype(np.float32).ravel()


def closure_bar(p):
    return 5e-8 * bfref.cells(p)


def test_periodic_box_conserves_and_has_no_sources():
    rng = np.random.default_rng(2)
    p = bfref.make_params((24, 16, 12), omega_plus=1.3, omega_minus=1.1)
    f0 = perturbed(p, rng)
    f1, _, _, _ = bfref.step(p, f0)
    z = np.zeros(bfref.cells(p), np.uint8)
    sums, counts = bfref.ledger_sources(p, f0, f1, z, z, np.zeros((bfref.cells(p), 4), np.float32))
    assert np.abs(sums[0]).max() <= closure_bar(p)
    assert np.array_equal(sums[1:], np.zeros((3, 4))) and np.array_equal(counts, [0, 0, 0, 0])
    # the fluid totals move by the same change
    d = bfref.ledger_fluid(p, f1, z) - bfref.ledger_fluid(p, f0, z)
    assert np.abs(d - sums[0]).max() <= 1e-9 * bfref.cells(p)


@pytest.mark.parametrize("faces", ["farfield", "outlet"])
def test_open_faces_and_sponge_close_the_fluid_change(faces):
    rng = np.random.default_rng(3)
    bc = bfref.BC_FARFIELD if faces == "farfield" else bfref.BC_OUTLET
    p = bfref.make_params((20, 18, 14), omega_plus=1.4, omega_minus=1.2,
                          bc=((bfref.BC_FARFIELD, bc), (bc, bc), (bfref.BC_PERIODIC, bfref.BC_PERIODIC)),
                          sponge_strength=0.2, sponge_width=3.0, far_u=(0.04, 0.0, 0.01))
    f0 = perturbed(p, rng)
    f1, _, _, _ = bfref.step(p, f0)
    z = np.zeros(bfref.cells(p), np.uint8)
    sums, counts = bfref.ledger_sources(p, f0, f1, z, z, np.zeros((bfref.cells(p), 4), np.float32))
    residual = sums[0] - sums[1] - sums[2]
    assert np.abs(residual).max() <= closure_bar(p)
    assert np.abs(sums[1]).max() > 100 * closure_bar(p)
    assert np.abs(sums[2]).max() > 100 * closure_bar(p)
    assert counts[0] > 0 and counts[1] > 0 and counts[2] == 0 and counts[3] == 0
    # without the sponge term the balance fails by the sponge's size
    assert np.abs(sums[0] - sums[1]).max() > 50 * closure_bar(p)


def test_static_moving_wall_exchange_matches_the_step_force():
    rng = np.random.default_rng(4)
    p = bfref.make_params((24, 20, 16), omega_plus=1.5, omega_minus=1.2, wall_rule=bfref.WALL_HALFWAY,
                          force_mode=6, accumulate_loads=True, far_u=(0.02, 0.0, 0.0))
    n = bfref.cells(p)
    solid = np.zeros((16, 20, 24), np.uint8)
    solid[5:11, 6:14, 8:16] = 1
    solid = solid.ravel()
    wall = np.zeros((n, 4), np.float32)
    wall[:, 0] = 0.03
    wall[:, 2] = -0.01
    f0 = perturbed(p, rng)
    f1, loads, _, _ = bfref.step(p, f0, solid, solid, wall)
    sums, counts = bfref.ledger_sources(p, f0, f1, solid, solid, wall)
    boundary = sums[0] - sums[1] - sums[2]
    assert counts[2] > 0 and counts[3] == 0
    assert np.abs(boundary - sums[3]).max() <= closure_bar(p)
    # momentum given to the fluid at the wall = minus the force on the body
    assert np.allclose(sums[3][1:], -np.asarray(loads[:3]), rtol=1e-5, atol=closure_bar(p))
    assert np.abs(sums[3][1:]).max() > 100 * closure_bar(p)
