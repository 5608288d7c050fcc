"""Pins the oracle's collision operators and its interpolated walls by what
the mathematics fixes, independently of the oracle's own formulas.

The FP64 instantiation of the oracle step (oracle/bfref_fluid.inc) is run on
a single periodic cell (every pull source is the cell itself, so one step is
one collision) or on small lattices, and checked in moment space with Hermite
tensors built here by einsum over the D3Q19 velocities:

  * every operator conserves mass and momentum, and leaves an equilibrium
    unchanged (BirdFlow.metal:2888-3174);
  * regularized: the post-collision second-order non-equilibrium moment is
    (1 - omega+) Pi_neq and every third-order Hermite moment is zero;
  * RR3 (Malaspinas 2015 recursive closure; the D3Q19-supported mixed set of
    Coreixas et al. 2017): second order as above, and the third-order moments
    equal (1 - omega+) / (2 cs^6) G a3, with a3_abg = u_a Pi_bg + u_b Pi_ag +
    u_g Pi_ab and G the Gram matrix of the six mixed Hermite polynomials on
    D3Q19; RR3 reduces to regularized at u = 0;
  * positivity limiter (BirdFlow.metal:2967-3037): a convex scale s = 1 -
    reported restriction of the unlimited non-equilibrium part, with the
    binding population exactly at its floor (s is the largest admissible);
  * symmetric TRT limiter (BirdFlow.metal:3095-3150): the antisymmetric part
    equals TRT's, the symmetric increment is TRT's scaled by one s, and the
    binding population sits at its floor; inactive, the result is TRT's bit
    for bit;
  * the step's inline Bouzidi branches (BirdFlow.metal:2729-2806) against the
    KAT-pinned two-branch formula (InterpolatedBounceBackReference.swift:28-42,
    MetalExecutionTests.swift:2244-2284) on single links with a moving wall,
    q in {0.25, 0.5, 0.75}, including the halfway fallback and the link's
    mode-6 force and torque;
  * a uniform flow past co-moving walls at arbitrary link fractions is a fixed
    point of every branch (Galilean consistency of the moving-wall term);
  * steady plane Couette flow between off-lattice walls (q > 1/2 on both
    sides) reproduces the exact linear profile to round-off: linear
    interpolation of populations that are linear in space is exact
    (the reference's own moving-wall gate, MetalExecutionTests.swift:1076-1129,
    runs the halfway case).
"""
import numpy as np
import pytest

import bfref

C = bfref.C.astype(np.float64)
W = np.asarray(bfref.W, np.float64)
OPP = np.asarray(bfref.OPP)
CS2 = 1.0 / 3.0
EYE = np.eye(3)
H2 = np.einsum("qa,qb->qab", C, C) - CS2 * EYE[None]
H3 = (np.einsum("qa,qb,qg->qabg", C, C, C)
      - CS2 * (np.einsum("qa,bg->qabg", C, EYE) + np.einsum("qb,ag->qabg", C, EYE)
               + np.einsum("qg,ab->qabg", C, EYE)))
# the six mixed third-order Hermite components D3Q19 supports
MIXED = [(0, 0, 1), (0, 0, 2), (0, 1, 1), (0, 2, 2), (1, 1, 2), (1, 2, 2)]
B3 = np.stack([H3[:, a, b, g] for a, b, g in MIXED])          # [6, 19]
GRAM = np.einsum("kq,q,jq->kj", B3, W, B3)


def feq_np(rho, u):
    cu = C @ u
    return W * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * (u @ u))


def collide(f, collision, omega_plus, omega_minus=None):
    """One collision of one cell through the FP64 oracle step (1x1x1 periodic)."""
    if omega_minus is None:
        omega_minus = omega_plus
    p = bfref.make_params((1, 1, 1), omega_plus=omega_plus, omega_minus=omega_minus, collision=collision,
                          accumulate_loads=True)
    out, loads, *_ = bfref.step(p, np.asarray(f, np.float64).copy())
    return out, loads, float(np.float32(omega_plus)), float(np.float32(omega_minus))


def macro(f):
    rho = f.sum()
    return rho, (f @ C) / rho


def pi_neq(f):
    rho, u = macro(f)
    return np.einsum("q,qab->ab", f - feq_np(rho, u), H2)


def closure_a3(u, pi):
    return np.einsum("a,bg->abg", u, pi) + np.einsum("b,ag->abg", u, pi) + np.einsum("g,ab->abg", u, pi)


def near_equilibrium(rng, scale=0.02):
    rho = 1.0 + 0.05 * rng.standard_normal()
    u = 0.06 * rng.standard_normal(3)
    f = feq_np(rho, u) * (1.0 + scale * rng.standard_normal(19))
    return f


OPERATORS = [bfref.COLLISION_TRT, bfref.COLLISION_TRT_SYMMETRIC_LIMITER, bfref.COLLISION_REGULARIZED,
             bfref.COLLISION_RR3]


@pytest.mark.parametrize("collision", OPERATORS)
def test_collisions_conserve_mass_and_momentum_and_keep_equilibrium(collision):
    rng = np.random.default_rng(11 + collision)
    for trial in range(40):
        f = near_equilibrium(rng, 0.02 if trial < 20 else 0.6)
        post, *_ = collide(f, collision, 1.6, 1.1)
        assert abs(post.sum() - f.sum()) < 1e-14
        assert np.abs(post @ C - f @ C).max() < 1e-15
    for _ in range(5):
        rho, u = 1.0 + 0.05 * rng.standard_normal(), 0.06 * rng.standard_normal(3)
        f = feq_np(rho, u)
        post, *_ = collide(f, collision, 1.7, 0.31)
        assert np.abs(post - f).max() < 2e-16


@pytest.mark.parametrize("omega", [0.8, 1.25, 1.9])
def test_regularized_relaxes_stress_and_removes_third_order(omega):
    rng = np.random.default_rng(3)
    for _ in range(30):
        f = near_equilibrium(rng)
        post, loads, op, _ = collide(f, bfref.COLLISION_REGULARIZED, omega)
        assert loads[6] == 0.0
        rho, u = macro(f)
        neq = post - feq_np(rho, u)
        m2 = np.einsum("q,qab->ab", neq, H2)
        assert np.abs(m2 - (1.0 - op) * pi_neq(f)).max() < 1e-15
        m3 = np.einsum("q,qabg->abg", neq, H3)
        assert np.abs(m3).max() < 1e-16


@pytest.mark.parametrize("omega", [0.8, 1.25, 1.9])
def test_rr3_third_order_moments_follow_recursive_closure(omega):
    # On D3Q19 each mixed Hermite polynomial is normalised by 1/(2 cs^6) = 13.5
    assert np.allclose(np.diag(GRAM) * 13.5, 1.0, atol=1e-15)
    rng = np.random.default_rng(5)
    for _ in range(30):
        f = near_equilibrium(rng)
        post, loads, op, _ = collide(f, bfref.COLLISION_RR3, omega)
        assert loads[6] == 0.0
        rho, u = macro(f)
        pi = pi_neq(f)
        neq = post - feq_np(rho, u)
        m2 = np.einsum("q,qab->ab", neq, H2)
        assert np.abs(m2 - (1.0 - op) * pi).max() < 1e-15
        a3 = closure_a3(u, pi)
        a_mixed = np.array([a3[k] for k in MIXED])
        m3_mixed = B3 @ neq
        expect = (1.0 - op) * 0.5 / CS2 ** 3 * (GRAM @ a_mixed)
        assert np.abs(m3_mixed - expect).max() < 1e-16 + 1e-12 * np.abs(expect).max()
        # the third-order part is not negligible in these cases
        assert np.abs(expect).max() > 1e-7
        # unsupported xyz moment and pure cubic moments vanish on D3Q19
        m3 = np.einsum("q,qabg->abg", neq, H3)
        assert abs(m3[0, 1, 2]) < 1e-17 and max(abs(m3[a, a, a]) for a in range(3)) < 1e-17


def test_rr3_reduces_to_regularized_at_rest():
    rng = np.random.default_rng(9)
    for _ in range(10):
        f = feq_np(1.0 + 0.05 * rng.standard_normal(), np.zeros(3)) * (1 + 0.02 * rng.standard_normal(19))
        # remove the momentum the noise added so that u = 0 exactly
        f = f - W * 3.0 * (C @ (f @ C))
        assert np.abs(f @ C).max() < 1e-16
        a, *_ = collide(f, bfref.COLLISION_RR3, 1.3)
        b, *_ = collide(f, bfref.COLLISION_REGULARIZED, 1.3)
        assert np.array_equal(a, b)


def _floor(feq):
    return np.maximum(1e-12, 1e-6 * np.maximum(feq, 0.0))


@pytest.mark.parametrize("collision", [bfref.COLLISION_REGULARIZED, bfref.COLLISION_RR3])
def test_positivity_limiter_scales_to_the_largest_admissible(collision):
    rng = np.random.default_rng(21 + collision)
    active = 0
    for _ in range(400):
        rho, u = 1.0 + 0.05 * rng.standard_normal(), 0.1 * rng.standard_normal(3)
        f = feq_np(rho, u) * (1.0 + 0.8 * rng.standard_normal(19))
        f = np.abs(f)
        post, loads, op, _ = collide(f, collision, 1.95)
        rho, u = macro(f)
        feq = feq_np(rho, u)
        if loads[6] == 0.0:
            continue
        active += 1
        s = 1.0 - loads[7]
        assert 0.0 <= s < 1.0
        # conserved moments survive the convex scale
        assert abs(post.sum() - f.sum()) < 1e-13 and np.abs(post @ C - f @ C).max() < 1e-14
        # the limited non-equilibrium part is the unlimited one scaled by s
        neq = post - feq
        m2 = np.einsum("q,qab->ab", neq, H2)
        assert np.abs(m2 - s * (1.0 - op) * pi_neq(f)).max() < 1e-13
        if collision == bfref.COLLISION_RR3:
            a_mixed = np.array([closure_a3(u, pi_neq(f))[k] for k in MIXED])
            expect = s * (1.0 - op) * 13.5 * (GRAM @ a_mixed)
            assert np.abs(B3 @ neq - expect).max() < 1e-13
        else:
            assert np.abs(np.einsum("q,qabg->abg", neq, H3)).max() < 1e-13
        if (feq < _floor(feq)).any():
            # an equilibrium below its own floor admits no positive scale
            assert s == 0.0 and np.abs(post - feq).max() < 1e-15
            continue
        # every population at or above its floor; the binding one exactly on it
        gap = post - _floor(feq)
        assert gap.min() > -1e-15
        assert abs(gap.min()) < 1e-14
    assert active >= 20, active


def test_symmetric_limiter_keeps_antisymmetric_part_and_scales_symmetric():
    rng = np.random.default_rng(31)
    active = 0
    for _ in range(400):
        rho, u = 1.0 + 0.05 * rng.standard_normal(), 0.1 * rng.standard_normal(3)
        f = np.abs(feq_np(rho, u) * (1.0 + 0.8 * rng.standard_normal(19)))
        post, loads, op, om = collide(f, bfref.COLLISION_TRT_SYMMETRIC_LIMITER, 1.95, 1.2)
        trt, *_ = collide(f, bfref.COLLISION_TRT, 1.95, 1.2)
        if loads[6] == 0.0:
            assert np.array_equal(post, trt)
            continue
        active += 1
        s = 1.0 - loads[7]
        sym = lambda g: 0.5 * (g + g[OPP])
        anti = lambda g: 0.5 * (g - g[OPP])
        assert np.abs(anti(post) - anti(trt)).max() < 1e-16
        assert np.abs((sym(post) - sym(f)) - s * (sym(trt) - sym(f))).max() < 1e-15
        rho, u = macro(f)
        feq = feq_np(rho, u)
        # binding population: the symmetric increment pushes it down to its floor
        dec = (sym(trt) - sym(f)) < 0
        if s == 0.0:
            continue
        gap = (post - _floor(feq))[dec]
        assert gap.min() > -1e-15 and abs(gap.min()) < 1e-14
    assert active >= 20, active


# ---------------------------------------------------------------- walls

def _link_case(q, lf, n=7, wall_solid=(0.03, -0.02, 0.01), wall_fluid=(0.0, 0.0, 0.0), farther_solid=False,
               seed=0):
    """A single solid cell with its 18 fluid neighbours; the link under test is
    direction q at the fluid cell x = s + c_q.  Returns everything the
    expected value needs."""
    rng = np.random.default_rng(seed)
    p = bfref.make_params((n, n, n), omega_plus=0.0, omega_minus=0.0, wall_rule=bfref.WALL_INTERPOLATED_SDF,
                          accumulate_loads=True, force_scale=1.0, dx=1.0, body_position=(1.25, -0.5, 2.0))
    N = n ** 3
    gid = lambda x, y, z: x + n * (y + n * z)
    s = np.array([3, 3, 3])
    x = s + bfref.C[q]
    far = x + bfref.C[q]
    mask = np.zeros(N, np.uint8)
    mask[gid(*s)] = 1
    if farther_solid:
        mask[gid(*far)] = 1
    wall = np.zeros((N, 4), np.float32)
    wall[:, 3] = 16.0
    wall[:, :3] = wall_fluid
    wall[gid(*s), :3] = wall_solid
    wall[gid(*s), 3] = np.float32(lf - 1.0)
    wall[gid(*x), 3] = np.float32(lf)
    f = (np.repeat(W, N) * (1.0 + 0.3 * rng.random(19 * N))).astype(np.float64)
    return p, N, gid, s, x, far, mask, wall, f


@pytest.mark.parametrize("q", [1, 4, 7, 12, 17])
@pytest.mark.parametrize("lf", [0.25, 0.5, 0.75])
def test_step_bouzidi_link_matches_kat_formula(q, lf):
    p, N, gid, s, x, far, mask, wall, f = _link_case(q, lf, seed=q)
    out, loads, *_ = bfref.step(p, f, mask, mask, wall)      # omega = 0: the step streams only
    xi, fi = gid(*x), gid(*far)
    o = OPP[q]
    assert bfref.link_fraction_f64(float(wall[xi, 3]), float(wall[gid(*s), 3])) == lf
    uw = wall[xi, :3].astype(np.float64) + (wall[gid(*s), :3].astype(np.float64) - wall[xi, :3]) * lf
    corr = 2.0 * W[q] * 1.0 * (C[q] @ uw) / CS2
    expect = bfref.bouzidi_population_f64(lf, f[o * N + xi], f[o * N + fi], f[q * N + xi], corr)
    assert abs(out[q * N + xi] - expect) < 1e-15, (out[q * N + xi], expect)


@pytest.mark.parametrize("q", [2, 9, 15])
def test_step_bouzidi_falls_back_to_halfway_when_far_side_is_solid(q):
    # q <= 1/2 needs the population of x + c_q; a solid far side (or one out of
    # the domain) selects halfway bounce-back with the solid's own velocity
    p, N, gid, s, x, far, mask, wall, f = _link_case(q, 0.25, farther_solid=True, wall_fluid=(0.01, 0.0, 0.0),
                                                     seed=40 + q)
    out, *_ = bfref.step(p, f, mask, mask, wall)
    xi = gid(*x)
    corr = 2.0 * W[q] * (C[q] @ wall[gid(*s), :3].astype(np.float64)) / CS2
    assert abs(out[q * N + xi] - (f[OPP[q] * N + xi] + corr)) < 1e-15


@pytest.mark.parametrize("mode", [6, 0])
@pytest.mark.parametrize("lf", [0.25, 0.75])
def test_step_link_force_and_torque_use_the_wall_point(lf, mode):
    # Every fluid neighbour of the single solid cell has the same fluid-side
    # distance, so all 18 links share the fraction lf.  Per link, mode 6:
    # F = -(value + reflected) c_q; mode 0 (momentum flux in the wall frame,
    # BirdFlow.metal:2812-2873): F = -(value (c_q - u_w) + reflected (c_q + u_w));
    # torque arm x - lf c_q dx - body (mode 0's force is not along c_q, so the
    # arm's offset to the wall point shows).
    p, N, gid, s, x, far, mask, wall, f = _link_case(7, lf, wall_fluid=(0.0, 0.0, 0.0), seed=77)
    p.force_mode = mode
    sv = gid(*s)
    for q in range(1, 19):
        wall[gid(*(s + bfref.C[q])), 3] = np.float32(lf)
    out, loads, *_ = bfref.step(p, f, mask, mask, wall)
    force = np.zeros(3)
    torque = np.zeros(3)
    body = np.array([1.25, -0.5, 2.0])
    for q in range(1, 19):
        xq = s + bfref.C[q]
        xi, fi = gid(*xq), gid(*(xq + bfref.C[q]))
        uw = wall[sv, :3].astype(np.float64) * lf
        corr = 2.0 * W[q] * (C[q] @ uw) / CS2
        value = bfref.bouzidi_population_f64(lf, f[OPP[q] * N + xi], f[OPP[q] * N + fi], f[q * N + xi], corr)
        reflected = f[OPP[q] * N + xi]
        if mode == 6:
            fl = -(value + reflected) * C[q]
        else:
            fl = -(value * (C[q] - uw) + reflected * (C[q] + uw))
        force += fl
        torque += np.cross(xq + 0.5 - lf * C[q] - body, fl)
    assert np.abs(loads[:3] - force).max() < 1e-14
    assert np.abs(loads[3:6] - torque).max() < 1e-13


@pytest.mark.parametrize("collision", OPERATORS)
def test_uniform_flow_past_comoving_walls_is_a_fixed_point(collision):
    # f = feq(1, U) everywhere and every wall moving at U: each Bouzidi branch
    # returns feq_q(U) exactly (reflected + 6 w c.U = feq_q; the interpolation
    # weights sum to one), whatever the link fraction.
    n = 10
    rng = np.random.default_rng(collision)
    U = np.array([0.04, -0.015, 0.02], np.float32).astype(np.float64)   # the wall field is FP32
    p = bfref.make_params((n, n, n), omega_plus=1.6, omega_minus=1.1, collision=collision,
                          wall_rule=bfref.WALL_INTERPOLATED_SDF)
    N = n ** 3
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    r = np.sqrt((x - 4.6) ** 2 + (y - 5.1) ** 2 + (z - 4.8) ** 2).ravel()
    mask = (r < 2.3).astype(np.uint8)
    wall = np.zeros((N, 4), np.float32)
    wall[:, :3] = U
    wall[:, 3] = np.where(mask == 1, -rng.random(N), rng.random(N)).astype(np.float32)
    f = np.concatenate([np.full(N, feq_np(1.0, U)[q]) for q in range(19)])
    g = f.copy()
    for _ in range(5):
        g, *_ = bfref.step(p, g, mask, mask, wall)
    fluid = np.tile(mask == 0, 19)
    assert np.abs(g - f)[fluid].max() < 5e-16


@pytest.mark.parametrize("collision", [bfref.COLLISION_TRT, bfref.COLLISION_REGULARIZED])
@pytest.mark.parametrize("q_lo,q_hi", [(0.75, 0.6), (0.55, 0.95), (0.5, 0.5)])
def test_couette_between_offlattice_walls_is_exact(collision, q_lo, q_hi):
    nx, ny, nz = 2, 12, 2
    U = 1e-3
    tau = 0.8
    omega_plus = 1.0 / tau
    omega_minus = 1.0 / (0.5 + 0.25 / (tau - 0.5))
    p = bfref.make_params((nx, ny, nz), omega_plus=omega_plus, omega_minus=omega_minus, collision=collision,
                          wall_rule=bfref.WALL_INTERPOLATED_SDF)
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[:, 0, :] = 1
    solid[:, ny - 1, :] = 1
    wall = np.zeros((nz, ny, nx, 4), np.float64)
    wall[..., 3] = 16.0
    wall[:, 1, :, 3], wall[:, 0, :, 3] = q_lo, q_lo - 1.0
    wall[:, ny - 2, :, 3], wall[:, ny - 1, :, 3] = q_hi, q_hi - 1.0
    wall[:, ny - 2:, :, 0] = U                 # top wall moves in x, bottom at rest
    y_lo, y_hi = 1.0 - q_lo, ny - 2 + q_hi      # wall planes in cell-centre coordinates
    wall = wall.reshape(-1, 4)
    f = bfref.init_equilibrium(p, solid.ravel(), wall.astype(np.float32), np.float64)
    for _ in range(3000):
        f, *_ = bfref.step(p, f, solid.ravel(), solid.ravel(), wall)
    _, u = bfref.moments(p, f)
    u = u.reshape(nz, ny, nx, 3)
    y = np.arange(1, ny - 1)
    exact = U * (y - y_lo) / (y_hi - y_lo)
    assert np.abs(u[:, 1:ny - 1, :, 0] - exact[None, :, None]).max() < 1e-13
    assert np.abs(u[:, 1:ny - 1, :, 1:]).max() < 1e-14
