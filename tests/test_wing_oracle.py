"""Pins the oracle's prescribed beta-planform wing (BirdFlow.metal:940-1297)
on the reference's Li & Nabawy case with closed forms, none retyping the
oracle's expressions:

  * stroke: +A at phase 0 with zero rate, 0 at mid-downstroke, -A at phase
    1/2 with zero rate; angle and rate continuous at every segment boundary;
    rate = d angle / d phase;
  * pitch: 90 deg at the stroke reversals, 45 deg through the downstroke,
    135 deg through the upstroke, rate = d pitch / d phase;
  * frame: span (cos s, sin s, 0), chord the horizontal tangent turned by the
    pitch about the span (the paper's sign), orthonormal; angular velocity
    reproduces frame differences;
  * pow against FP64 on the planform's range (its only transcendental);
  * the bisected link fraction finds a flat face to its 2^-10 resolution and
    the trailing edge at the closed-form beta chord;
  * the planform's mean chord is the mean chord (beta normalization);
  * the link table holds a fraction for exactly the fluid cells with a solid
    source.
"""
import math

import numpy as np
import pytest

import bfref

CYCLE = 200.0
HALF = 80.0 * math.pi / 180.0


def wing(chord=16.0, root=(40.5, 40.5, 40.5)):
    return bfref.li_nabawy_wing(chord, root, CYCLE)


def kin(w, phase):
    f = bfref.wing_prepare(w, phase * CYCLE)
    return float(f[5, 0]), float(f[5, 1]), float(f[5, 2]) * CYCLE, float(f[5, 3]) * CYCLE   # rates per cycle


def test_stroke_closed_forms_and_continuity():
    w = wing()
    s, _, r, _ = kin(w, 0.0)
    assert s == pytest.approx(HALF, abs=2e-6) and abs(r) < 2e-5
    s, _, _, _ = kin(w, 0.25)
    assert abs(s) < 2e-6
    s, _, r, _ = kin(w, 0.5)
    assert s == pytest.approx(-HALF, abs=2e-6) and abs(r) < 2e-5
    for b in (0.125, 0.375, 0.625, 0.875):
        lo, hi = kin(w, b - 1e-4), kin(w, b + 1e-4)
        assert abs(lo[0] - hi[0]) < 2e-3 and abs(lo[2] - hi[2]) < 2e-2
    for ph in np.linspace(0.01, 0.99, 37):
        h = 1e-3
        d = (kin(w, ph + h)[0] - kin(w, ph - h)[0]) / (2 * h)
        assert d == pytest.approx(kin(w, ph)[2], rel=2e-3, abs=2e-3)


def test_pitch_closed_forms():
    w = wing()
    assert kin(w, 0.0)[1] == pytest.approx(math.pi / 2, abs=2e-6)
    assert kin(w, 0.5)[1] == pytest.approx(math.pi / 2, abs=2e-6)
    assert kin(w, 0.25)[1] == pytest.approx(math.pi / 4, abs=1e-7) and kin(w, 0.25)[3] == 0.0
    assert kin(w, 0.75)[1] == pytest.approx(3 * math.pi / 4, abs=1e-7) and kin(w, 0.75)[3] == 0.0
    for ph in np.linspace(0.02, 0.98, 25):
        h = 1e-3
        d = (kin(w, ph + h)[1] - kin(w, ph - h)[1]) / (2 * h)
        assert d == pytest.approx(kin(w, ph)[3], rel=5e-3, abs=5e-3)


def test_frame_closed_form_and_angular_velocity():
    w = wing()
    for ph in (0.1, 0.25, 0.4, 0.6, 0.83):
        f = bfref.wing_prepare(w, ph * CYCLE)
        s, p = float(f[5, 0]), float(f[5, 1])
        span = np.array([math.cos(s), math.sin(s), 0.0])
        tangent = np.array([-math.sin(s), math.cos(s), 0.0])
        chord = math.cos(p) * tangent - math.sin(p) * np.array([0.0, 0.0, 1.0])
        assert np.allclose(f[2, :3], span, atol=2e-6)
        assert np.allclose(f[1, :3], chord, atol=2e-6)
        assert np.allclose(f[3, :3], np.cross(span, chord), atol=2e-6)
        h = 0.2
        fp, fm = bfref.wing_prepare(w, ph * CYCLE + h), bfref.wing_prepare(w, ph * CYCLE - h)
        om = f[4, :3].astype(np.float64)
        for row in (1, 2, 3):
            d = (fp[row, :3].astype(np.float64) - fm[row, :3]) / (2 * h)
            assert np.linalg.norm(d - np.cross(om, f[row, :3])) <= 2e-3 * np.linalg.norm(om) + 1e-6


def test_pow_against_fp64():
    rng = np.random.default_rng(2)
    for x in np.concatenate([rng.uniform(1e-6, 0.25, 2000), [0.25, 1e-7, 0.5, 1.0]]):
        for y in (0.48915066, 1.0, 0.3):
            # |y log2 x| <= 20 here: the exponent's rounding bounds the error (~1e-6)
            got = bfref.lib().bfo_wing_pow(float(np.float32(x)), float(np.float32(y)))
            want = float(np.float32(x)) ** float(np.float32(y))
            assert abs(got - want) <= 1.5e-6 * want + 1e-38
    assert bfref.lib().bfo_wing_pow(0.0, 0.489) == 0.0


def test_link_fraction_locates_faces():
    w = wing()
    f = bfref.wing_prepare(w, 0.25 * CYCLE)          # span +x, chord (0, .707, -.707), normal (0, .707, .707)
    root = np.array([40.5, 40.5, 40.5])
    span, chord, normal = (f[k, :3].astype(np.float64) for k in (2, 1, 3))
    half = 0.5 * max(0.05 * 16, 1.0)
    r = 0.5 * 48.0
    # a link along the normal through the upper face at r = R/2, chord 0
    for t_in, t_out in ((0.1, 1.9), (-0.3, 1.7), (0.45, 1.05)):
        solid = root + r * span + t_in * half * normal
        fluid = root + r * span + t_out * half * normal
        q = bfref.wing_link_fraction(w, f, 1.0, fluid, solid)
        exact = (t_out - 1.0) / (t_out - t_in)
        assert q == pytest.approx(exact, abs=2.0 ** -10 + 1e-6)
    # trailing edge along the chord at radial stations: lead + c(r)
    a = 1.4891506583557845
    norm = 2.497931737168364
    for rh in (0.15, 0.3, 0.5, 0.8):
        c = 16.0 * norm * (rh * (1 - rh)) ** (a - 1)
        te = -0.25 * 16.0 + c
        solid = root + rh * 48.0 * span + (te - 0.6) * chord
        fluid = root + rh * 48.0 * span + (te + 0.4) * chord
        q = bfref.wing_link_fraction(w, f, 1.0, fluid, solid)
        assert q == pytest.approx(0.4, abs=2.0 ** -10 + 2e-4)


def test_planform_mean_chord():
    a = 1.4891506583557845
    norm = 2.497931737168364
    r = (np.arange(200000) + 0.5) / 200000
    assert float(np.mean(norm * (r * (1 - r)) ** (a - 1))) == pytest.approx(1.0, abs=1e-3)


def solid_sources(mask, shape, q, periodic):
    """Fluid cells whose pull source cell - c_q is solid, [N] bool: wrapped
    across the faces when periodic, else only in-domain sources."""
    nx, ny, nz = shape
    m3 = mask.reshape(nz, ny, nx)
    cx, cy, cz = bfref.C[q]
    if periodic:
        src = np.roll(m3, (cz, cy, cx), axis=(0, 1, 2))
    else:
        src = np.zeros_like(m3)
        src[max(cz, 0):nz + min(cz, 0), max(cy, 0):ny + min(cy, 0), max(cx, 0):nx + min(cx, 0)] = \
            m3[max(-cz, 0):nz + min(-cz, 0), max(-cy, 0):ny + min(-cy, 0), max(-cx, 0):nx + min(-cx, 0)]
    return (src.ravel() != 0) & (mask == 0)


@pytest.mark.parametrize("root,phase,periodic", [((20.5, 20.5, 20.5), 0.3, True), ((20.5, 20.5, 20.5), 0.3, False),
                                                 ((44.5, 20.5, 20.5), 0.0, True), ((44.5, 20.5, 20.5), 0.0, False)])
def test_link_table_covers_exactly_the_solid_sources(root, phase, periodic):
    """Every fluid cell whose pull source is on the wing has a fraction, and
    no other cell: with periodic faces the source wraps (a wing rooted 3.5
    cells from the x = 48 face at the stroke reversal reaches its last plane, so cells of the first
    plane pull from it), behind a far-field face there is no source."""
    shape = (48, 40, 40)
    w = wing(chord=8.0, root=root)
    bc = ((bfref.BC_PERIODIC, bfref.BC_PERIODIC) if periodic else (bfref.BC_FARFIELD, bfref.BC_FARFIELD),) * 3
    p = bfref.make_params(shape, omega_plus=1.0, omega_minus=1.0, bc=bc)
    f = bfref.wing_prepare(w, phase * CYCLE)
    mask, wall = bfref.wing_build(p, w, f)
    assert (mask != 0).sum() > 100
    links = bfref.wing_links(p, w, f, mask).reshape(19, -1)
    wrapped = 0
    for q in range(1, 19):
        want = solid_sources(mask, shape, q, periodic)
        wrapped += int((want & ~solid_sources(mask, shape, q, False)).sum())
        have = links[q] != 0
        assert np.array_equal(want, have)
        assert (links[q][have] >= 1e-4).all() and (links[q][have] <= 1.0).all()
    if periodic and root[0] > 40:
        # the wing crosses the face; its wrapped links are bisected, not
        # left at the minimum fraction
        assert wrapped > 20
        m3 = mask.reshape(40, 40, 48)
        assert m3[:, :, -1].any()
    else:
        assert wrapped == 0
    # on-wing cells carry -1, their guard neighbourhood +1, and the rigid rotation
    assert (wall[mask != 0, 3] == -1).all()
    near = (mask == 0) & (wall[:, 3] == 1.0)
    assert near.sum() > 0
