"""Pins the oracle's exact link fractions against a meshed shell (the
ray-root wall rule, bfo_ray_links / bfo_ray_shell) without retyping them:

* closed forms of the first entry of a ray F + t D (t >= 0; 2 reports none
  below 2) into one triangle's shell (the triangle swept by a ball of radius
  h): through an offset face, an edge cylinder, a vertex sphere; a miss; a
  start inside;
* random triangles and segments against an independent FP64 reference: a
  dense scan of the point-triangle distance along the segment (projection
  onto the plane, else the nearest edge point) refined by bisection;
* an axis-aligned box plate (a hollow shell of 12 triangles): every axis
  link through a face interior has the closed-form fraction (fluid node to
  the offset face) / link length, outside and inside the box, and
  the table covers exactly the fluid cells with a solid source;
* the reference's own exact roots: the 15 links of its link-ray-root audit
  (deetjen-dove-moving-wall-link-ray-root.json: the indexed dove at 26.5 ms
  on the D12 and D16 grids, global union, 7.5 mm half thickness; the
  reference scans 256 intervals and bisects 48 times), to 1e-4 of the link
  (measured 4.2e-5 at most, 14 of 15 within 5e-6).
"""
import json
import math
import os

import numpy as np
import pytest

import bfref
from birdflow_hip import workloads

TRI = (np.array([0.0, 0.0, 0.0]), np.array([1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0]))


def shell(F, D, tri=TRI, h=0.5):
    return bfref.ray_shell(F, D, tri[0], tri[1], tri[2], h)


def test_shell_closed_forms():
    # straight down onto the face interior: the offset plane z = h
    assert shell((0.2, 0.2, 3.0), (0.0, 0.0, -4.0)) == pytest.approx((3.0 - 0.5) / 4.0, abs=1e-7)
    # from below: the plane z = -h
    assert shell((0.2, 0.3, -2.0), (0.0, 0.0, 4.0)) == pytest.approx((2.0 - 0.5) / 4.0, abs=1e-7)
    # across the edge y = 0 (beside the face): the edge cylinder at y = -h
    assert shell((0.5, -3.0, 0.0), (0.0, 4.0, 0.0)) == pytest.approx((3.0 - 0.5) / 4.0, abs=1e-7)
    # a slanted approach to the same edge: |(p - a) x e| = h on the cylinder
    t = shell((0.5, -3.0, 1.0), (0.0, 4.0, -1.0))
    y, z = -3.0 + 4.0 * t, 1.0 - t
    assert math.hypot(y, z) == pytest.approx(0.5, abs=2e-6) and y < 0
    # along the line of edge AB toward B: the vertex sphere at B, x = 1 + h
    assert shell((3.0, 0.0, 0.0), (-4.0, 0.0, 0.0)) == pytest.approx((3.0 - 1.5) / 4.0, abs=1e-7)
    # past the far vertex C diagonally: the sphere at C
    t = shell((-2.0, 3.0, 0.0), (4.0, 0.0, 0.0))
    assert math.hypot(-2.0 + 4.0 * t, 3.0 - 1.0) == pytest.approx(0.5 + 1e-9, abs=1e-5) or t == 2.0
    # a miss: parallel to the face 0.6 above it
    assert shell((-1.0, 0.2, 0.6), (3.0, 0.0, 0.0)) == 2.0
    # too short to reach it
    assert shell((0.2, 0.2, 3.0), (0.0, 0.0, -1.0)) == 2.0
    # starting inside: 0
    assert shell((0.2, 0.2, 0.3), (0.0, 0.0, 1.0)) == 0.0
    assert shell((1.2, 0.0, 0.0), (1.0, 0.0, 0.0)) == 0.0


def point_triangle_distance(p, a, b, c):
    """FP64 distance from p to triangle abc: the plane projection when it lies
    inside, else the nearest point of the three edges."""
    n = np.cross(b - a, c - a)
    n = n / np.linalg.norm(n)
    proj = p - np.dot(p - a, n) * n
    inside = all(np.dot(np.cross(v - u, proj - u), n) >= 0 for u, v in ((a, b), (b, c), (c, a)))
    if inside:
        return abs(np.dot(p - a, n))
    best = np.inf
    for u, v in ((a, b), (b, c), (c, a)):
        e = v - u
        s = min(max(np.dot(p - u, e) / np.dot(e, e), 0.0), 1.0)
        best = min(best, np.linalg.norm(p - (u + s * e)))
    return best


def fp64_entry(F, D, tri, h, samples=3000):
    """First t in [0, 1] with distance <= h: dense scan, then bisection."""
    g = lambda t: point_triangle_distance(F + t * D, *tri) - h
    if g(0.0) <= 0:
        return 0.0
    ts = np.linspace(0.0, 1.0, samples + 1)
    prev = 0.0
    for t in ts[1:]:
        if g(t) <= 0:
            lo, hi = prev, t
            for _ in range(60):
                mid = 0.5 * (lo + hi)
                if g(mid) <= 0:
                    hi = mid
                else:
                    lo = mid
            return hi
        prev = t
    return 2.0


def test_shell_against_fp64_scan_and_bisection():
    rng = np.random.default_rng(11)
    hits = 0
    for trial in range(120):
        tri = tuple(rng.uniform(-1.0, 1.0, 3) for _ in range(3))
        h = float(rng.uniform(0.1, 0.6))
        F = rng.uniform(-2.0, 2.0, 3)
        # aimed near a random point of the triangle (barycentric weights
        # slightly outside [0, 1] reach the edges and vertices too), the
        # segment ending short of it, across it or beyond
        w = rng.dirichlet((1.0, 1.0, 1.0)) * 1.3 - 0.1
        target = w[0] * tri[0] + w[1] * tri[1] + w[2] * tri[2] + rng.normal(scale=0.3, size=3)
        D = (target - F) * rng.uniform(0.5, 1.6)
        tri32 = tuple(v.astype(np.float32) for v in tri)
        F32, D32 = F.astype(np.float32), D.astype(np.float32)
        got = shell(F32, D32, tri32, np.float32(h))
        # the shell reports the first entry along the ray (below 2); the link
        # counts an entry only within the segment
        got = got if got <= 1.0 else 2.0
        want = fp64_entry(F32.astype(np.float64), D32.astype(np.float64),
                          tuple(v.astype(np.float64) for v in tri32), float(np.float32(h)))
        if want == 2.0 or got == 2.0:
            # a miss on one side only may be a graze: the segment's closest
            # approach is then within rounding of h
            if want != got:
                closest = min(point_triangle_distance(F32 + t * D32, *[v.astype(np.float64) for v in tri32])
                              for t in np.linspace(0, 1, 4001))
                assert abs(closest - np.float32(h)) < 1e-4, (trial, got, want)
            continue
        hits += 1
        assert got == pytest.approx(want, abs=3e-5), (trial, got, want)
    assert hits > 30


def box_case(shape=(24, 20, 16)):
    """An axis-aligned box (the flapping plate's 12-triangle box, 6 x 8 x 3
    cells, not grid-aligned) with a half thickness of 0.5 cell."""
    seq = workloads.flapping_plate(1, center=(11.3, 9.8, 7.7), span=8.0, chord=6.0, thickness=3.0, heave=0.0,
                                   pitch_deg=0.0, half_thickness=0.5)
    p = bfref.make_params(shape, omega_plus=1.0, omega_minus=1.0, bc=((bfref.BC_FARFIELD, bfref.BC_FARFIELD),) * 3)
    first, _, _ = bfref.frame_interval(seq.times, 0.0)
    pos, vel = bfref.prepare_vertices(seq.points, seq.times, first, 0.0)
    mask, wall, _ = bfref.raster_resolve(p, pos, vel, seq.triangles, seq.parts, seq.half_thickness, 64)
    return p, seq, pos, mask, wall


def test_ray_links_of_a_box_plate_are_the_closed_form():
    p, seq, pos, mask, wall = box_case()
    links = bfref.ray_links(p, pos, seq.triangles, seq.half_thickness, mask, wall).reshape(19, -1)
    lo, hi = pos.min(axis=0).astype(np.float64), pos.max(axis=0).astype(np.float64)
    h = seq.half_thickness
    nx, ny, nz = p.nx, p.ny, p.nz
    checked = 0
    for q in range(1, 7):          # the six axis directions
        c = np.array(bfref.C[q])
        a = int(np.nonzero(c)[0][0])
        for gid in np.nonzero(links[q])[0]:
            x, y, z = gid % nx, (gid // nx) % ny, gid // (nx * ny)
            F = np.array([x + 0.5, y + 0.5, z + 0.5])
            others = [b for b in range(3) if b != a]
            # a link through the interior of a face: the other two
            # coordinates inside the face, away from its edges
            if not all(lo[b] + 0.05 < F[b] < hi[b] - 0.05 for b in others):
                continue
            # the box is a hollow shell (the raster marks cells within h of
            # its triangles): the ray from F toward its source F - c enters
            # the first face slab [s - h, s + h], s = lo or hi, ahead of it
            entries = [(F[a] - (s + h)) if c[a] > 0 else ((s - h) - F[a]) for s in (lo[a], hi[a])]
            entries = [t for t in entries if t >= 0.0]
            if not entries:
                continue
            want = min(entries)
            assert links[q][gid] == pytest.approx(min(max(want, 1e-4), 1.0), abs=2e-6), (q, gid)
            checked += 1
    assert checked > 50


def test_ray_link_table_covers_exactly_the_solid_sources():
    p, seq, pos, mask, wall = box_case()
    links = bfref.ray_links(p, pos, seq.triangles, seq.half_thickness, mask, wall).reshape(19, -1)
    nx, ny, nz = p.nx, p.ny, p.nz
    m3 = mask.reshape(nz, ny, nx)
    for q in range(1, 19):
        cx, cy, cz = bfref.C[q]
        src = np.zeros_like(m3)
        src[max(cz, 0):nz + min(cz, 0), max(cy, 0):ny + min(cy, 0), max(cx, 0):nx + min(cx, 0)] = \
            m3[max(-cz, 0):nz + min(-cz, 0), max(-cy, 0):ny + min(-cy, 0), max(-cx, 0):nx + min(-cx, 0)]
        want = (src.ravel() != 0) & (mask == 0)
        have = links[q] != 0
        assert np.array_equal(want, have)
        assert (links[q][have] >= 1e-4).all() and (links[q][have] <= 1.0).all()
    assert links[0].max() == 0.0


def test_the_reference_exact_roots(golden_dir):
    d = np.load(os.path.join(golden_dir, "deetjen_subset.npz"))
    ref = json.load(open(os.path.join(golden_dir, "deetjen_ray_roots.json")))
    frames = {int(f): d["points"][i] for i, f in enumerate(d["frames"])}
    t = np.float32(ref["timeSeconds"])
    times = d["times"][26:28]
    first, _, _ = bfref.frame_interval(times, t)
    pos, _ = bfref.prepare_vertices(np.stack([frames[26], frames[27]]), times, first, t)
    tri = d["triangles"]
    h = np.float32(ref["halfThicknessMeters"])
    worst = 0.0
    for case, cd in ref["cases"].items():
        o = np.asarray(cd["origin"], np.float32)
        dx = np.float32(cd["cellSize"])
        for link in cd["links"]:
            q = link["directionIndex"]
            c = np.asarray(bfref.C[q], np.float32)
            S = (o + (np.asarray(link["solidCell"], np.float32) + np.float32(0.5)) * dx).astype(np.float32)
            assert np.array_equal(np.asarray(link["fluidCell"]) - np.asarray(link["solidCell"]), bfref.C[q])
            F = (S + c * dx).astype(np.float32)
            D = (-c * dx).astype(np.float32)
            got = min(bfref.ray_shell(F, D, pos[k[0]], pos[k[1]], pos[k[2]], h) for k in tri)
            want = link["exactGlobalFluidToIntersectionFraction"]
            worst = max(worst, abs(got - want))
            assert got == pytest.approx(want, abs=1e-4), (case, link["solidCell"], q)
            # the signed-distance ratio the production path uses is far off
            # for these links (they were selected as its outliers)
    assert worst > 0.0


def test_fp64_shell_matches_closed_forms_and_the_fp32_shell():
    # bfo_ray_shell_f64 (the reference for differences of the exact roots):
    # the closed forms above, and the FP32 shell on random rays to its
    # rounding (grazing rays, where the two may disagree on a hit, skipped)
    f64 = lambda F, D, tri=TRI, h=0.5: float(bfref.ray_shell_f64(F, D, tri[0], tri[1], tri[2], h)[0])
    assert f64((0.2, 0.2, 3.0), (0.0, 0.0, -4.0)) == pytest.approx((3.0 - 0.5) / 4.0, abs=1e-15)
    assert f64((0.5, -3.0, 0.0), (0.0, 4.0, 0.0)) == pytest.approx((3.0 - 0.5) / 4.0, abs=1e-15)
    assert f64((3.0, 0.0, 0.0), (-4.0, 0.0, 0.0)) == pytest.approx((3.0 - 1.5) / 4.0, abs=1e-15)
    assert f64((-1.0, 0.2, 0.6), (3.0, 0.0, 0.0)) == 2.0
    assert f64((0.2, 0.2, 0.3), (0.0, 0.0, 1.0)) == 0.0
    rng = np.random.default_rng(19)
    compared = 0
    for _ in range(600):
        tri = tuple(rng.uniform(-1.0, 1.0, 3).astype(np.float32) for _ in range(3))
        h = np.float32(rng.uniform(0.1, 0.6))
        F = rng.uniform(-2.0, 2.0, 3).astype(np.float32)
        D = ((rng.uniform(-1.0, 1.0, 3) - F) * 1.2).astype(np.float32)
        t32 = bfref.ray_shell(F, D, *tri, h)
        t64 = f64(F.astype(np.float64), D.astype(np.float64), tuple(v.astype(np.float64) for v in tri), float(h))
        if (t32 <= 1.0) != (t64 <= 1.0):
            continue
        if t32 <= 1.0:
            assert abs(t32 - t64) <= 3e-5
            compared += 1
    assert compared > 150


def test_entry_triangles_attain_the_table():
    # bfo_ray_links_winners: for every link with an entry, its triangle's
    # shell gives exactly the table's fraction; links without an entry (the
    # signed-distance fallback) and non-links report none
    p, seq, pos, mask, wall = box_case()
    links, win = bfref.ray_links(p, pos, seq.triangles, seq.half_thickness, mask, wall, winners=True)
    assert np.array_equal(links, bfref.ray_links(p, pos, seq.triangles, seq.half_thickness, mask, wall))
    n = p.nx * p.ny * p.nz
    have = np.nonzero(win != 0xFFFFFFFF)[0]
    assert have.size > 100
    assert (links[have] != 0).all()
    assert (win[links == 0] == 0xFFFFFFFF).all()
    for i in have[::7]:
        q, g = divmod(int(i), n)
        x, y, z = g % p.nx, (g // p.nx) % p.ny, g // (p.nx * p.ny)
        c = np.asarray(bfref.C[q])
        S = np.array([x, y, z]) - c                       # far-field faces: no wrap in this box case
        F = ((S + 0.5) + c).astype(np.float32)
        D = (-c).astype(np.float32)
        a, b, cc = (pos[k] for k in seq.triangles[win[i]])
        t = bfref.ray_shell(F, D, a, b, cc, np.float32(seq.half_thickness))
        assert np.float32(min(max(t, 1e-4), 1.0)) == links[i]
