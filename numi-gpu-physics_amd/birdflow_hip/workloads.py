"""Synthetic moving-boundary workloads as indexed-mesh frame sequences.

Both produce what the reference's indexed-surface replay consumes
(MeasuredBirdSurfaceSequence.swift:36-60): frame-major vertex positions,
frame times, a fixed triangle topology and per-triangle part ids
(1 body, 2 left wing, 3 right wing, 4 tail).  One frame per lattice step, so
the prepare kernel's piecewise-linear interpolation lands exactly on frames
and the wall velocity is the one-step finite difference of positions.

* flapping_plate — a rigid rectangular plate (span x chord x thickness,
  12 triangles: 2 per face of the box) heaving h0 sin(2 pi t/T) and pitching
  a0 sin(2 pi t/T + pi/2) about its mid-chord span axis.
* synthetic_bird — ellipsoid body, two tapered wings and a fan tail
  (about 40k triangles); per-wing stroke, deviation, pitch and twist are
  3-harmonic Fourier series whose coefficients come from a seed.
Lengths are in cells (world units = cells, dx = 1).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class MeshSequence:
    points: np.ndarray          # [frames, vertices, 3] float32
    times: np.ndarray           # [frames] float32
    triangles: np.ndarray       # [triangles, 3] uint32
    parts: np.ndarray           # [triangles] uint8
    half_thickness: float
    meta: dict = None

    @property
    def frames(self):
        return self.points.shape[0]


def _box(span, chord, thickness):
    """Closed box centred at the origin: x chord, y span, z thickness."""
    hx, hy, hz = chord / 2.0, span / 2.0, thickness / 2.0
    v = np.array([[sx * hx, sy * hy, sz * hz] for sz in (-1, 1) for sy in (-1, 1) for sx in (-1, 1)], np.float64)
    # vertex index = (sz>0)*4 + (sy>0)*2 + (sx>0)
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for a, b, c, d in quads:
        tris += [(a, b, c), (a, c, d)]
    return v, np.array(tris, np.uint32)


def _rot_y(angle):
    c, s = np.cos(angle), np.sin(angle)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def flapping_plate(steps: int, center=(64.0, 64.0, 64.0), span=48.0, chord=16.0, thickness=2.0,
                   heave=8.0, pitch_deg=30.0, period=800.0, half_thickness=0.5, t0=0) -> MeshSequence:
    """Config: plate 48 x 16 x 2 at x = 64, heave 8 sin(2 pi t/800), pitch 30 deg sin(2 pi t/800 + pi/2)."""
    rest, tri = _box(span, chord, thickness)
    times = np.arange(t0, t0 + steps + 1, dtype=np.float64)
    pts = np.empty((len(times), rest.shape[0], 3), np.float64)
    for i, t in enumerate(times):
        ph = 2 * math.pi * t / period
        r = _rot_y(math.radians(pitch_deg) * math.sin(ph + math.pi / 2))
        off = np.array([center[0], center[1], center[2] + heave * math.sin(ph)])
        pts[i] = rest @ r.T + off
    return MeshSequence(pts.astype(np.float32), times.astype(np.float32), tri, np.ones(len(tri), np.uint8),
                        float(half_thickness))


PLATE_PARAMETERS = ("heave_1", "heave_phase_1", "heave_2", "heave_phase_2",
                    "pitch_1", "pitch_phase_1", "pitch_2", "pitch_phase_2")


def parametric_plate(theta, steps: int, center=(64.0, 64.0, 64.0), span=48.0, chord=16.0, thickness=2.0,
                     period=800.0, half_thickness=0.5, t0=0):
    """Flapping plate with two-harmonic kinematics and its parameter tangents.

    theta = (H1, phi1, H2, phi2, P1, psi1, P2, psi2) (cells, radians):
    heave(t) = H1 sin(w t + phi1) + H2 sin(2 w t + phi2),
    pitch(t) = P1 sin(w t + psi1) + P2 sin(2 w t + psi2), w = 2 pi / period;
    vertex p(t) = R_y(pitch(t)) p_rest + center + (0, 0, heave(t)).
    Returns (MeshSequence with FP32 frames, FP64 frames [F, V, 3], tangents
    d frames / d theta [8, F, V, 3] in FP64)."""
    th = np.asarray(theta, np.float64)
    rest, tri = _box(span, chord, thickness)
    times = np.arange(t0, t0 + steps + 1, dtype=np.float64)
    F, V = len(times), rest.shape[0]
    pts = np.empty((F, V, 3))
    tan = np.zeros((8, F, V, 3))
    w = 2 * math.pi / period
    for i, t in enumerate(times):
        a1, a2 = w * t + th[1], 2 * w * t + th[3]
        b1, b2 = w * t + th[5], 2 * w * t + th[7]
        heave = th[0] * math.sin(a1) + th[2] * math.sin(a2)
        pitch = th[4] * math.sin(b1) + th[6] * math.sin(b2)
        c, s_ = math.cos(pitch), math.sin(pitch)
        r = np.array([[c, 0, s_], [0, 1, 0], [-s_, 0, c]])
        dr = np.array([[-s_, 0, c], [0, 0, 0], [-c, 0, -s_]])      # d R_y / d pitch
        pts[i] = rest @ r.T + np.array([center[0], center[1], center[2] + heave])
        dpitch = rest @ dr.T
        tan[0, i, :, 2] = math.sin(a1)
        tan[1, i, :, 2] = th[0] * math.cos(a1)
        tan[2, i, :, 2] = math.sin(a2)
        tan[3, i, :, 2] = th[2] * math.cos(a2)
        tan[4, i] = dpitch * math.sin(b1)
        tan[5, i] = dpitch * th[4] * math.cos(b1)
        tan[6, i] = dpitch * math.sin(b2)
        tan[7, i] = dpitch * th[6] * math.cos(b2)
    seq = MeshSequence(pts.astype(np.float32), times.astype(np.float32), tri, np.ones(len(tri), np.uint8),
                       float(half_thickness), meta={"theta": th.tolist(), "parameters": PLATE_PARAMETERS})
    return seq, pts, tan


def _grid_sheet(nu, nv):
    """Triangulated (nu x nv)-quad sheet; returns (u, v) in [0,1]^2 and triangles."""
    u, v = np.meshgrid(np.linspace(0, 1, nu + 1), np.linspace(0, 1, nv + 1), indexing="ij")
    idx = np.arange((nu + 1) * (nv + 1)).reshape(nu + 1, nv + 1)
    a, b = idx[:-1, :-1].ravel(), idx[1:, :-1].ravel()
    c, d = idx[1:, 1:].ravel(), idx[:-1, 1:].ravel()
    tri = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
    return u.ravel(), v.ravel(), tri.astype(np.uint32)


def _rot(axis, angle):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    x, y, z = axis
    c, s = math.cos(angle), math.sin(angle)
    C = 1 - c
    return np.array([[c + x * x * C, x * y * C - z * s, x * z * C + y * s],
                     [y * x * C + z * s, c + y * y * C, y * z * C - x * s],
                     [z * x * C - y * s, z * y * C + x * s, c + z * z * C]])


@dataclass
class BirdKinematics:
    """3-harmonic Fourier coefficients (radians) per wing and angle."""

    period: float
    stroke: np.ndarray      # [2 wings, 3 harmonics, 2 (sin, cos)]
    deviation: np.ndarray
    pitch: np.ndarray
    twist: np.ndarray

    @classmethod
    def from_seed(cls, seed: int, period: float = 1200.0):
        rng = np.random.default_rng(seed)

        def series(max_deg):
            # amplitudes decay with harmonic; the total stays within max_deg
            a = rng.uniform(-1.0, 1.0, size=(2, 3, 2)) * np.array([1.0, 0.35, 0.15])[None, :, None]
            a *= math.radians(max_deg) / max(np.abs(a).sum(axis=(1, 2)).max(), 1e-12)
            return a

        return cls(period, series(60.0), series(20.0), series(20.0), series(20.0))

    def angle(self, coeff, wing, t):
        ph = 2 * math.pi * t / self.period
        k = np.arange(1, 4)
        return float((coeff[wing, :, 0] * np.sin(k * ph) + coeff[wing, :, 1] * np.cos(k * ph)).sum())


SOUND_SPEED = math.sqrt(1.0 / 3.0)


class SyntheticBird:
    """Rest geometry, topology and Fourier kinematics of the synthetic bird."""

    def __init__(self, seed=7, center=(160.0, 128.0, 128.0), body_radii=(40.0, 14.0, 14.0), half_span=100.0,
                 root_chord=40.0, tip_chord=12.0, tail=(30.0, 20.0), period=1200.0, body_segments=(96, 64),
                 wing_segments=(160, 40), tail_segments=(48, 24)):
        self.kin = BirdKinematics.from_seed(seed, period)
        self.center = np.asarray(center, np.float64)
        # body: ellipsoid with single pole vertices (no zero-area triangles)
        nu, nv = body_segments
        th = np.linspace(0, math.pi, nv + 1)[1:-1]
        ph = np.linspace(0, 2 * math.pi, nu, endpoint=False)
        T, P = np.meshgrid(th, ph, indexing="ij")
        rings = np.stack([body_radii[0] * np.cos(T), body_radii[1] * np.sin(T) * np.cos(P),
                          body_radii[2] * np.sin(T) * np.sin(P)], -1).reshape(-1, 3)
        self.body = np.concatenate([[[body_radii[0], 0.0, 0.0]], rings, [[-body_radii[0], 0.0, 0.0]]])
        nr = nv - 1
        ring = lambda i, j: 1 + i * nu + (j % nu)
        tri = []
        for j in range(nu):
            tri.append((0, ring(0, j + 1), ring(0, j)))
            tri.append((len(self.body) - 1, ring(nr - 1, j), ring(nr - 1, j + 1)))
        for i in range(nr - 1):
            for j in range(nu):
                a, b, c, d = ring(i, j), ring(i, j + 1), ring(i + 1, j + 1), ring(i + 1, j)
                tri += [(a, b, c), (a, c, d)]
        body_tri = np.asarray(tri, np.uint32)
        su, sv, wing_tri = _grid_sheet(*wing_segments)
        self.span_fraction = su
        chord_len = root_chord + (tip_chord - root_chord) * su
        self.wing_local = np.stack([(sv - 0.25) * chord_len, su * half_span, np.zeros_like(su)], -1)
        self.root = np.array([0.0, body_radii[1] * 0.8, body_radii[2] * 0.3])
        tu, tv, tail_tri = _grid_sheet(*tail_segments)
        fan = (tu - 0.5) * math.radians(60.0)
        rad = 2.0 + tv * (tail[0] - 2.0)
        width_scale = tail[1] / (2 * tail[0] * math.sin(math.radians(30.0)))
        self.tail_local = np.stack([-body_radii[0] * 0.9 - rad * np.cos(fan), rad * np.sin(fan) * width_scale * 2,
                                    np.zeros_like(tu)], -1)
        nb, nw = len(self.body), len(self.wing_local)
        self.triangles = np.concatenate([body_tri, wing_tri + nb, wing_tri + nb + nw,
                                         tail_tri + nb + 2 * nw]).astype(np.uint32)
        self.parts = np.concatenate([np.full(len(body_tri), 1), np.full(len(wing_tri), 2),
                                     np.full(len(wing_tri), 3), np.full(len(tail_tri), 4)]).astype(np.uint8)
        self.amplitude_scale = 1.0

    def wing(self, w, side, t):
        k, s = self.kin, self.amplitude_scale
        stroke = s * k.angle(k.stroke, w, t)
        dev = s * k.angle(k.deviation, w, t)
        pit = s * k.angle(k.pitch, w, t)
        tw = s * k.angle(k.twist, w, t)
        loc = self.wing_local.copy()
        loc[:, 1] *= side
        # pitch plus a twist growing linearly with span, about the span axis
        ang = pit + tw * self.span_fraction
        cth, sth = np.cos(ang), np.sin(ang)
        x0, z0 = loc[:, 0].copy(), loc[:, 2].copy()
        loc[:, 0] = cth * x0 + sth * z0
        loc[:, 2] = -sth * x0 + cth * z0
        r = _rot((1.0, 0.0, 0.0), side * stroke) @ _rot((0.0, 0.0, 1.0), side * dev)
        return loc @ r.T + self.root * np.array([1.0, side, 1.0])

    def frame(self, t):
        parts = [self.body, self.wing(0, 1.0, t), self.wing(1, -1.0, t), self.tail_local]
        return np.concatenate(parts) + self.center

    def articulated(self, half_thickness: float = 1.0):
        """The same bird as an articulated mesh (on-device kinematics): the
        body and tail rigid (joint 0), the wings on joints 1 (+y) and 2 (-y)
        with this bird's Fourier series; positions agree with frame(t) to
        FP32 rounding."""
        from .simulation import ArticulatedMesh, Joint
        nb, nw, nt = len(self.body), len(self.wing_local), len(self.tail_local)
        rest = np.concatenate([self.body, self.wing_local, self.wing_local, self.tail_local]).astype(np.float32)
        joint = np.concatenate([np.zeros(nb), np.ones(nw), np.full(nw, 2), np.zeros(nt)]).astype(np.uint8)
        span = np.concatenate([np.zeros(nb), self.span_fraction, self.span_fraction, np.zeros(nt)]).astype(np.float32)
        k = self.kin
        joints = [Joint(side, self.root.tolist(), k.stroke[w], k.deviation[w], k.pitch[w], k.twist[w])
                  for w, side in ((0, 1.0), (1, -1.0))]
        return ArticulatedMesh(rest, joint, span, self.triangles, self.parts, float(half_thickness), float(k.period),
                               float(self.amplitude_scale), self.center.tolist(), joints, harmonics=3)

    def max_wing_speed(self, samples=240):
        """Largest wing-vertex speed over one period (central differences)."""
        period = self.kin.period
        best = 0.0
        for t in np.linspace(0.0, period, samples, endpoint=False):
            for w, side in ((0, 1.0), (1, -1.0)):
                v = self.wing(w, side, t + 0.5) - self.wing(w, side, t - 0.5)
                best = max(best, float(np.sqrt((v * v).sum(axis=1)).max()))
        return best

    def limit_mach(self, relative_translation, mach=0.15):
        """Scale every drawn angle amplitude by one factor so that
        (relative translation + wing speed) / c_s <= mach, the guard of
        BirdModel.swift:345-362."""
        limit = mach * SOUND_SPEED - relative_translation
        for _ in range(4):
            speed = self.max_wing_speed()
            if speed <= limit * (1 - 1e-6):
                break
            self.amplitude_scale *= limit / speed * 0.999
        return self.amplitude_scale


def synthetic_bird(steps: int, seed: int = 7, center=(160.0, 128.0, 128.0), inflow: float = 0.04,
                   mach: float = 0.15, half_thickness: float = 1.0, t0: int = 0, **geometry) -> MeshSequence:
    """Synthetic articulated bird of about 40k triangles (512 x 256 x 256 domain)
    with one frame per step.  Amplitudes are drawn from the seed (stroke <= 60
    deg, deviation / pitch / twist <= 20 deg) and then scaled uniformly to the
    reference's lattice-Mach guard for a body held in the inflow."""
    bird = SyntheticBird(seed=seed, center=center, **geometry)
    scale = bird.limit_mach(abs(inflow), mach)
    times = np.arange(t0, t0 + steps + 1, dtype=np.float64)
    pts = np.empty((len(times), bird.frame(0.0).shape[0], 3), np.float32)
    for i, t in enumerate(times):
        pts[i] = bird.frame(t).astype(np.float32)
    return MeshSequence(pts, times.astype(np.float32), bird.triangles, bird.parts, float(half_thickness),
                        meta={"amplitude_scale": scale, "max_wing_speed": bird.max_wing_speed(),
                              "triangles": int(len(bird.triangles)), "vertices": int(pts.shape[1])})


def synthetic_bird_articulated(seed: int = 7, center=(160.0, 128.0, 128.0), inflow: float = 0.04, mach: float = 0.15,
                               half_thickness: float = 1.0, **geometry):
    """The synthetic bird of synthetic_bird() as an articulated mesh: the
    same rest geometry, Fourier series and Mach-limited amplitude scale, with
    the kinematics evaluated on the device at every step's time."""
    bird = SyntheticBird(seed=seed, center=center, **geometry)
    bird.limit_mach(abs(inflow), mach)
    return bird.articulated(half_thickness)
