"""CPU checks of the hand-written adjoint math (FP64 instantiations of the
device code) against central finite differences of its forward twin:
collision (TRT, TRT with the symmetric limiter, regularized, RR3 with the
positivity limiter), the moving-wall bounce-back value in every branch, and
the equilibrium's velocity derivative.  The forward twin of the symmetric
limiter is itself checked against the FP64 oracle's collision."""
import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "numi-gpu-physics_amd", "birdflow_hip", "libbirdflow_adjoint_check.so")

TRT, SYM, REG, RR3 = 0, 1, 2, 3
CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("adjoint check library not built")
    l = ctypes.CDLL(LIB)
    d = ctypes.POINTER(ctypes.c_double)
    l.bfa_collide_f64.argtypes = [ctypes.c_int, d, ctypes.c_double, ctypes.c_double, d, ctypes.POINTER(ctypes.c_int)]
    l.bfa_collide_adjoint_f64.argtypes = [ctypes.c_int, d, ctypes.c_double, ctypes.c_double, d, d]
    l.bfa_link_f64.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, d, d, d, d]
    l.bfa_feq_velocity_adjoint_f64.argtypes = [d, d, d]
    l.bfa_link_table_f64.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, d, d, d, ctypes.c_double, d]
    return l


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def collide(lib, coll, f, op, om):
    f = np.ascontiguousarray(f, np.float64)
    post = np.zeros(19)
    lim = ctypes.c_int(0)
    assert lib.bfa_collide_f64(coll, _p(f), op, om, _p(post), ctypes.byref(lim)) == 0
    return post, lim.value


def equilibrium(rho, u):
    cu = CX * u[0] + CY * u[1] + CZ * u[2]
    return W * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * np.dot(u, u))


@pytest.mark.parametrize("coll", [TRT, SYM, REG, RR3])
@pytest.mark.parametrize("strong", [False, True])
def test_collision_adjoint_matches_finite_differences(lib, coll, strong):
    rng = np.random.default_rng(17 + coll + 10 * strong)
    op, om = 1.0 / 0.55, 1.0 / 1.1
    checked = limited_seen = 0
    for trial in range(40):
        u = rng.uniform(-0.08, 0.08, 3)
        amp = 0.9 if strong else 0.05
        f = equilibrium(1.0 + rng.uniform(-0.05, 0.05), u) * rng.uniform(1 - amp, 1 + amp, 19)
        lp = rng.normal(size=19)
        lam = np.zeros(19)
        assert lib.bfa_collide_adjoint_f64(coll, _p(f), op, om, _p(lp), _p(lam)) == 0
        delta = rng.normal(size=19) * f
        h = 1e-7
        pp, lim_p = collide(lib, coll, f + h * delta, op, om)
        pm, lim_m = collide(lib, coll, f - h * delta, op, om)
        _, lim0 = collide(lib, coll, f, op, om)
        if not (lim_p == lim_m == lim0):
            continue            # limiter switches inside the stencil: not differentiable there
        fd = np.dot(lp, (pp - pm) / (2 * h))
        ad = np.dot(lam, delta)
        assert abs(fd - ad) <= 1e-6 * max(1.0, abs(fd)), (trial, fd, ad)
        checked += 1
        limited_seen += lim0
    assert checked >= 30
    if strong and coll != TRT:
        assert limited_seen >= 3


def test_symmetric_limiter_twin_matches_the_oracle(lib):
    # the adjoint's forward twin of the symmetric-limiter TRT against the
    # FP64 oracle's collision of one cell (a 1 x 1 x 1 periodic step), both
    # when the limiter binds and when it does not
    import bfref
    rng = np.random.default_rng(5)
    op, om = 1.95, 1.2
    limited = 0
    for _ in range(200):
        rho, u = 1.0 + 0.05 * rng.standard_normal(), 0.1 * rng.standard_normal(3)
        f = np.abs(equilibrium(rho, u) * (1.0 + 0.8 * rng.standard_normal(19)))
        post, lim = collide(lib, SYM, f, op, om)
        p = bfref.make_params((1, 1, 1), omega_plus=op, omega_minus=om,
                              collision=bfref.COLLISION_TRT_SYMMETRIC_LIMITER, accumulate_loads=True)
        want, loads, *_ = bfref.step(p, f.copy())
        # the oracle keeps its rates in FP32: compare at their rounding
        assert np.abs(post - want).max() <= 1e-6 * np.abs(want).max(), (post, want)
        assert lim == int(loads[6] != 0.0)
        limited += lim
    assert limited >= 20


WALL_HALFWAY, WALL_SDF, WALL_HALF = 0, 1, 2


def test_link_table_fraction_derivative(lib):
    # the link-table rule (the exact-root fractions): d value / d lf against
    # central differences in lf, both Bouzidi branches, with and without the
    # far side (halfway fallback below 1/2 without it: no dependence)
    rng = np.random.default_rng(8)
    seen = set()
    for trial in range(200):
        q = int(rng.integers(1, 19))
        ok = int(rng.integers(0, 2))
        x = np.ascontiguousarray(rng.uniform(0.01, 0.1, 3))
        wg = np.ascontiguousarray(np.concatenate([rng.uniform(-0.05, 0.05, 3), [0.5]]))
        ws = np.ascontiguousarray(np.concatenate([rng.uniform(-0.05, 0.05, 3), [-0.5]]))
        lf = float(rng.uniform(0.02, 0.98))

        def val(l):
            o = np.zeros(3)
            lib.bfa_link_table_f64(q, ok, 1.0, _p(x), _p(wg), _p(ws), l, _p(o))
            return o

        o = val(lf)
        assert o[2] == (0.5 if (not ok and lf <= 0.5) else lf)      # halfway fallback reports 1/2
        h = 1e-7
        fd = (val(lf + h)[0] - val(lf - h)[0]) / (2 * h)
        assert abs(fd - o[1]) <= 1e-6 * max(1.0, abs(fd)), (q, ok, lf, fd, o[1])
        if not ok and lf <= 0.5:
            assert o[1] == 0.0
        seen.add((ok, lf <= 0.5))
    assert len(seen) == 4


@pytest.mark.parametrize("rule", [WALL_HALFWAY, WALL_SDF, WALL_HALF])
def test_link_value_derivatives_match_finite_differences(lib, rule):
    rng = np.random.default_rng(5 + rule)
    branches = set()
    for trial in range(200):
        q = int(rng.integers(1, 19))
        ok = int(rng.integers(0, 2))
        x = rng.uniform(0.01, 0.1, 3)                     # reflected, farther, pin
        wg = np.concatenate([rng.uniform(-0.05, 0.05, 3), rng.uniform(0.05, 1.0, 1)])
        ws = np.concatenate([rng.uniform(-0.05, 0.05, 3), rng.uniform(-1.0, -0.05, 1)])
        out = np.zeros(13)

        def val(x_, wg_, ws_):
            o = np.zeros(13)
            lib.bfa_link_f64(q, rule, ok, 1.0, _p(np.ascontiguousarray(x_)), _p(np.ascontiguousarray(wg_)),
                             _p(np.ascontiguousarray(ws_)), _p(o))
            return o

        out = val(x, wg, ws)
        lf = out[12]
        branches.add((ok, lf <= 0.5))
        h = 1e-7
        grads = [out[1], out[2], out[3]]
        for i in range(3):
            e = np.zeros(3)
            e[i] = h
            fd = (val(x + e, wg, ws)[0] - val(x - e, wg, ws)[0]) / (2 * h)
            assert abs(fd - grads[i]) <= 1e-6 * max(1, abs(fd)), (q, rule, ok, i, fd, grads[i])
        for i in range(4):
            e = np.zeros(4)
            e[i] = h
            a, b = val(x, wg + e, ws), val(x, wg - e, ws)
            if (a[12] <= 0.5) == (b[12] <= 0.5) == (lf <= 0.5):
                fd = (a[0] - b[0]) / (2 * h)
                assert abs(fd - out[4 + i]) <= 1e-6 * max(1, abs(fd)), (q, rule, ok, "wg", i, fd, out[4 + i])
            a, b = val(x, wg, ws + e), val(x, wg, ws - e)
            if (a[12] <= 0.5) == (b[12] <= 0.5) == (lf <= 0.5):
                fd = (a[0] - b[0]) / (2 * h)
                assert abs(fd - out[8 + i]) <= 1e-6 * max(1, abs(fd)), (q, rule, ok, "ws", i, fd, out[8 + i])
    if rule == WALL_SDF:
        assert {(1, True), (1, False), (0, True)} <= branches


def test_equilibrium_velocity_adjoint(lib):
    rng = np.random.default_rng(2)
    for _ in range(10):
        u = rng.uniform(-0.1, 0.1, 3)
        l = rng.normal(size=19)
        g = np.zeros(3)
        lib.bfa_feq_velocity_adjoint_f64(_p(l), _p(u), _p(g))
        h = 1e-6
        for i in range(3):
            e = np.zeros(3)
            e[i] = h
            fd = np.dot(l, (equilibrium(1.0, u + e) - equilibrium(1.0, u - e)) / (2 * h))
            assert abs(fd - g[i]) <= 1e-8 * max(1, abs(fd))
