"""Pins of the oracle's splined pair potential (fast-CMD path,
forces.cpp:452-466, 634-691) and of its restatement of Boost.Math's
cubic_b_spline (uniform cubic B-spline interpolation, endpoint slopes from
the 5-point one-sided difference).  Boost is not in this image and the
reference ships no spline files; what fixes the restatement:
  * a cubic polynomial is reproduced exactly (value and slope), since the
    5-point slope is exact for it and the clamped cubic spline then is the
    polynomial;
  * the 5-point end slope is exact for quartics;
  * the spline interpolates its knots;
  * the interpolation error falls as h^4 for a smooth function;
  * the splined forces are minus the gradient of the splined energy.
"""
import numpy as np
import pytest

import spline_tables
from qtip4pf_water_amd import config as cfg


def test_spline_reproduces_cubic(oracle_mod):
    h, n, a = 0.05, 60, 0.0
    c = [0.3, -1.2, 0.7, 0.25]
    f = lambda x: c[0] + c[1] * x + c[2] * x ** 2 + c[3] * x ** 3
    df = lambda x: c[1] + 2 * c[2] * x + 3 * c[3] * x ** 2
    xs = a + h * np.arange(n)
    beta = oracle_mod.spline_fit(f(xs), h)
    for x in np.random.default_rng(1).uniform(a, xs[-1], 200):
        v, d = oracle_mod.spline_eval(beta, n, a, h, x)
        assert v == pytest.approx(f(x), abs=1e-11)
        assert d == pytest.approx(df(x), abs=1e-10)


def test_end_slopes_exact_for_quartic(oracle_mod):
    h, n = 0.1, 30
    f = lambda x: 1.0 - x + 0.5 * x ** 2 - 0.2 * x ** 3 + 0.05 * x ** 4
    df = lambda x: -1.0 + x - 0.6 * x ** 2 + 0.2 * x ** 3
    xs = h * np.arange(n)
    beta = oracle_mod.spline_fit(f(xs), h)
    _, d0 = oracle_mod.spline_eval(beta, n, 0.0, h, 0.0)
    _, d1 = oracle_mod.spline_eval(beta, n, 0.0, h, xs[-1])
    assert d0 == pytest.approx(df(0.0), abs=1e-10)
    assert d1 == pytest.approx(df(xs[-1]), abs=1e-9)


def test_spline_interpolates_knots_and_converges_h4(oracle_mod):
    g = lambda x: np.exp(-x) * np.sin(3 * x)
    errs = []
    for h in (0.04, 0.02):
        n = int(round(4.0 / h)) + 1
        xs = h * np.arange(n)
        beta = oracle_mod.spline_fit(g(xs), h)
        for x in xs[::7]:
            assert oracle_mod.spline_eval(beta, n, 0.0, h, x)[0] == pytest.approx(g(x), abs=1e-14)
        mids = xs[:-1] + 0.5 * h
        errs.append(max(abs(oracle_mod.spline_eval(beta, n, 0.0, h, x)[0] - g(x)) for x in mids))
    assert 12.0 < errs[0] / errs[1] < 20.0


def test_wspline_forces_are_minus_gradient(oracle_mod):
    oo, oh, hh = spline_tables.tables()
    oracle_mod.spline_set(oo, oh, hh, spline_tables.START, spline_tables.STEP)
    q0, box = cfg.lattice_water_box(3, 9.4, seed=7, jitter_bohr=0.1)
    o = oracle_mod.Oracle(27, 1, box)
    point, lst, _ = o.all_lists(q0)
    F0 = np.zeros(o.n_dof)
    e_plain = o.get_forces(q0, F0, point, lst)
    o.use_spline = True
    F = np.zeros(o.n_dof)
    e = o.get_forces(q0, F, point, lst)
    assert abs(e - e_plain) > 1e-4          # the tables matter at this size
    assert np.abs(F - F0).max() > 1e-4
    h = 1e-5
    for k in (0, 4, 8, 40, 130, 200):
        qp, qm = q0.copy(), q0.copy()
        qp[k] += h
        qm[k] -= h
        ep = o.get_forces(qp, np.zeros(o.n_dof), point, lst)
        em = o.get_forces(qm, np.zeros(o.n_dof), point, lst)
        assert -(ep - em) / (2 * h) == pytest.approx(F[k], rel=2e-6, abs=2e-9)
