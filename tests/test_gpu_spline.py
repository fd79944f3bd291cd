"""GPU fast-CMD path: the splined pair potential (forces.cpp:634-691) in the
*_wspline calls and the resident mode, against the oracle with the same
tables (tests/spline_tables.py; the reference ships none).  Tolerances as
for the plain path: per-atom forces 1e-10, energies 1e-12 relative."""
import sys

import numpy as np
import pytest

import spline_tables
from qtip4pf_water_amd import config as cfg

from energy_tol import term_scale, trajectory_bound

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def rel_atom_err(F, Fr):
    d = np.linalg.norm((F - Fr).reshape(-1, 3), axis=1)
    n = np.linalg.norm(Fr.reshape(-1, 3), axis=1)
    return float(np.max(d / np.maximum(n, 1e-300)))


def system(n_side, L, P, seed, jitter=0.05, noise=0.03):
    q0, box = cfg.lattice_water_box(n_side, L, seed=seed, jitter_bohr=jitter)
    q = cfg.replicate_beads(q0, P)
    if P > 1:
        q = q + np.random.default_rng(seed + 100).normal(0, noise, q.shape)
    return np.ascontiguousarray(q), box


@pytest.fixture(scope="module")
def tabs(oracle_mod):
    t = spline_tables.tables()
    oracle_mod.spline_set(*t, spline_tables.START, spline_tables.STEP)
    return t


def _water(W, n, P, box, tabs, **kw):
    w = W(n, P, box, **kw)
    w.set_splines(*tabs, spline_tables.START, spline_tables.STEP)
    return w


@pytest.mark.parametrize("n_side,L,P,seed", [(6, 18.64, 1, 11), (3, 9.4, 2, 12)])
def test_get_forces_wspline(oracle_mod, tabs, n_side, L, P, seed):
    from qtip4pf_water_amd import Water
    say("get_forces_wspline", n_side ** 3, "P", P)
    q, box = system(n_side, L, P, seed)
    o = oracle_mod.Oracle(n_side ** 3, P, box)
    point, lst, _ = o.all_lists(q)
    with _water(Water, n_side ** 3, P, box, tabs) as w:
        for b in range(P):
            o.use_spline = True
            Fr = np.zeros(P * o.n_dof)
            er = o.get_forces(q, Fr, point, lst, b)
            F = np.zeros(P * o.n_dof)
            e = w.get_forces_wspline(q, F, o.charge, point, lst, b)
            sl = slice(b * o.n_dof, (b + 1) * o.n_dof)
            assert rel_atom_err(F[sl], Fr[sl]) < 1e-10
            assert abs(e - er) <= 1e-12 * abs(er)
            # the plain call ignores the loaded tables
            o.use_spline = False
            Fr0 = np.zeros(P * o.n_dof)
            er0 = o.get_forces(q, Fr0, point, lst, b)
            F0 = np.zeros(P * o.n_dof)
            e0 = w.get_forces_nospline(q, F0, o.charge, point, lst, b)
            assert rel_atom_err(F0[sl], Fr0[sl]) < 1e-10
            assert abs(e0 - er0) <= 1e-12 * abs(er0)
            assert abs(e - e0) > 1e-5


def test_force_eval_sc_wspline(oracle_mod, tabs):
    from qtip4pf_water_amd import Water
    P = 4
    say("force_eval_sc_wspline 216 P", P)
    q, box = system(6, 18.64, P, 21)
    o = oracle_mod.Oracle(216, P, box)
    o.use_spline = True
    point, lst, _ = o.all_lists(q)
    Fr, fr = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
    er, scr = o.force_eval_sc(q, Fr, fr, point, lst)
    with _water(Water, 216, P, box, tabs) as w:
        F, f = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
        e, sc = w.force_eval_sc_wspline(q, F, o.charge, o.mass, point, lst, fphys=f)
    assert rel_atom_err(f, fr) < 1e-10
    assert abs(e - er) <= 1e-12 * abs(er)
    assert abs(sc - scr) <= 1e-12 * abs(scr)
    for b in range(P):
        sl = slice(b * o.n_dof, (b + 1) * o.n_dof)
        assert rel_atom_err(F[sl], Fr[sl]) < (1e-10 if b % 2 == 0 else 1e-9), b


def test_update_sc_wspline_steps(oracle_mod, tabs):
    from qtip4pf_water_amd import HostRng, Water
    P = 4
    say("update_sc_wspline 216 P", P)
    q0, box = system(6, 18.64, P, 31)
    o = oracle_mod.Oracle(216, P, box)
    o.use_spline = True
    p0 = cfg.sample_momenta(216, P, o.sys.beta_n, seed=31)
    point, lst, disp = o.all_lists(q0)
    F, fph = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
    o.force_eval_sc(q0, F, fph, point, lst)
    ref = [q0.copy(), p0.copy(), F.copy(), fph.copy(), point.copy(), lst.copy(), disp.copy()]
    gpu = [a.copy() for a in ref]
    rng = HostRng(9)
    hr = hg = 0.0
    with _water(Water, 216, P, box, tabs) as w:
        for step in range(4):
            nvt = step < 2
            dev = rng.draw_pile(o.n_atoms, P) if nvt else None
            er, hr, scr, _ = o.update_sc(*ref[:4], ref[4], ref[5], ref[6], nvt, dev, hr)
            eg, hg, scg = w.update_sc_wspline(gpu[0], gpu[1], o.mass, gpu[2], gpu[3], o.charge,
                                              gpu[4], gpu[5], gpu[6], dev, nvt, hg)
            assert np.max(np.abs(gpu[0] - ref[0])) <= 1e-12 * np.max(np.abs(ref[0])), step
            assert rel_atom_err(gpu[3], ref[3]) < 1e-9, step
            assert abs(eg - er) <= 1e-11 * abs(er), step


def test_update_trajectory_class_wspline_steps(oracle_mod, tabs):
    from qtip4pf_water_amd import Water
    say("update_trajectory_class_wspline 216")
    q0, box = system(6, 18.64, 1, 41)
    o = oracle_mod.Oracle(216, 1, box)
    o.use_spline = True
    p0 = cfg.sample_momenta(216, 1, o.sys.beta, seed=41)
    point, lst, disp = o.all_lists(q0)
    F = np.zeros(o.n_dof)
    o.get_forces(q0, F, point, lst, 0)
    ref = [q0.copy(), p0.copy(), F.copy(), point.copy(), lst.copy(), disp.copy()]
    gpu = [a.copy() for a in ref]
    S = term_scale(o, q0, point, lst)
    with _water(Water, 216, 1, box, tabs) as w, oracle_mod.exact_energy():
        for step in range(6):
            er, _ = o.update_class(*ref)
            eg = w.update_trajectory_class_wspline(gpu[0], gpu[1], o.mass, gpu[2], o.charge,
                                                   gpu[3], gpu[4], gpu[5])
            assert np.max(np.abs(gpu[0] - ref[0])) <= 1e-12 * np.max(np.abs(ref[0])), step
            assert rel_atom_err(gpu[2], ref[2]) < 1e-9, step
            tol = trajectory_bound(er, ref[2], gpu[0] - ref[0], scale=S)
            assert abs(eg - er) <= tol, (step, eg - er, tol)


def test_resident_run_with_splines_matches_drop_in(oracle_mod, tabs):
    from qtip4pf_water_amd import Water
    P = 4
    say("resident wspline NVE 216 P", P)
    q0, box = system(6, 18.64, P, 51)
    o = oracle_mod.Oracle(216, P, box)
    p0 = cfg.sample_momenta(216, P, o.sys.beta_n, seed=51)
    with _water(Water, 216, P, box, tabs) as w:
        w.set_state(q0, p0)
        w.run(4, is_nvt=False)
        w.sync()
        qr = w.get_state()[0]
        e_res = w.scalars()["energy"]
        point, lst, disp = o.all_lists(q0)
        F, fph = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
        w.force_eval_sc_wspline(q0, F, o.charge, o.mass, point, lst, fphys=fph)
        q, p = q0.copy(), p0.copy()
        for _ in range(4):
            e, _, _ = w.update_sc_wspline(q, p, o.mass, F, fph, o.charge, point, lst, disp,
                                          None, False, 0.0)
    assert np.max(np.abs(qr - q)) <= 1e-12 * np.max(np.abs(q))
    assert abs(e_res - e) <= 1e-11 * abs(e)


def test_wspline_without_tables_is_an_error():
    from qtip4pf_water_amd import QtpError, Water
    q, box = system(2, 6.3, 1, 61)
    with Water(8, 1, box) as w:
        point, lst, _ = w.new_lists()
        w.neighbour_list(q, point, lst, None, 0)
        with pytest.raises(QtpError):
            w.get_forces_wspline(q, np.zeros(q.size), None, point, lst, 0)
