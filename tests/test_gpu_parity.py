"""GPU parity: the HIP library against the CPU oracle on the same seeded inputs.

Tolerances (the targets of the north-star statement):
  forces        per atom |F_gpu - F_ref| <= 1e-10 |F_ref|
  energies      relative 1e-12
  lists         bit-exact (point and list arrays equal)
  normal modes  1e-12 relative to the array's scale
SC forces on odd beads contain a finite-difference Hessian term whose
rounding noise is amplified by 1/eps (forces.cpp:755-775); they are held to
1e-9 per atom.
"""
import sys

import numpy as np
import pytest

from qtip4pf_water_amd import config as cfg

from energy_tol import EPS, single, term_scale, trajectory_bound

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def rel_atom_err(F, Fr):
    d = np.linalg.norm((F - Fr).reshape(-1, 3), axis=1)
    n = np.linalg.norm(Fr.reshape(-1, 3), axis=1)
    return float(np.max(d / np.maximum(n, 1e-300)))


@pytest.fixture(scope="module")
def W():
    from qtip4pf_water_amd import Water
    return Water


def system(n_side, L_angs, P, seed, jitter=0.05, bead_noise=0.03):
    q0, box = cfg.lattice_water_box(n_side, L_angs, seed=seed, jitter_bohr=jitter)
    q = cfg.replicate_beads(q0, P)
    if P > 1 and bead_noise > 0:
        q = q + np.random.default_rng(seed + 100).normal(0, bead_noise, q.shape)
    return np.ascontiguousarray(q), box


CASES = [(6, 18.64, 1, 11), (2, 6.3, 2, 12), (8, 24.86, 2, 13), (3, 9.4, 4, 14)]


@pytest.mark.parametrize("n_side,L,P,seed", CASES)
def test_get_forces_nospline(W, oracle_mod, n_side, L, P, seed):
    say("get_forces_nospline", n_side ** 3, "mols P", P)
    q, box = system(n_side, L, P, seed)
    o = oracle_mod.Oracle(n_side ** 3, P, box)
    point, lst, disp = o.all_lists(q)
    with W(n_side ** 3, P, box) as w:
        assert w.derived["k_max"] == o.ewald()["k_max"]
        for b in range(P):
            Fr = np.zeros(P * o.n_dof)
            er = o.get_forces(q, Fr, point, lst, b)
            F = np.zeros(P * o.n_dof)
            e = w.get_forces_nospline(q, F, o.charge, point, lst, b)
            sl = slice(b * o.n_dof, (b + 1) * o.n_dof)
            assert rel_atom_err(F[sl], Fr[sl]) < 1e-10
            assert abs(e - er) <= 1e-12 * abs(er)
            # other beads' slices untouched (forces.cpp:481-485 zeroes only its own)
            assert np.all(np.delete(F, np.arange(sl.start, sl.stop)) == 0.0)


@pytest.mark.parametrize("n_side,L,P,seed", [(6, 18.64, 4, 21), (3, 9.4, 8, 22), (2, 6.3, 2, 23)])
def test_force_eval_sc_nospline(W, oracle_mod, n_side, L, P, seed):
    say("force_eval_sc_nospline", n_side ** 3, "mols P", P)
    q, box = system(n_side, L, P, seed)
    o = oracle_mod.Oracle(n_side ** 3, P, box)
    point, lst, disp = o.all_lists(q)
    Fr, fr = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
    er, scr = o.force_eval_sc(q, Fr, fr, point, lst)
    with W(n_side ** 3, P, box) as w:
        F, f = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
        e, sc = w.force_eval_sc_nospline(q, F, o.charge, o.mass, point, lst, fphys=f)
    assert rel_atom_err(f, fr) < 1e-10
    assert abs(e - er) <= 1e-12 * abs(er)
    assert abs(sc - scr) <= 1e-12 * abs(scr)
    for b in range(P):
        sl = slice(b * o.n_dof, (b + 1) * o.n_dof)
        tol = 1e-10 if b % 2 == 0 else 1e-9
        assert rel_atom_err(F[sl], Fr[sl]) < tol, b


@pytest.mark.parametrize("n_side,L,P,seed", [(6, 18.64, 2, 31), (8, 24.86, 1, 32), (3, 9.4, 2, 33)])
def test_neighbour_list_bit_exact(W, oracle_mod, n_side, L, P, seed):
    say("neighbour_list", n_side ** 3, "mols P", P)
    q, box = system(n_side, L, P, seed, jitter=0.4, bead_noise=0.2)
    o = oracle_mod.Oracle(n_side ** 3, P, box)
    pr, lr, dr = o.new_lists()
    dr[:] = 1.0
    with W(n_side ** 3, P, box) as w:
        pg, lg, dg = w.new_lists()
        dg[:] = 1.0
        for b in range(P):
            o.neighbour_list(q, pr, lr, dr, b)
            w.neighbour_list(q, pg, lg, dg, b)
    assert np.array_equal(pg, pr)
    assert np.array_equal(lg, lr)
    assert np.array_equal(dg, dr)


def test_neighbour_list_matches_reference_fixture(W):
    """Golden vector written by the reference's own neighbour.cpp."""
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "nlist_216.npz")
    g = np.load(path)
    box = tuple(g["boxl"])
    P = int(g["n_beads"])
    say("neighbour_list golden 216 P", P)
    with W(216, P, box) as w:
        pg, lg, dg = w.new_lists()
        for b in range(P):
            w.neighbour_list(g["q"], pg, lg, dg, b)
    assert np.array_equal(pg, g["point"])
    assert np.array_equal(lg, g["list"])


@pytest.mark.parametrize("P", [2, 8, 32])
def test_ftransform(W, oracle_mod, P):
    say("ftransform P", P)
    q, box = system(3, 9.4, P, 41)
    o = oracle_mod.Oracle(27, P, box)
    with W(27, P, box) as w:
        for length in (o.n_dof, 3 * o.n_mols):
            a = np.random.default_rng(P + length).normal(0, 5.0, P * length)
            ar, ag = a.copy(), a.copy()
            o.ftransform(ar, length, True)
            w.ftransform(ag, length, True)
            assert np.max(np.abs(ag - ar)) <= 1e-12 * np.max(np.abs(ar))
            o.ftransform(ar, length, False)
            w.ftransform(ag, length, False)
            assert np.max(np.abs(ag - ar)) <= 1e-12 * np.max(np.abs(ar))
            assert np.max(np.abs(ag - a)) <= 1e-12 * np.max(np.abs(a))


@pytest.mark.parametrize("P", [2, 8])
def test_pile_g_with_reference_draws(W, oracle_mod, P):
    from qtip4pf_water_amd import HostRng
    say("pile_g P", P)
    q, box = system(3, 9.4, P, 51)
    o = oracle_mod.Oracle(27, P, box)
    p = np.random.default_rng(P).normal(0, 20.0, P * o.n_dof)
    dev = HostRng(99 + P).draw_pile(o.n_atoms, P)
    pr, pg = p.copy(), p.copy()
    hr = o.pile_g(pr, dev)
    with W(27, P, box) as w:
        hg = w.pile_g(pg, o.mass, dev)
    assert np.max(np.abs(pg - pr)) <= 1e-13 * np.max(np.abs(pr))
    ke = 0.5 * np.sum(p.reshape(P, -1) ** 2 / o.mass)
    assert abs(hg - hr) <= 1e-12 * ke


def test_update_sc_nospline_steps(W, oracle_mod):
    """Three NVT steps (reference-order host draws) then two NVE steps."""
    from qtip4pf_water_amd import HostRng
    P = 4
    say("update_sc_nospline 216 P", P)
    q0, box = system(6, 18.64, P, 61)
    o = oracle_mod.Oracle(216, P, box)
    p0 = cfg.sample_momenta(216, P, o.sys.beta_n, seed=61)
    point, lst, disp = o.all_lists(q0)
    F, fph = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
    o.force_eval_sc(q0, F, fph, point, lst)
    ref = [q0.copy(), p0.copy(), F.copy(), fph.copy(), point.copy(), lst.copy(), disp.copy()]
    gpu = [a.copy() for a in ref]
    rng = HostRng(5)
    hr = hg = 0.0
    with W(216, P, box) as w:
        for step in range(5):
            nvt = step < 3
            dev = rng.draw_pile(o.n_atoms, P) if nvt else None
            er, hr, scr, _ = o.update_sc(*ref[:4], ref[4], ref[5], ref[6], nvt, dev, hr)
            eg, hg, scg = w.update_sc_nospline(gpu[0], gpu[1], o.mass, gpu[2], gpu[3], o.charge,
                                               gpu[4], gpu[5], gpu[6], dev, nvt, hg)
            qs = np.max(np.abs(ref[0]))
            assert np.max(np.abs(gpu[0] - ref[0])) <= 1e-12 * qs, step
            assert np.max(np.abs(gpu[1] - ref[1])) <= 1e-10 * np.max(np.abs(ref[1])), step
            assert rel_atom_err(gpu[3], ref[3]) < 1e-9, step
            assert abs(eg - er) <= 1e-11 * abs(er), step
            assert abs(scg - scr) <= 1e-11 * abs(scr), step
            assert np.array_equal(gpu[4], ref[4]) and np.array_equal(gpu[5], ref[5])


def test_update_trajectory_class_nospline_steps(W, oracle_mod):
    say("update_trajectory_class_nospline 216")
    q0, box = system(6, 18.64, 1, 71)
    o = oracle_mod.Oracle(216, 1, box)
    p0 = cfg.sample_momenta(216, 1, o.sys.beta, seed=71)
    point, lst, disp = o.all_lists(q0)
    F = np.zeros(o.n_dof)
    o.get_forces(q0, F, point, lst, 0)
    ref = [q0.copy(), p0.copy(), F.copy(), point.copy(), lst.copy(), disp.copy()]
    gpu = [a.copy() for a in ref]
    S = term_scale(o, q0, point, lst)
    with W(216, 1, box) as w, oracle_mod.exact_energy():
        for step in range(10):
            er, _ = o.update_class(*ref)
            eg = w.update_trajectory_class_nospline(gpu[0], gpu[1], o.mass, gpu[2], o.charge,
                                                    gpu[3], gpu[4], gpu[5])
            assert np.max(np.abs(gpu[0] - ref[0])) <= 1e-12 * np.max(np.abs(ref[0])), step
            assert rel_atom_err(gpu[2], ref[2]) < 1e-9, step
            # 80-bit oracle energy; 1e-12 |V| plus the first-order effect of
            # the measured position difference of the two trajectories
            tol = trajectory_bound(er, ref[2], gpu[0] - ref[0], scale=S)
            assert abs(eg - er) <= tol, (step, eg - er, tol)


def test_observables(W, oracle_mod):
    P = 8
    say("get_kcv / get_conserved P", P)
    q, box = system(3, 9.4, P, 81)
    o = oracle_mod.Oracle(27, P, box)
    p = cfg.sample_momenta(27, P, o.sys.beta_n, seed=81)
    f = np.random.default_rng(3).normal(0, 0.01, P * o.n_dof)
    kr = o.get_kcv(q, f)
    cr = o.get_conserved(q.copy(), p.copy(), -1.25)
    with W(27, P, box) as w:
        kg = w.get_kcv(q, f)
        cg = w.get_conserved(q, p, o.mass, -1.25)
    assert abs(kg - kr) <= 1e-12 * abs(kr)
    assert abs(cg - cr) <= 1e-12 * abs(cr)


def test_resident_run_matches_drop_in_nve(W, oracle_mod):
    """Resident graph replay == reference-shaped stepping (NVE, same kernels)."""
    P = 4
    say("resident run NVE 216 P", P)
    q0, box = system(6, 18.64, P, 91)
    o = oracle_mod.Oracle(216, P, box)
    p0 = cfg.sample_momenta(216, P, o.sys.beta_n, seed=91)
    with W(216, P, box) as w:
        w.set_state(q0, p0)
        w.run(6, is_nvt=False)
        w.sync()
        qr, pr, Fr, fr, dr = w.get_state()
        s = w.scalars()
        assert s["steps"] == 6
    point, lst, disp = o.all_lists(q0)
    F, fph = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
    o.force_eval_sc(q0, F, fph, point, lst)
    q, p = q0.copy(), p0.copy()
    h = 0.0
    for _ in range(6):
        e, h, sc, _ = o.update_sc(q, p, F, fph, point, lst, disp, False, None, h)
    assert np.max(np.abs(qr - q)) <= 1e-11 * np.max(np.abs(q))
    assert abs(s["energy"] - e) <= 1e-10 * abs(e)


def test_list_overflow_is_reported(W, oracle_mod):
    from qtip4pf_water_amd import ListOverflow
    say("list overflow")
    q, box = system(6, 18.64, 1, 101)
    with W(216, 1, box, list_size=1000) as w:
        pg, lg, dg = w.new_lists()
        with pytest.raises(ListOverflow):
            w.neighbour_list(q, pg, lg, dg, 0)


def test_nonfinite_positions_rejected(W):
    from qtip4pf_water_amd import QtpError
    say("nonfinite input")
    q, box = system(2, 6.3, 1, 111)
    q[5] = np.nan
    with W(8, 1, box) as w:
        pg, lg, dg = w.new_lists()
        with pytest.raises(QtpError):
            w.neighbour_list(q, pg, lg, dg, 0)


def test_pile_g_matches_reference_golden(W):
    """Reference thermostat.cpp output with std::mt19937(seed): the GPU pile_g
    fed the same stream's draws."""
    import os
    from qtip4pf_water_amd import HostRng
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "pile_g_27x8.npz"))
    P = int(g["n_beads"])
    say("pile_g golden 27 P", P)
    rng = HostRng(int(g["seed"]))
    p = g["p_in"].copy()
    with W(27, P, tuple(g["boxl"]), T_K=float(g["T_K"]), dtfs=float(g["dtfs"])) as w:
        h1 = w.pile_g(p, g["mass"], rng.draw_pile(81, P))
        assert np.max(np.abs(p - g["p_out1"])) <= 1e-13 * np.max(np.abs(g["p_out1"]))
        h2 = w.pile_g(p, g["mass"], rng.draw_pile(81, P))
        assert np.max(np.abs(p - g["p_out2"])) <= 1e-13 * np.max(np.abs(g["p_out2"]))
    ke = 0.5 * np.sum(g["p_in"].reshape(P, -1) ** 2 / g["mass"])
    assert abs(h1 - float(g["heat1"])) <= 1e-12 * ke
    assert abs(h2 - float(g["heat2"])) <= 1e-12 * ke


@pytest.mark.parametrize("P,nvt", [(4, True), (1, False)])
def test_reference_adapter_driver(oracle_mod, tmp_path, P, nvt):
    """The reference-named entry points of include/qtip4pf_ref_adapter.hpp,
    driven by a C++ loop in the role of simulation.cpp, against the oracle
    fed the same std::mt19937 stream."""
    import os
    import struct
    import subprocess
    from qtip4pf_water_amd import HostRng
    drv = os.path.join(os.path.dirname(__file__), "_bin", "qtp_adapter_driver")
    assert os.path.exists(drv), "build the driver: python -m qtip4pf_water_amd.build"
    n_side, L, steps, seed = 3, 9.4, 4, 77
    say("adapter driver 27 mols P", P, "nvt", nvt)
    q, box = system(n_side, L, P, 121)
    o = oracle_mod.Oracle(27, P, box)
    p = cfg.sample_momenta(27, P, o.sys.beta_n, seed=5)
    fin = tmp_path / "in.bin"
    with open(fin, "wb") as f:
        f.write(struct.pack("5i", 27, P, steps, 1 if nvt else 0, seed))
        f.write(struct.pack("7d", *box, 298.0, 0.25, 1e-6, 1e-5))
        f.write(q.tobytes())
        f.write(p.tobytes())
    fout = tmp_path / "out.bin"
    r = subprocess.run([drv, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = np.fromfile(fout, dtype=np.float64)
    N = P * o.n_dof
    qg, pg, Fg, eg = out[:N], out[N:2 * N], out[2 * N:3 * N], out[3 * N:3 * N + steps]
    # oracle run with the same stream
    point, lst, disp = o.all_lists(q)
    S = sum(term_scale(o, q, point, lst, b) for b in range(P))
    qr, pr = q.copy(), p.copy()
    F, fph = np.zeros(N), np.zeros(N)
    rng = HostRng(seed)
    heat = 0.0
    with oracle_mod.exact_energy():
        if P > 1:
            o.force_eval_sc(qr, F, fph, point, lst)
        else:
            o.get_forces(qr, F, point, lst, 0)
        for s in range(steps):
            if P > 1:
                dev = rng.draw_pile(o.n_atoms, P) if nvt else None
                er, heat, _, _ = o.update_sc(qr, pr, F, fph, point, lst, disp, nvt, dev, heat)
            else:
                er, _ = o.update_class(qr, pr, F, point, lst, disp)
            # the driver returns its positions after the last step only; the
            # trajectories drift apart monotonically, so the final position
            # difference bounds every step's (80-bit oracle energies)
            tol = P * trajectory_bound(er, F, qg - qr) + 16.0 * EPS * S
            assert abs(eg[s] - er) <= tol, (s, eg[s] - er, tol)
    assert np.max(np.abs(qg - qr)) <= 1e-12 * np.max(np.abs(qr))
    assert np.max(np.abs(pg - pr)) <= 1e-10 * np.max(np.abs(pr))


def test_orthorhombic_ice_box(W, oracle_mod):
    """Non-cubic box: proton-disordered ice from the reference's own start
    (initialise_orth, 4 x 3 x 3 orthorhombic cells, 288 molecules), forces,
    energies and lists against the oracle."""
    from qtip4pf_water_amd import HostRng
    say("orthorhombic ice 288 mols")
    P = 4
    q, box, _ = HostRng(2024).initialise_orth(4, 3, 3, 0.92, n_beads=P)
    assert len({round(b, 6) for b in box}) == 3        # three different edges
    q = q + np.random.default_rng(5).normal(0.0, 0.03, q.shape)
    n = 288
    o = oracle_mod.Oracle(n, P, box)
    point, lst, disp = o.all_lists(q)
    with W(n, P, box) as w:
        assert w.derived["k_max"] == o.ewald()["k_max"]
        assert w.derived["rcut"] == pytest.approx(o.ewald()["rcut"], rel=1e-15)
        pg, lg, dg = w.new_lists()
        for b in range(P):
            w.neighbour_list(q, pg, lg, dg, b)
        assert np.array_equal(pg, point) and np.array_equal(lg, lst)
        with oracle_mod.exact_energy():
            for b in (0, 3):
                Fr, F = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
                er = o.get_forces(q, Fr, point, lst, b)
                e = w.get_forces_nospline(q, F, o.charge, point, lst, b)
                sl = slice(b * o.n_dof, (b + 1) * o.n_dof)
                assert rel_atom_err(F[sl], Fr[sl]) < 1e-10
                assert single(e, er), (e - er) / er
            Fr, fr = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
            er, scr = o.force_eval_sc(q, Fr, fr, point, lst)
        F, f = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
        e, sc = w.force_eval_sc_nospline(q, F, o.charge, o.mass, point, lst, fphys=f)
    assert rel_atom_err(f, fr) < 1e-10
    assert single(e, er), (e - er) / er
