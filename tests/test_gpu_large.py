"""The large single system: 4096 H2O in a 49.72 A box (k_max = 14 and
r_c = 14.8 A with the reference's parameter derivation, forces.cpp:573-584),
against the oracle on the same seeded inputs, at the tolerances of the
smaller systems.  The oracle's single-bead force evaluation at this size
takes a few seconds, so each check uses one or two beads; the 32-bead run
checks that the resident step runs at full size."""
import sys

import numpy as np
import pytest

from qtip4pf_water_amd import config as cfg

from energy_tol import single

pytestmark = pytest.mark.gpu

N_SIDE, L_ANGS = 16, 49.72


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def e_self_scale(o):
    """|Ewald self energy| (forces.cpp:622-624), the largest single term of V.
    At 4096 waters V (~10 Eh) is the difference of Ewald terms of ~600 Eh; the
    reference's FP64 running sum over ~8M real-space pair terms rounds by
    ~1e-11 Eh by itself, so that order is held to 1e-12 of this scale, the
    GPU to 1e-12 of V against the 80-bit oracle sum (energy_tol)."""
    return o.ewald()["prefactor"] * 0.5 * float(np.sum(o.charge ** 2))


def rel_atom_err(F, Fr):
    d = np.linalg.norm((F - Fr).reshape(-1, 3), axis=1)
    n = np.linalg.norm(Fr.reshape(-1, 3), axis=1)
    return float(np.max(d / np.maximum(n, 1e-300)))


def _system(P, seed):
    q0, box = cfg.lattice_water_box(N_SIDE, L_ANGS, seed=seed, jitter_bohr=0.05)
    q = cfg.replicate_beads(q0, P)
    if P > 1:
        q = q + np.random.default_rng(seed + 100).normal(0, 0.03, q.shape)
    return np.ascontiguousarray(q), box


def test_large_neighbour_list_and_forces(oracle_mod):
    from qtip4pf_water_amd import Water
    n = N_SIDE ** 3
    say("large get_forces_nospline", n, "mols")
    q, box = _system(1, 41)
    o = oracle_mod.Oracle(n, 1, box)
    pr, lr, dr = o.new_lists()
    o.neighbour_list(q, pr, lr, dr, 0)
    Fr = np.zeros(o.n_dof)
    er, parts = o.get_forces(q, Fr, pr, lr, 0, parts=True)
    ex = parts[5]
    with Water(n, 1, box) as w:
        assert w.derived["k_max"] == o.ewald()["k_max"] == 14
        pg, lg, dg = w.new_lists()
        w.neighbour_list(q, pg, lg, dg, 0)
        assert np.array_equal(pg, pr) and np.array_equal(lg, lr) and np.array_equal(dg, dr)
        F = np.zeros(o.n_dof)
        e = w.get_forces_nospline(q, F, o.charge, pg, lg, 0)
    assert rel_atom_err(F, Fr) < 1e-10
    assert single(e, ex), (e - ex) / ex
    assert abs(er - ex) <= 1e-12 * e_self_scale(o)   # the reference's FP64 sum


def test_large_resident_sc_forces(oracle_mod):
    """Resident state at P = 2: lists built on the device and the SC forces of
    qtp_set_state against force_eval_sc_nospline of the oracle."""
    from qtip4pf_water_amd import Water
    n, P = N_SIDE ** 3, 2
    say("large resident SC forces", n, "mols P", P)
    q, box = _system(P, 42)
    o = oracle_mod.Oracle(n, P, box)
    point, lst, _ = o.all_lists(q)
    Fr, fr = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
    with oracle_mod.exact_energy():
        er, scr = o.force_eval_sc(q, Fr, fr, point, lst)
    with Water(n, P, box) as w:
        w.set_state(q, np.zeros_like(q))
        _, _, F, f, _ = w.get_state()
        s = w.scalars()
        pg, lg = w.get_lists()
    assert np.array_equal(pg, point) and np.array_equal(lg, lst)
    assert rel_atom_err(f, fr) < 1e-10
    assert rel_atom_err(F[: o.n_dof], Fr[: o.n_dof]) < 1e-10   # even bead
    assert rel_atom_err(F[o.n_dof:], Fr[o.n_dof:]) < 1e-9      # odd bead (FD Hessian term)
    assert single(s["energy"], er), (s["energy"] - er) / er
    assert single(s["sc_pot"], scr), (s["sc_pot"] - scr) / scr


def test_large_32_bead_steps():
    """4096 x 32 resident NVT steps: runs, stays finite, reports no overflow."""
    from qtip4pf_water_amd import Water
    n, P = N_SIDE ** 3, 32
    say("large resident run", n, "mols P", P)
    q0, box = cfg.lattice_water_box(N_SIDE, L_ANGS, seed=43, jitter_bohr=0.05)
    beta_n = 1.0 / (cfg.K2AU * 298.0) / P
    q = cfg.free_ring_polymer_beads(q0, P, beta_n, seed=44)
    p = cfg.sample_momenta(n, P, beta_n, seed=45)
    with Water(n, P, box) as w:
        w.set_seed(5)
        w.set_state(q, p)
        w.run(3, is_nvt=True)
        s = w.scalars()
        q1, p1, F1, _, _ = w.get_state()
    assert s["status"] == 0 and s["steps"] == 3
    assert np.all(np.isfinite(q1)) and np.all(np.isfinite(p1)) and np.all(np.isfinite(F1))
    assert np.isfinite(s["energy"])


def test_large_checkpoint_continues_bit_for_bit(tmp_path):
    """A restart from a checkpoint continues exactly at this size too (the
    four-kernel list path, k_max = 14, 196 real-space blocks per bead)."""
    from qtip4pf_water_amd import Water
    n, P = N_SIDE ** 3, 8  # (P = 2 from a free-ring-polymer start is unstable at any size)
    say("large checkpoint", n, "mols P", P)
    q0, box = cfg.lattice_water_box(N_SIDE, L_ANGS, seed=46, jitter_bohr=0.05)
    beta_n = 1.0 / (cfg.K2AU * 298.0) / P
    q = cfg.free_ring_polymer_beads(q0, P, beta_n, seed=47)
    p = cfg.sample_momenta(n, P, beta_n, seed=48)

    def snap(w):
        qq, pp, F, f, d = w.get_state()
        return qq, pp, F, f, d, w.scalars()["heat"]

    with Water(n, P, box) as w:
        w.set_seed(9)
        w.set_state(q, p)
        w.run(12, is_nvt=True)
        ref = snap(w)
    path = tmp_path / "large.qtp"
    with Water(n, P, box) as w:
        w.set_seed(9)
        w.set_state(q, p)
        w.run(5, is_nvt=True)
        w.save_checkpoint(path)
    with Water(n, P, box) as w:
        w.load_checkpoint(path)
        w.run(7, is_nvt=True)
        got = snap(w)
    assert np.all(np.isfinite(ref[0])) and np.all(np.isfinite(ref[2]))
    for a, b in zip(got, ref):
        assert np.array_equal(np.asarray(a), np.asarray(b))


def _nominal_ewald(box):
    """alpha = 0.3 / A, r_c = 9 A, every k = 2 pi n / L with |n| <= 18."""
    L = float(box[0])
    return {"kappa": 0.3 / cfg.ANGS2BOHR, "rcut": 9.0 * cfg.ANGS2BOHR,
            "k_cut": 2.0 * np.pi * 18.0 / L * (1.0 + 1e-9), "k_max": 18}


def test_large_nominal_ewald_kmax18(oracle_mod):
    """The Ewald point stated for this system (alpha = 0.3 / A, r_c = 9 A,
    |n| <= 18: the twelve-k-step matrix-core kernels) with cell-binned lists:
    lists, forces and energy against the oracle given the same parameters."""
    from qtip4pf_water_amd import Water
    n = N_SIDE ** 3
    say("large nominal Ewald k_max 18", n, "mols")
    q, box = _system(1, 51)
    ew = _nominal_ewald(box)
    o = oracle_mod.Oracle(n, 1, box, ewald=ew)
    assert o.ewald()["k_max"] == 18
    pr, lr, dr = o.new_lists()
    o.neighbour_list(q, pr, lr, dr, 0)
    Fr = np.zeros(o.n_dof)
    er, parts = o.get_forces(q, Fr, pr, lr, 0, parts=True)
    with Water(n, 1, box, ewald=ew) as w:
        assert w.derived["k_max"] == 18 and w.list_info()["cells"]
        pg, lg, dg = w.new_lists()
        w.neighbour_list(q, pg, lg, dg, 0)
        assert np.array_equal(pg, pr) and np.array_equal(lg, lr)
        F = np.zeros(o.n_dof)
        e = w.get_forces_nospline(q, F, o.charge, pg, lg, 0)
    assert rel_atom_err(F, Fr) < 1e-10
    assert single(e, parts[5]), (e - parts[5]) / parts[5]


def test_large_nominal_32_bead_run_with_forced_rebuilds():
    """4096 x 32, nominal Ewald point, cell lists rebuilt every 10 steps on
    top of the displacement test: 12 NVT steps run clean, and every bead was
    rebuilt at step 10."""
    from qtip4pf_water_amd import Water
    n, P = N_SIDE ** 3, 32
    say("large nominal resident run", n, "mols P", P)
    q0, box = cfg.lattice_water_box(N_SIDE, L_ANGS, seed=53, jitter_bohr=0.05)
    beta_n = 1.0 / (cfg.K2AU * 298.0) / P
    q = cfg.free_ring_polymer_beads(q0, P, beta_n, seed=54)
    p = cfg.sample_momenta(n, P, beta_n, seed=55)
    with Water(n, P, box, ewald=_nominal_ewald(box)) as w:
        w.set_lists("auto", rebuild_every=10)
        w.set_seed(6)
        w.set_state(q, p)
        r0 = w.scalars()["rebuilds"]
        w.run(12, is_nvt=True)
        s = w.scalars()
        q1, p1, F1, _, _ = w.get_state()
    assert s["status"] == 0 and s["steps"] == 12
    assert s["rebuilds"] - r0 >= P
    assert np.all(np.isfinite(q1)) and np.all(np.isfinite(p1)) and np.all(np.isfinite(F1))
