"""Real-space tile culling (the CULL form of k_realspace, boxes at least 2.5
r_c wide): a tile whose two 64-site blocks have bounding boxes farther apart
than r_c under every periodic image is skipped.  That is exact only if the
box gap bounds every pair's minimum-image distance from below; these cases
load it: molecules moved by whole box vectors (unwrapped coordinates, blocks
whose boxes span several periods), against the oracle and against the same
system wrapped."""
import sys

import numpy as np
import pytest

from qtip4pf_water_amd import config as cfg

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def rel_atom_err(F, Fr):
    d = np.linalg.norm((F - Fr).reshape(-1, 3), axis=1)
    n = np.linalg.norm(Fr.reshape(-1, 3), axis=1)
    return float(np.max(d / np.maximum(n, 1e-300)))


def _sites_launches(n, P, box, q, o, point, lst):
    """get_forces with per-kernel timing on: the number of k_sites launches
    (the CULL form's site kernel) and the forces and energy."""
    from qtip4pf_water_amd import Water
    with Water(n, P, box) as w:
        w.enable_timing(True)
        F = np.zeros(P * o.n_dof)
        e = w.get_forces_nospline(q, F, o.charge, point, lst, 0)
        return w.timing("sites")[1], F[: o.n_dof], e


def test_culling_only_for_wide_boxes(oracle_mod):
    for n_side, L, on in ((6, 18.64, False), (10, 31.07, True)):
        n = n_side ** 3
        q, box = cfg.lattice_water_box(n_side, L, seed=71, jitter_bohr=0.05)
        o = oracle_mod.Oracle(n, 1, box)
        say("culling on/off", n, "mols, box / r_c", box[0] / o.ewald()["rcut"])
        point, lst, _ = o.all_lists(q)
        launches, _, _ = _sites_launches(n, 1, box, q, o, point, lst)
        assert (launches > 0) == on, (n, launches)


@pytest.mark.parametrize("rsw,split", [(4, 0), (2, 3)])
def test_culled_forces_unwrapped_molecules(oracle_mod, rsw, split):
    """1000 H2O x 2 beads; 30% of the molecules moved by random whole box
    vectors (components -2..2 periods).  SC forces, one-bead forces and
    energies against the oracle on the same coordinates at the parity
    tolerances, and the two device runs against each other; real-space
    tiles of rsw waves, split eighths of them on the second stream."""
    from qtip4pf_water_amd import Water
    n_side, L, P = 10, 31.07, 2
    n = n_side ** 3
    q0, box = cfg.lattice_water_box(n_side, L, seed=72, jitter_bohr=0.05)
    q = cfg.replicate_beads(q0, P) + np.random.default_rng(73).normal(0, 0.03, P * 9 * n)
    rng = np.random.default_rng(74)
    shift = rng.integers(-2, 3, size=(n, 3)) * np.asarray(box)
    shift[rng.random(n) > 0.3] = 0.0
    qs = (q.reshape(P, n, 3, 3) + shift[None, :, None, :]).reshape(-1)
    say("unwrapped molecules", n, "mols P", P, "waves", rsw, "moved", int(np.count_nonzero(shift.any(axis=1))))
    o = oracle_mod.Oracle(n, P, box)
    assert box[0] >= 2.5 * o.ewald()["rcut"]
    point, lst, _ = o.all_lists(qs)
    scale = o.ewald()["prefactor"] * 0.5 * float(np.sum(o.charge ** 2))
    got = {}
    for name, x in (("unwrapped", qs), ("wrapped", q)):
        Fr, fr = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
        er, scr = o.force_eval_sc(x, Fr, fr, point, lst)
        F1r = np.zeros(P * o.n_dof)
        e1r = o.get_forces(x, F1r, point, lst, 0)
        with Water(n, P, box) as w:
            w.set_tuning(rs_waves=rsw, rs_split=split)
            F, f = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
            e, sc = w.force_eval_sc_nospline(x, F, o.charge, o.mass, point, lst, fphys=f)
            F1 = np.zeros(P * o.n_dof)
            e1 = w.get_forces_nospline(x, F1, o.charge, point, lst, 0)
        got[name] = (f, F1[: o.n_dof], e1)
        # a wrongly culled tile would move forces by ~1e-3 relative
        assert rel_atom_err(f, fr) < 1e-10, name
        assert rel_atom_err(F1[: o.n_dof], F1r[: o.n_dof]) < 1e-10, name
        # energies: V (a few Eh here) is the difference of Ewald terms of
        # ~200 Eh, so 1e-12 of the self-energy scale (as at 4096 and 5832)
        assert abs(e - er) <= 1e-12 * scale, name
        assert abs(e1 - e1r) <= 1e-12 * scale, name
        assert abs(sc - scr) <= 1e-12 * scale, name
        for b in range(P):
            sl = slice(b * o.n_dof, (b + 1) * o.n_dof)
            assert rel_atom_err(F[sl], Fr[sl]) < (1e-10 if b % 2 == 0 else 1e-9), (name, b)
    assert rel_atom_err(got["unwrapped"][0], got["wrapped"][0]) < 1e-11
    assert rel_atom_err(got["unwrapped"][1], got["wrapped"][1]) < 1e-11


def _pair_distances_near(sites, box, lo, hi):
    """Minimum-image distances of all site pairs that fall in [lo, hi]."""
    L = np.asarray(box)
    out = []
    for i in range(len(sites) - 1):
        d = sites[i + 1:] - sites[i]
        d -= L * np.round(d / L)
        r = np.sqrt(np.einsum("ij,ij->i", d, d))
        out.append(r[(r >= lo) & (r <= hi)])
    return np.sort(np.concatenate(out))


@pytest.mark.parametrize("split", [0, 3])
def test_culled_virial_matches_energy_differences(oracle_mod, split):
    """The virial form of the culled kernel: 1000 H2O (box 2.66 r_c), the
    diagonal of W = -dV/d(strain) against central differences of the GPU's
    own energy in stretched boxes at fixed Ewald parameters.  Strain 1e-8,
    and r_c moved into a gap of the site-pair distances at least 2e-7 r_c
    wide on either side, so no pair crosses the cutoff; tolerance 1e-6 of the
    largest component (the differences carry ~1e-13 Eh of rounding over
    2e-8).  split: eighths of the tiles on the second stream (the skipped
    ones' virial partials zeroed ahead of both launches)."""
    from qtip4pf_water_amd import Water
    n_side, L, h = 10, 31.07, 1e-8
    n = n_side ** 3
    q0, box = cfg.lattice_water_box(n_side, L, seed=75, jitter_bohr=0.1)
    g = oracle_mod.consts()["gamma_vs"]
    x = q0.reshape(n, 3, 3)
    sites = x.copy()
    sites[:, 0] = g * x[:, 0] + 0.5 * (1.0 - g) * (x[:, 1] + x[:, 2])
    with Water(n, 1, box) as w0:
        d = w0.derived
        ew = {"kappa": d["kappa"], "rcut": d["rcut"], "k_cut": d["k_cut"], "k_max": d["k_max"]}
    r = _pair_distances_near(sites.reshape(-1, 3), box, 0.995 * ew["rcut"], ew["rcut"])
    gaps = np.diff(r)
    k = int(np.argmax(gaps))
    assert gaps[k] > 4e-7 * ew["rcut"], gaps[k]
    ew["rcut"] = 0.5 * (r[k] + r[k + 1])
    assert box[0] >= 2.5 * ew["rcut"]
    say("culled virial FD", n, "mols, r_c", ew["rcut"], "gap", gaps[k], "split", split)

    def energy_virial(q, b):
        o = oracle_mod.Oracle(n, 1, b)
        point, lst, _ = o.all_lists(q)
        with Water(n, 1, b, ewald=ew) as w:
            w.set_tuning(rs_split=split)
            return w.get_forces_virial(q, np.zeros(o.n_dof), o.charge, point, lst, 0)

    _, W = energy_virial(q0, box)
    fd = np.zeros(3)
    for a in range(3):
        es = []
        for sgn in (1.0, -1.0):
            b2 = list(box)
            b2[a] *= 1.0 + sgn * h
            q2 = q0.reshape(-1, 3).copy()
            q2[:, a] *= 1.0 + sgn * h
            es.append(energy_virial(q2.ravel(), tuple(b2))[0])
        fd[a] = -(es[0] - es[1]) / (2.0 * h)
    say("W diag", np.diag(W), "FD", fd)
    assert np.max(np.abs(np.diag(W) - fd)) <= 1e-6 * np.max(np.abs(W))


@pytest.mark.parametrize("rsw", [4, 2])
def test_skipped_tiles_leave_no_stale_slots(oracle_mod, rsw):
    """A skipped tile writes no slot entries, and the force kernels must not
    read them.  One Water object evaluates 1000 H2O x 2 beads with the
    molecules in random order (every tile within r_c: every slot written),
    then the same molecules in lattice order (most tiles skipped, their
    slots still holding the first evaluation's values): forces and energies
    of the second evaluation against the oracle at the parity tolerances,
    both the single-bead and the Suzuki-Chin forms."""
    from qtip4pf_water_amd import Water
    n_side, L, P = 10, 31.07, 2
    n = n_side ** 3
    q0, box = cfg.lattice_water_box(n_side, L, seed=76, jitter_bohr=0.05)
    q = cfg.replicate_beads(q0, P) + np.random.default_rng(77).normal(0, 0.03, P * 9 * n)
    perm = np.random.default_rng(78).permutation(n)
    qp = q.reshape(P, n, 9)[:, perm].reshape(-1)
    o = oracle_mod.Oracle(n, P, box)
    assert box[0] >= 2.5 * o.ewald()["rcut"]
    say("stale slots", n, "mols P", P, "waves", rsw)
    pp, lp, _ = o.all_lists(qp)
    point, lst, _ = o.all_lists(q)
    scale = o.ewald()["prefactor"] * 0.5 * float(np.sum(o.charge ** 2))
    Fr, fr = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
    er, _ = o.force_eval_sc(q, Fr, fr, point, lst)
    F1r = np.zeros(P * o.n_dof)
    e1r = o.get_forces(q, F1r, point, lst, 0)
    with Water(n, P, box) as w:
        w.set_tuning(rs_waves=rsw)
        w.force_eval_sc_nospline(qp, np.zeros(P * o.n_dof), o.charge, o.mass, pp, lp,
                                 fphys=np.zeros(P * o.n_dof))
        F, f = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
        e, _ = w.force_eval_sc_nospline(q, F, o.charge, o.mass, point, lst, fphys=f)
        w.get_forces_nospline(qp, np.zeros(P * o.n_dof), o.charge, pp, lp, 0)
        F1 = np.zeros(P * o.n_dof)
        e1 = w.get_forces_nospline(q, F1, o.charge, point, lst, 0)
    assert rel_atom_err(f, fr) < 1e-10
    assert rel_atom_err(F1[: o.n_dof], F1r[: o.n_dof]) < 1e-10
    assert abs(e - er) <= 1e-12 * scale
    assert abs(e1 - e1r) <= 1e-12 * scale


@pytest.mark.parametrize("rsw", [4, 2])
def test_culled_partial_clusters(oracle_mod, rsw):
    """999 H2O x 2 beads (the 1000-molecule lattice less its last molecule):
    2,997 sites, so the last 64-site block holds 53 -- its seventh 8-site
    cluster 5 sites, its eighth none -- and the cluster-pair test meets a
    partly filled and an empty cluster.  SC and single-bead forces and
    energies against the oracle at the parity tolerances."""
    from qtip4pf_water_amd import Water
    n_side, L, P = 10, 31.07, 2
    q0, box = cfg.lattice_water_box(n_side, L, seed=79, jitter_bohr=0.05)
    n = n_side ** 3 - 1
    q0 = q0[: 9 * n]
    q = cfg.replicate_beads(q0, P) + np.random.default_rng(80).normal(0, 0.03, P * 9 * n)
    assert (3 * n) % 64 == 53
    o = oracle_mod.Oracle(n, P, box)
    assert box[0] >= 2.5 * o.ewald()["rcut"]
    say("partial clusters", n, "mols P", P, "waves", rsw)
    point, lst, _ = o.all_lists(q)
    scale = o.ewald()["prefactor"] * 0.5 * float(np.sum(o.charge ** 2))
    Fr, fr = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
    er, scr = o.force_eval_sc(q, Fr, fr, point, lst)
    F1r = np.zeros(P * o.n_dof)
    e1r = o.get_forces(q, F1r, point, lst, 0)
    with Water(n, P, box) as w:
        w.set_tuning(rs_waves=rsw)
        F, f = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
        e, sc = w.force_eval_sc_nospline(q, F, o.charge, o.mass, point, lst, fphys=f)
        F1 = np.zeros(P * o.n_dof)
        e1 = w.get_forces_nospline(q, F1, o.charge, point, lst, 0)
    assert rel_atom_err(f, fr) < 1e-10
    assert rel_atom_err(F1[: o.n_dof], F1r[: o.n_dof]) < 1e-10
    assert abs(e - er) <= 1e-12 * scale
    assert abs(e1 - e1r) <= 1e-12 * scale
    assert abs(sc - scr) <= 1e-12 * scale
    for b in range(P):
        sl = slice(b * o.n_dof, (b + 1) * o.n_dof)
        assert rel_atom_err(F[sl], Fr[sl]) < (1e-10 if b % 2 == 0 else 1e-9), b
