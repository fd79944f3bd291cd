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


def test_culled_forces_unwrapped_molecules(oracle_mod):
    """1000 H2O x 2 beads; 30% of the molecules moved by random whole box
    vectors (components -2..2 periods).  SC forces, one-bead forces and
    energies against the oracle on the same coordinates at the parity
    tolerances, and the two device runs against each other."""
    from qtip4pf_water_amd import Water
    n_side, L, P = 10, 31.07, 2
    n = n_side ** 3
    q0, box = cfg.lattice_water_box(n_side, L, seed=72, jitter_bohr=0.05)
    q = cfg.replicate_beads(q0, P) + np.random.default_rng(73).normal(0, 0.03, P * 9 * n)
    rng = np.random.default_rng(74)
    shift = rng.integers(-2, 3, size=(n, 3)) * np.asarray(box)
    shift[rng.random(n) > 0.3] = 0.0
    qs = (q.reshape(P, n, 3, 3) + shift[None, :, None, :]).reshape(-1)
    say("unwrapped molecules", n, "mols P", P, "moved", int(np.count_nonzero(shift.any(axis=1))))
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
