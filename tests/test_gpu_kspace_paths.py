"""The two implementations of the reciprocal-space sums and forces: the
matrix-core kernels (k_ksum_mf / k_kforce_mf, used for k_max <= 23) and the
vector kernels (k_ksum / k_kforce, used above that or with QTP_KSPACE=vector
set when a context is created).  Both against the oracle at the parity
tolerances, and against each other (both FP64; they differ only in summation
order)."""
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


def system(n_side, L_angs, P, seed):
    q0, box = cfg.lattice_water_box(n_side, L_angs, seed=seed, jitter_bohr=0.05)
    q = cfg.replicate_beads(q0, P)
    if P > 1:
        q = q + np.random.default_rng(seed + 100).normal(0, 0.03, q.shape)
    return np.ascontiguousarray(q), box


def _forces(monkeypatch, path, n, P, box, q, o, point, lst, ewald=None):
    from qtip4pf_water_amd import Water
    if path == "vector":
        monkeypatch.setenv("QTP_KSPACE", "vector")
    else:
        monkeypatch.delenv("QTP_KSPACE", raising=False)
    with Water(n, P, box, ewald=ewald) as w:
        assert w.tuning()["matrix_kspace"] == (path == "matrix")
        F, f = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
        e, sc = w.force_eval_sc_nospline(q, F, o.charge, o.mass, point, lst, fphys=f)
        F1 = np.zeros(P * o.n_dof)
        e1 = w.get_forces_nospline(q, F1, o.charge, point, lst, 0)
    return F, f, e, sc, F1[: o.n_dof], e1


def _stated(box, n_max):
    """alpha = 0.3 / A, r_c = 9 A, every k = 2 pi n / L with |n| <= n_max."""
    return {"kappa": 0.3 / cfg.ANGS2BOHR, "rcut": 9.0 * cfg.ANGS2BOHR,
            "k_cut": 2.0 * np.pi * n_max / float(box[0]) * (1.0 + 1e-9), "k_max": n_max}


@pytest.mark.parametrize("n_side,L,P,seed,n_max", [
    (6, 18.64, 4, 51, None), (3, 9.4, 2, 52, None),
    (6, 18.64, 2, 54, 18),    # 2 K = 38 z-factor entries: 10 k-steps, three n tiles
    (6, 18.64, 2, 55, 23)])   # the largest k-set of the matrix-core kernels
def test_both_kspace_paths_match_oracle_and_each_other(monkeypatch, oracle_mod, n_side, L, P, seed,
                                                       n_max):
    n = n_side ** 3
    say("k-space paths", n, "mols P", P, "n_max", n_max)
    q, box = system(n_side, L, P, seed)
    ew = None if n_max is None else _stated(box, n_max)
    o = oracle_mod.Oracle(n, P, box, ewald=ew)
    point, lst, _ = o.all_lists(q)
    Fr, fr = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
    er, scr = o.force_eval_sc(q, Fr, fr, point, lst)
    F1r = np.zeros(P * o.n_dof)
    e1r = o.get_forces(q, F1r, point, lst, 0)
    got = {}
    for path in ("matrix", "vector"):
        F, f, e, sc, F1, e1 = _forces(monkeypatch, path, n, P, box, q, o, point, lst, ew)
        assert rel_atom_err(f, fr) < 1e-10, path
        assert rel_atom_err(F1, F1r[: o.n_dof]) < 1e-10, path
        assert abs(e - er) <= 1e-12 * abs(er), path
        assert abs(sc - scr) <= 1e-12 * abs(scr), path
        assert abs(e1 - e1r) <= 1e-12 * abs(e1r), path
        for b in range(P):
            sl = slice(b * o.n_dof, (b + 1) * o.n_dof)
            assert rel_atom_err(F[sl], Fr[sl]) < (1e-10 if b % 2 == 0 else 1e-9), (path, b)
        got[path] = (f, F1, e1)
    # the two paths: same sums in another order
    assert rel_atom_err(got["matrix"][0], got["vector"][0]) < 1e-12
    assert rel_atom_err(got["matrix"][1], got["vector"][1]) < 1e-12
    assert abs(got["matrix"][2] - got["vector"][2]) <= 1e-13 * abs(got["vector"][2])


def test_largest_matrix_core_k_max(monkeypatch, oracle_mod):
    """k_max = 15, the largest k-set of the eight-k-step matrix-core kernels
    (2 K = 32 z-factor entries): 5,832 H2O at liquid density, one bead."""
    from qtip4pf_water_amd import Water
    n_side, L = 18, 55.94
    n = n_side ** 3
    say("k_max 15 box", n, "mols")
    q, box = system(n_side, L, 1, 53)
    o = oracle_mod.Oracle(n, 1, box)
    assert o.ewald()["k_max"] == 15
    point, lst, _ = o.all_lists(q)
    Fr = np.zeros(o.n_dof)
    er = o.get_forces(q, Fr, point, lst, 0)
    scale = o.ewald()["prefactor"] * 0.5 * float(np.sum(o.charge ** 2))
    for path in ("matrix", "vector"):
        if path == "vector":
            monkeypatch.setenv("QTP_KSPACE", "vector")
        else:
            monkeypatch.delenv("QTP_KSPACE", raising=False)
        with Water(n, 1, box) as w:
            assert w.derived["k_max"] == 15
            F = np.zeros(o.n_dof)
            e = w.get_forces_nospline(q, F, o.charge, point, lst, 0)
        assert rel_atom_err(F, Fr) < 1e-10, path
        # energies at this size: 1e-12 of the self-energy scale (as for 4096)
        assert abs(e - er) <= 1e-12 * scale, path


def test_nve_conservation_both_paths(monkeypatch):
    """216 x 32 SC-PIMD NVE, 4000 steps from one start on each k-space path:
    the conserved quantity's largest excursion of the matrix-core path is no
    worse than the vector path's (the oracle is too slow for this length at
    this size; the vector path is the one the shorter oracle comparisons of
    test_gpu_drift.py also hold), and both stay small against the kinetic
    energy."""
    from qtip4pf_water_amd import Water
    n_side, L, P = 6, 18.64, 32
    n = n_side ** 3
    say("NVE conservation, both paths", n, "mols P", P)
    q0, box = cfg.lattice_water_box(n_side, L, seed=61, jitter_bohr=0.05)
    beta_n = 1.0 / (cfg.K2AU * 298.0) / P
    q = cfg.free_ring_polymer_beads(q0, P, beta_n, seed=62)
    p = cfg.sample_momenta(n, P, beta_n, seed=63)
    exc = {}
    for path in ("matrix", "vector"):
        if path == "vector":
            monkeypatch.setenv("QTP_KSPACE", "vector")
        else:
            monkeypatch.delenv("QTP_KSPACE", raising=False)
        with Water(n, P, box) as w:
            w.set_state(q, p)
            c0 = w.observables()["conserved"]
            tr = w.run_observed(4000, is_nvt=False, every=50)
            assert w.scalars()["status"] == 0
        e = np.concatenate([[c0], tr[:, 1]])
        assert np.all(np.isfinite(e)), path
        exc[path] = float(np.max(np.abs(e - e[0])))
    _, mass = cfg.charges_masses(n)
    ke = 0.5 * float(np.sum(p * p / np.tile(mass, P)))
    say(f"excursions: matrix {exc['matrix']:.3e} vector {exc['vector']:.3e} Eh (KE {ke:.3e})")
    assert exc["matrix"] <= 1.5 * exc["vector"] + 1e-6 * ke
    assert exc["matrix"] < 1e-2 * ke


@pytest.mark.parametrize("n_side,L,P,seed,n_max", [
    (6, 18.64, 4, 61, None),   # 11 site chunks of 64
    (6, 18.64, 2, 62, 18)])    # 13 chunks of 52; several unit slices per chunk group
def test_structure_factor_chunk_groups(oracle_mod, n_side, L, P, seed, n_max):
    """k_ksum_mf with 1, 2, 4, 8 and 16 site chunks per block (the block's
    accumulators carried across its chunks, one partial per chunk group for
    k_sreduce): forces and energies against the oracle at the parity
    tolerances and against the one-chunk form to 1e-13 relative (the same
    sums in another order)."""
    from qtip4pf_water_amd import Water
    n = n_side ** 3
    q, box = system(n_side, L, P, seed)
    ew = None if n_max is None else _stated(box, n_max)
    o = oracle_mod.Oracle(n, P, box, ewald=ew)
    point, lst, _ = o.all_lists(q)
    Fr, fr = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
    with oracle_mod.exact_energy():
        er, _ = o.force_eval_sc(q, Fr, fr, point, lst)
    out = {}
    with Water(n, P, box, ewald=ew) as w:
        assert w.tuning()["matrix_kspace"] == 1
        for cpb in (1, 2, 4, 8, 16):
            say("chunk groups", n, "n_max", n_max, "chunks per block", cpb)
            w.set_tuning(ksum_chunks_per_block=cpb)
            F, f = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
            e, _ = w.force_eval_sc_nospline(q, F, o.charge, o.mass, point, lst, fphys=f)
            assert rel_atom_err(f, fr) < 1e-10, cpb
            assert abs(e - er) <= 1e-12 * abs(er), cpb
            out[cpb] = (f, e)
    for cpb in (2, 4, 8, 16):
        assert rel_atom_err(out[cpb][0], out[1][0]) < 1e-13, cpb
        assert abs(out[cpb][1] - out[1][1]) <= 1e-13 * abs(out[1][1]), cpb
