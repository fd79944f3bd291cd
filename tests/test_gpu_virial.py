"""GPU virial tensor (qtp_get_forces_virial, qtp_get_virial).

The reference computes no virial; the tensor W = -dV/d(strain) at fixed
Ewald parameters is pinned by
  * central differences of the GPU's own energy in stretched orthorhombic
    boxes (diagonal components), at the production Ewald parameters
    (eps_ewald = 1e-6) and at converged ones;
  * all six components against central differences of an independent
    periodic-cell energy (tests/ewald_bruteforce.py, itself pinned to the
    oracle in test_virial_pin.py) at converged Ewald parameters, where the
    virial no longer depends on how the Coulomb sum is split.
Tolerance: 1e-6 of the largest component (the differences carry ~1e-9
rounding and truncation error).
"""
import sys

import numpy as np
import pytest

import ewald_bruteforce as bf
from qtip4pf_water_amd import config as cfg

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _system(P):
    q0, box = cfg.lattice_water_box(4, 18.64, seed=21, jitter_bohr=0.1)
    q = cfg.replicate_beads(q0, P)
    if P > 1:
        q = q + np.random.default_rng(121).normal(0, 0.03, q.shape)
    return np.ascontiguousarray(q), box


def _override(w):
    d = w.derived
    return {"kappa": d["kappa"], "rcut": d["rcut"], "k_cut": d["k_cut"], "k_max": d["k_max"]}


def _gpu_energy(q0, box, eps, ew, oracle_mod):
    from qtip4pf_water_amd import Water
    o = oracle_mod.Oracle(64, 1, box)
    point, lst, _ = o.all_lists(q0)
    with Water(64, 1, box, eps_ewald=eps, ewald=ew) as w:
        e, W = w.get_forces_virial(q0, np.zeros(o.n_dof), o.charge, point, lst, 0)
    return e, W


def _diag_fd(q0, box, eps, ew, h, oracle_mod):
    out = np.zeros(3)
    for a in range(3):
        es = []
        for sgn in (1.0, -1.0):
            b2 = list(box)
            b2[a] *= 1.0 + sgn * h
            q2 = q0.reshape(-1, 3).copy()
            q2[:, a] *= 1.0 + sgn * h
            es.append(_gpu_energy(q2.ravel(), tuple(b2), eps, ew, oracle_mod)[0])
        out[a] = -(es[0] - es[1]) / (2.0 * h)
    return out


@pytest.mark.parametrize("eps,h", [(1e-6, 1e-6), (1e-16, 1e-5)])
def test_virial_diagonal_matches_energy_differences(oracle_mod, eps, h):
    """r_c is the derived L/2 pulled in by 2h so that it stays <= L/2 in the
    compressed boxes.  No site pair of this configuration lies within 3.6e-5
    (relative) of that r_c at eps 1e-6, so a strain of 1e-6 crosses no cutoff."""
    from qtip4pf_water_amd import Water
    say("virial diagonal FD", eps)
    q0, box = _system(1)
    with Water(64, 1, box, eps_ewald=eps) as w0:
        ew = _override(w0)
    ew["rcut"] *= 1.0 - 2.0 * h
    e, W = _gpu_energy(q0, box, eps, ew, oracle_mod)
    fd = _diag_fd(q0, box, eps, ew, h, oracle_mod)
    scale = np.max(np.abs(W))
    assert np.max(np.abs(np.diag(W) - fd)) <= 1e-6 * scale, (np.diag(W), fd)


def test_virial_all_components_match_bruteforce(oracle_mod):
    from qtip4pf_water_amd import Water
    say("virial vs brute force")
    q0, box = _system(1)
    with Water(64, 1, box, eps_ewald=1e-16) as w0:
        ew = _override(w0)
    e, W = _gpu_energy(q0, box, 1e-16, ew, oracle_mod)
    c = oracle_mod.consts()
    Wb = bf.virial_fd(q0, box, c, ew["kappa"], ew["rcut"], ew["k_cut"], h=1e-5)
    Wg = np.array([W[a, b] for a, b in bf.VOIGT])
    say("W_gpu", Wg, "W_bf", Wb)
    assert np.max(np.abs(Wg - Wb)) <= 1e-6 * np.max(np.abs(Wb))
    eb = bf.energy(q0, box, c, ew["kappa"], ew["rcut"], ew["k_cut"])
    assert abs(e - eb) <= 1e-10 * 23.0


def test_virial_same_for_both_realspace_block_shapes(oracle_mod):
    """k_realspace with two and with four waves per tile: the same pairs
    summed in a different order, so energy, forces and virial agree to
    rounding (1e-13 relative)."""
    from qtip4pf_water_amd import Water
    q0, box = _system(1)
    o = oracle_mod.Oracle(64, 1, box)
    point, lst, _ = o.all_lists(q0)
    out = []
    for rsw in (4, 2):
        say("virial, real-space waves", rsw)
        with Water(64, 1, box) as w:
            w.set_tuning(rs_waves=rsw)
            F = np.zeros(o.n_dof)
            e, W = w.get_forces_virial(q0, F, o.charge, point, lst, 0)
        out.append((e, W, F))
    (e4, W4, F4), (e2, W2, F2) = out
    assert abs(e4 - e2) <= 1e-13 * abs(e4)
    assert np.max(np.abs(W4 - W2)) <= 1e-13 * np.max(np.abs(W4))
    assert np.max(np.abs(F4 - F2)) <= 1e-13 * np.max(np.abs(F4))


def test_resident_virial_matches_per_bead_calls(oracle_mod):
    from qtip4pf_water_amd import Water
    say("resident virial P=2")
    P = 2
    q, box = _system(P)
    o = oracle_mod.Oracle(64, P, box)
    point, lst, _ = o.all_lists(q)
    p = cfg.sample_momenta(64, P, o.sys.beta_n, seed=5)
    with Water(64, P, box) as w:
        ref = [w.get_forces_virial(q, np.zeros(P * o.n_dof), o.charge, point, lst, b)
               for b in range(P)]
        w.set_state(q, p)
        Wr, er = w.virial()
    for b in range(P):
        assert abs(er[b] - ref[b][0]) <= 1e-13 * 23.0
        assert np.max(np.abs(Wr[b] - ref[b][1])) <= 1e-13 * np.max(np.abs(ref[b][1]))
    # and the bead energies are the oracle's
    for b in range(P):
        e_or = o.get_forces(q, np.zeros(P * o.n_dof), point, lst, b)
        assert abs(er[b] - e_or) <= 1e-12 * 23.0


def test_virial_under_chunk_groups_and_realspace_split(oracle_mod):
    """The virial form with several structure-factor chunks per block
    (k_sreduce's k-space virial from chunk-group partials) and with real-space
    tiles shared between the streams (its virial partials indexed by tile):
    against the default launch, energy and virial to 1e-13 relative with
    chunk groups (summation order), bit for bit with the split."""
    from qtip4pf_water_amd import Water
    q0, box = _system(1)
    o = oracle_mod.Oracle(64, 1, box)
    point, lst, _ = o.all_lists(q0)
    out = {}
    for name, kw in (("default", {}), ("chunks", {"ksum_chunks_per_block": 4}),
                     ("split", {"rs_split": 3})):
        say("virial launch", name)
        with Water(64, 1, box) as w:
            if kw:
                w.set_tuning(**kw)
            F = np.zeros(o.n_dof)
            e, W = w.get_forces_virial(q0, F, o.charge, point, lst, 0)
        out[name] = (e, W, F)
    e0, W0, F0 = out["default"]
    e1, W1, F1 = out["chunks"]
    assert abs(e1 - e0) <= 1e-13 * abs(e0)
    assert np.max(np.abs(W1 - W0)) <= 1e-13 * np.max(np.abs(W0))
    assert np.max(np.abs(F1 - F0)) <= 1e-13 * np.max(np.abs(F0))
    e2, W2, F2 = out["split"]
    assert e2 == e0 and np.array_equal(W2, W0) and np.array_equal(F2, F0)
