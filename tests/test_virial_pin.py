"""Pins of the independent periodic-cell energy used to check the GPU virial
(tests/ewald_bruteforce.py), on the CPU:

  * its energy at zero strain equals the oracle's get_forces energy
    (forces.cpp:468-630) at converged Ewald parameters;
  * its strain derivatives on the diagonal equal central differences of the
    oracle's energy in stretched orthorhombic boxes (the oracle re-derives
    kappa for each box; at convergence the energy does not depend on it).

The reference computes no virial, so these two facts are what fixes the
off-diagonal components the GPU test compares against.
"""
import numpy as np

import ewald_bruteforce as bf
from qtip4pf_water_amd import config as cfg


def _case(oracle_mod):
    q0, box = cfg.lattice_water_box(4, 18.64, seed=21, jitter_bohr=0.1)
    o = oracle_mod.Oracle(64, 1, box, eps=1e-16)
    return q0, box, o, o.ewald(), oracle_mod.consts()


def test_bruteforce_energy_matches_oracle(oracle_mod):
    q0, box, o, ew, c = _case(oracle_mod)
    point, lst, _ = o.all_lists(q0)
    F = np.zeros(o.n_dof)
    e, parts = o.get_forces(q0, F, point, lst, 0, parts=True)
    eb, pb = bf.energy(q0, box, c, ew["kappa"], ew["rcut"], ew["k_cut"], parts=True)
    # the reference's truncated pi (qtip4pfparams.h:45) moves the k-space terms ~1e-13
    assert abs(eb - e) <= 1e-10 * abs(parts[4])
    for i, k in enumerate(["lj", "intra", "real", "recip", "self"]):
        assert abs(pb[k] - parts[i]) <= 1e-10 * abs(parts[4]), k


def test_bruteforce_diagonal_virial_matches_oracle_fd(oracle_mod):
    q0, box, o, ew, c = _case(oracle_mod)
    h = 1e-5
    W = bf.virial_fd(q0, box, c, ew["kappa"], ew["rcut"], ew["k_cut"], h=h)
    for a in range(3):
        es = []
        for sgn in (1.0, -1.0):
            b2 = list(box)
            b2[a] *= 1.0 + sgn * h
            q2 = q0.reshape(-1, 3).copy()
            q2[:, a] *= 1.0 + sgn * h
            q2 = q2.ravel()
            o2 = oracle_mod.Oracle(64, 1, b2, eps=1e-16)
            p2, l2, _ = o2.all_lists(q2)
            es.append(o2.get_forces(q2, np.zeros(o2.n_dof), p2, l2, 0))
        fd = -(es[0] - es[1]) / (2.0 * h)
        assert abs(fd - W[a]) <= 1e-7 * np.max(np.abs(W)), (a, fd, W[a])
    # the off-diagonal components are not small for this configuration
    assert np.min(np.abs(W[3:5])) > 1e-2
