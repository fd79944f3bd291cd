"""Independent qTIP4P/F energy of one bead in a general (triclinic) periodic
cell, for finite-difference checks of the virial tensor.

The cell is the orthorhombic box deformed by F = I + strain; positions are the
unwrapped molecules mapped by F.  Unlike the reference's routines
(forces.cpp), the electrostatics is written for an arbitrary cell: the real
space sum takes each site pair at its unstrained minimum image, the
reciprocal sum runs over the full sphere |k| < k_cut of the deformed cell's
reciprocal lattice with complex exponentials.  With converged Ewald
parameters (erfc(kappa r_c) and exp(-k_cut^2 / 4 kappa^2) near 1e-16) the
result is the Coulomb lattice energy, independent of how it is split, so a
derivative of this energy with respect to strain is the virial any correct
implementation must reproduce.

Test infrastructure only.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erfc


def _unwrap(q0, box):
    X = np.asarray(q0, float).reshape(-1, 3, 3).copy()
    L = np.asarray(box, float)
    for s in (1, 2):
        d = X[:, s] - X[:, 0]
        X[:, s] = X[:, 0] + d - L * np.round(d / L)
    return X


def energy(q0, box, c, kappa, rcut, kcut, strain=None, parts=False):
    """Energy (hartree) of the bead slice q0 (reference layout, bohr) in the
    box deformed by I + strain.  c: the reference's constants (dict)."""
    Fm = np.eye(3) + (np.zeros((3, 3)) if strain is None else np.asarray(strain, float))
    L = np.asarray(box, float)
    Xu = _unwrap(q0, box)
    n = Xu.shape[0]
    X = Xu @ Fm.T
    h = Fm @ np.diag(L)                 # cell vectors as columns
    V = abs(np.linalg.det(h))

    # intramolecular: quartic bond and harmonic angle
    r1 = X[:, 0] - X[:, 1]
    r2 = X[:, 0] - X[:, 2]
    e_intra = 0.0
    for r in (r1, r2):
        x = np.linalg.norm(r, axis=1) - c["roheq"]
        e_intra += np.sum(c["Dr"] * (c["ar2"] * x ** 2 - c["ar3"] * x ** 3 + c["ar4fac"] * x ** 4))
    cosang = np.sum(r1 * r2, axis=1) / (np.linalg.norm(r1, axis=1) * np.linalg.norm(r2, axis=1))
    e_intra += np.sum(0.5 * c["ktheta"] * (np.arccos(cosang) - c["theta_eq"]) ** 2)

    iu, ju = np.triu_indices(n, 1)

    # Lennard-Jones between oxygens, shifted at r_cut
    dO = Xu[iu, 0] - Xu[ju, 0]
    dO = (dO - L * np.round(dO / L)) @ Fm.T
    rO = np.linalg.norm(dO, axis=1)
    m = rO < c["r_cut"]
    sr6 = (c["sigmalj"] / rO[m]) ** 6
    e_lj = np.sum(4.0 * c["epslj"] * (sr6 * sr6 - sr6) - c["e_cut"])

    # charge sites: M, H1, H2
    g = c["gamma_vs"]
    Mu = g * Xu[:, 0] + 0.5 * (1.0 - g) * (Xu[:, 1] + Xu[:, 2])
    Su = np.stack([Mu, Xu[:, 1], Xu[:, 2]], axis=1).reshape(-1, 3)
    qs = np.tile([c["qo"], c["qh"], c["qh"]], n)
    S = Su @ Fm.T
    ns = Su.shape[0]
    ia, ja = np.triu_indices(ns, 1)
    same = (ia // 3) == (ja // 3)
    d = Su[ia] - Su[ja]
    d = np.where(same[:, None], d, d - L * np.round(d / L)) @ Fm.T
    r = np.linalg.norm(d, axis=1)
    qq = qs[ia] * qs[ja]
    keep = same | (r < rcut)
    ec = erfc(kappa * r[keep]) - same[keep]
    e_real = np.sum(qq[keep] * ec / r[keep])

    # reciprocal space over the deformed cell's reciprocal lattice, half sphere x2
    G = 2.0 * math.pi * np.linalg.inv(h).T   # rows: reciprocal vectors b_i
    mmax = [int(math.ceil(kcut / np.linalg.norm(G[i]))) + 1 for i in range(3)]
    mx, my, mz = np.meshgrid(np.arange(-mmax[0], mmax[0] + 1), np.arange(-mmax[1], mmax[1] + 1),
                             np.arange(-mmax[2], mmax[2] + 1), indexing="ij")
    mv = np.stack([mx.ravel(), my.ravel(), mz.ravel()], axis=1)
    half = (mv[:, 0] > 0) | ((mv[:, 0] == 0) & ((mv[:, 1] > 0) | ((mv[:, 1] == 0) & (mv[:, 2] > 0))))
    mv = mv[half]
    k = mv @ G
    k2 = np.sum(k * k, axis=1)
    sel = k2 < kcut * kcut
    k, k2 = k[sel], k2[sel]
    e_rec = 0.0
    for a in range(0, k.shape[0], 4096):
        kb = k[a:a + 4096]
        ph = kb @ S.T
        Sk = np.cos(ph) @ qs + 1j * (np.sin(ph) @ qs)
        w = np.exp(-k2[a:a + 4096] / (4.0 * kappa * kappa)) / k2[a:a + 4096]
        e_rec += np.sum(w * np.abs(Sk) ** 2)
    e_rec *= 2.0 * (2.0 * math.pi / V)

    e_self = -kappa / math.sqrt(math.pi) * np.sum(qs * qs)
    tot = e_lj + e_intra + e_real + e_rec + e_self
    if parts:
        return tot, dict(lj=e_lj, intra=e_intra, real=e_real, recip=e_rec, self=e_self)
    return tot


VOIGT = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]


def virial_fd(q0, box, c, kappa, rcut, kcut, h=1e-5):
    """W_ab = -dE/d(strain_ab) for the symmetric strain, by central
    differences; returned in Voigt order xx, yy, zz, xy, xz, yz."""
    out = np.zeros(6)
    for n, (a, b) in enumerate(VOIGT):
        e = np.zeros((3, 3))
        if a == b:
            e[a, a] = 1.0
        else:
            e[a, b] = e[b, a] = 0.5
        ep = energy(q0, box, c, kappa, rcut, kcut, strain=h * e)
        em = energy(q0, box, c, kappa, rcut, kcut, strain=-h * e)
        out[n] = -(ep - em) / (2.0 * h)
    return out
