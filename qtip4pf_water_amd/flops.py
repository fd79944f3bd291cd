"""Algorithmic work of one single-bead force evaluation (forces.cpp:468-630).

Useful operations only, FP64, each add/mul/sub/div/sqrt/exp/erfc one FLOP:

  W = 51 N_at n_xyz + 24 N_at (n_xz + n_xy) + 10 N_at n_x + N_at (14 + 20 k_max)
      + 32 N_ew + 27 N_LJ + 205 N_mol

n_x, n_xz, n_xy, n_xyz are the iteration counts of kwald's four branches
(forces.cpp:258, 283, 329, 376); N_ew = N_at(N_at-1)/2 * (4 pi r_c^3/3)/V and
N_LJ = N_mol(N_mol-1)/2 * (4 pi r_cut^3/3)/V are the pair counts inside the
cutoffs.  Per branch and site, kwald spends (structure-factor part, force part)
= (24, 27) on xyz, (12, 12) on xz/xy and (2, 8) on x-only; the factor tables
(14 + 20 k_max per site) are split evenly between the two kernels that build
them.  A Suzuki-Chin step does 2P such evaluations.
"""
from __future__ import annotations

import math

from . import config as cfg


def kgroup_counts(boxl, k_cut, k_max):
    """Branch iteration counts of kwald with the reference's loop bounds."""
    bk = [2.0 * cfg.PI_REF / L for L in boxl]
    kc2 = k_cut * k_cut
    n = {"x": 0, "xz": 0, "xy": 0, "xyz": 0}
    for kx in range(k_max + 1):
        r0 = bk[0] * kx
        ksq = r0 * r0
        if 0.0 < ksq < kc2:
            n["x"] += 1
        for kz in range(1, k_max + 1):
            r2 = bk[2] * kz
            if r0 * r0 + r2 * r2 > kc2:
                break
            n["xz"] += 1
        for ky in range(1, k_max + 1):
            r1 = bk[1] * ky
            if r0 * r0 + r1 * r1 > kc2:
                break
            n["xy"] += 1
            for kz in range(1, k_max + 1):
                r2 = bk[2] * kz
                if r0 * r0 + r1 * r1 + r2 * r2 > kc2:
                    break
                n["xyz"] += 1
    return n


def bead_eval_flops(n_mols, boxl, rcut, k_cut, k_max):
    """Algorithmic FLOPs of one single-bead evaluation, by kernel."""
    n_at = 3 * n_mols
    vol = boxl[0] * boxl[1] * boxl[2]
    n = kgroup_counts(boxl, k_cut, k_max)
    n_ew = 0.5 * n_at * (n_at - 1) * (4.0 * math.pi * rcut ** 3 / 3.0) / vol
    n_lj = 0.5 * n_mols * (n_mols - 1) * (4.0 * math.pi * cfg.R_CUT ** 3 / 3.0) / vol
    table = n_at * (14 + 20 * k_max)
    ksum = n_at * (24 * n["xyz"] + 12 * (n["xz"] + n["xy"]) + 2 * n["x"]) + 0.5 * table
    kforce = n_at * (27 * n["xyz"] + 12 * (n["xz"] + n["xy"]) + 8 * n["x"]) + 0.5 * table
    out = {
        "realspace": 32.0 * n_ew,
        "ksum": float(ksum),
        "kforce": float(kforce),
        "ljintra": 27.0 * n_lj + 205.0 * n_mols,
        "n_ew": n_ew,
        "n_lj": n_lj,
        "kgroups": n,
    }
    out["total"] = out["realspace"] + out["ksum"] + out["kforce"] + out["ljintra"]
    return out
