"""Cell-binned Verlet lists (k_cell_build, k_nl_cell) against the oracle's
all-pairs lists (neighbour.cpp:12-62 restated), bit for bit, and the forced
rebuild period of resident runs.  Neither has a counterpart in the
reference (all pairs, displacement-triggered rebuilds only); the lists must
come out identical to its, whatever the coordinates: unwrapped positions
several boxes away, O atoms on cell faces and on the periodic boundary."""
import sys

import numpy as np
import pytest

from qtip4pf_water_amd import config as cfg

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _adversarial(q, box, nc, seed):
    """Whole molecules moved: some by whole boxes (unwrapped state), some so
    that their O sits on a cell face or on the box boundary (at 0, at L, a
    hair inside either side).  Pair geometry changes, which a list test
    does not mind."""
    q = q.reshape(-1, 9).copy()
    rng = np.random.default_rng(seed)
    n = q.shape[0]
    idx = rng.permutation(n)
    for m in idx[: n // 8]:                          # shifted by whole boxes
        s = rng.integers(-3, 4, size=3) * np.asarray(box)
        q[m] += np.tile(s, 3)
    L = np.asarray(box)
    for k, m in enumerate(idx[n // 8: n // 8 + 64]):  # O on a face or boundary
        d = k % 3
        w = L[d] / nc[d]
        face = [0.0, L[d], w * (1 + k % (nc[d] - 1)), -0.0][k % 4]
        eps = [0.0, 1e-12, -1e-12, 4e-16 * L[d]][(k // 4) % 4]
        shift = face + eps - q[m, d]
        q[m, d::3] += shift
    return q.reshape(-1)


@pytest.mark.parametrize("n_side,L_angs,lj_angs,expect_S", [
    (16, 49.72, None, 2),      # 4096 H2O, 9 A cutoff: 8 cells a side, S = 2
    (24, 74.58, 4.0, 1),       # 13824 H2O, 4 A cutoff: S = 2 would need 28^3 cells
])
def test_cell_lists_equal_all_pairs(oracle_mod, n_side, L_angs, lj_angs, expect_S):
    from qtip4pf_water_amd import Water
    n = n_side ** 3
    lj = None if lj_angs is None else lj_angs * cfg.ANGS2BOHR
    say("cell lists", n, "mols, lj", lj_angs)
    q0, box = cfg.lattice_water_box(n_side, L_angs, seed=81, jitter_bohr=0.05)
    P = 2
    o = oracle_mod.Oracle(n, P, box, lj_rcut=lj)
    with Water(n, P, box, lj_rcut=lj) as w:
        info = w.list_info()
        assert info["cells"] and info["S"] == expect_S, info
        q = np.concatenate([_adversarial(q0, box, info["grid"], 82),
                            _adversarial(q0, box, info["grid"], 83)])
        point, lst, disp = o.all_lists(q)
        pg, lg, dg = w.new_lists()
        for b in range(P):
            w.neighbour_list(q, pg, lg, dg, b)
        assert np.array_equal(pg, point) and np.array_equal(lg, lst)
        w.set_lists("allpairs")
        pa, la, da = w.new_lists()
        w.neighbour_list(q, pa, la, da, 1)
        assert np.array_equal(pa[n:], point[n:]) and np.array_equal(la[o.list_size:], lst[o.list_size:])


def test_resident_runs_cells_and_all_pairs_identical():
    """1728 H2O x 8 beads, NVT: the cell-binned and the all-pairs build give
    the same lists, so the trajectories are the same bit for bit."""
    from qtip4pf_water_amd import Water
    n, P = 12 ** 3, 8
    say("resident cells vs all pairs", n, "P", P)
    q0, box = cfg.lattice_water_box(12, 37.29, seed=85, jitter_bohr=0.05)
    beta_n = 1.0 / (cfg.K2AU * 298.0) / P
    q = cfg.free_ring_polymer_beads(q0, P, beta_n, seed=86)
    p = cfg.sample_momenta(n, P, beta_n, seed=87)
    out = {}
    for method in ("cells", "allpairs"):
        with Water(n, P, box) as w:
            w.set_lists(method)
            assert w.list_info()["cells"] == (method == "cells")
            w.set_seed(3)
            w.set_state(q, p)
            w.run(25, is_nvt=True)
            s = w.scalars()
            assert s["status"] == 0
            out[method] = w.get_state()[:2] + (w.get_lists(),)
    assert np.array_equal(out["cells"][0], out["allpairs"][0])
    assert np.array_equal(out["cells"][1], out["allpairs"][1])
    assert np.array_equal(out["cells"][2][0], out["allpairs"][2][0])
    assert np.array_equal(out["cells"][2][1], out["allpairs"][2][1])


@pytest.mark.parametrize("method", ["cells", "allpairs"])
def test_forced_rebuild_period(method):
    """rebuild_every = 10: each bead's lists are rebuilt on steps 10, 20, 30
    (no displacement-triggered rebuild happens in 30 steps of this start, as
    the run without the period shows), and the trajectory agrees with the
    one without forced rebuilds to rounding (same pairs within the cutoff,
    other lane assignments in the pair sums)."""
    from qtip4pf_water_amd import Water
    n, P = 12 ** 3, 8
    say("forced rebuild period", method)
    q0, box = cfg.lattice_water_box(12, 37.29, seed=88, jitter_bohr=0.05)
    beta_n = 1.0 / (cfg.K2AU * 298.0) / P
    q = cfg.free_ring_polymer_beads(q0, P, beta_n, seed=89)
    p = cfg.sample_momenta(n, P, beta_n, seed=90)
    res = {}
    for every in (0, 10):
        with Water(n, P, box) as w:
            w.set_lists(method, rebuild_every=every)
            w.set_state(q, p)
            r0 = w.scalars()["rebuilds"]
            w.run(30, is_nvt=False)
            s = w.scalars()
            assert s["status"] == 0
            res[every] = (s["rebuilds"] - r0, w.get_state()[0])
    assert res[0][0] == 0
    assert res[10][0] == 3 * P
    qa, qb = res[0][1], res[10][1]
    assert np.max(np.abs(qa - qb)) <= 1e-9 * np.max(np.abs(qa))


def test_small_box_rejects_cells():
    from qtip4pf_water_amd import Water, QtpError
    q0, box = cfg.lattice_water_box(6, 18.64, seed=1)
    with Water(216, 1, box) as w:
        assert not w.list_info()["cells"]
        with pytest.raises(QtpError):
            w.set_lists("cells")
