"""Lossless checkpoint of the resident state: a run stopped, saved, loaded
into a fresh context and continued equals the uninterrupted run bit for bit
(positions, momenta, forces, displacements, lists, heat, rebuild count).

The reference restarts from one set of positions and copies bead 0 to every
bead (initialise.cpp:107-135), so it has no state to compare with: the check
is against the library's own uninterrupted trajectory."""
import sys

import numpy as np
import pytest

from qtip4pf_water_amd import config as cfg

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _state(n_side, L, P, seed, T=298.0):
    q0, box = cfg.lattice_water_box(n_side, L, seed=seed, jitter_bohr=0.05)
    n = n_side ** 3
    beta_n = 1.0 / (cfg.K2AU * T) / P
    q = cfg.free_ring_polymer_beads(q0, P, beta_n, seed=seed + 1) if P > 1 else q0
    p = cfg.sample_momenta(n, P, beta_n, seed=seed + 2)
    return q, p, box, n


def _snapshot(w):
    q, p, F, fph, disp = w.get_state()
    point, lst = w.get_lists()
    s = w.scalars()
    return dict(q=q, p=p, F=F, fph=fph, disp=disp, point=point, lst=lst,
                heat=s["heat"], rebuilds=s["rebuilds"], energy=s["energy"], sc_pot=s["sc_pot"])


def _assert_same(a, b):
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


@pytest.mark.parametrize("P,is_nvt", [(8, True), (8, False), (1, False)])
def test_checkpoint_continues_bit_for_bit(tmp_path, P, is_nvt):
    from qtip4pf_water_amd import Water
    say("checkpoint 216 P", P, "nvt", is_nvt)
    q0, p0, box, n = _state(6, 18.64, P, 11, T=600.0)
    # long enough at 600 K and 0.5 fs for molecules to leave the list skin
    # after the stop, so the restored displacements decide the next rebuild
    n1, n2 = 300, 300
    kw = dict(T_K=600.0, dtfs=0.5)
    with Water(n, P, box, **kw) as w:
        w.set_seed(42)
        w.set_state(q0, p0)
        w.run(n1 + n2, is_nvt=is_nvt)
        ref = _snapshot(w)
    path = tmp_path / "state.qtp"
    with Water(n, P, box, **kw) as w:
        w.set_seed(42)
        w.set_state(q0, p0)
        w.run(n1, is_nvt=is_nvt)
        mid = _snapshot(w)
        w.save_checkpoint(path)
    with Water(n, P, box, **kw) as w:
        w.set_seed(7)  # replaced by the checkpoint's seed
        w.load_checkpoint(path)
        _assert_same(_snapshot(w), mid)
        w.run(n2, is_nvt=is_nvt)
        got = _snapshot(w)
    assert ref["rebuilds"] > mid["rebuilds"]  # lists were rebuilt after the restart
    _assert_same(got, ref)


def test_checkpoint_rejects_another_system(tmp_path):
    from qtip4pf_water_amd import QtpError, Water
    say("checkpoint mismatch")
    q0, p0, box, n = _state(4, 12.43, 4, 3)
    path = tmp_path / "s.qtp"
    with Water(n, 4, box) as w:
        w.set_state(q0, p0)
        w.save_checkpoint(path)
    with Water(n, 8, box) as w:
        with pytest.raises(QtpError):
            w.load_checkpoint(path)
    with Water(n, 4, [b * 1.01 for b in box]) as w:
        with pytest.raises(QtpError):
            w.load_checkpoint(path)
    # same system, another temperature, time step or Ewald accuracy
    for kw, word in ((dict(T_K=300.0), "T_K"), (dict(dtfs=0.5), "dtfs"),
                     (dict(eps_ewald=1e-7), "eps_ewald")):
        with Water(n, 4, box, **kw) as w:
            with pytest.raises(QtpError, match=word):
                w.load_checkpoint(path)
    with Water(n, 4, box) as w:
        with pytest.raises(QtpError):
            w.save_checkpoint(tmp_path / "none.qtp")  # no state yet
