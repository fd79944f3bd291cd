"""Launch configurations and the autotuner (qtp_set_tuning, qtp_autotune; no
counterpart in the reference).  Every configuration the autotuner may pick
must give the reference's forces: each candidate is set on a resident
216 x 4 state and its Suzuki-Chin forces are checked against the oracle's
force_eval_sc_nospline (forces.cpp:698-789) at the usual tolerances."""
import itertools
import sys

import numpy as np
import pytest

from qtip4pf_water_amd import config as cfg

from energy_tol import single

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def rel_atom_err(F, Fr):
    d = np.linalg.norm((F - Fr).reshape(-1, 3), axis=1)
    n = np.linalg.norm(Fr.reshape(-1, 3), axis=1)
    return float(np.max(d / np.maximum(n, 1e-300)))


def _state(P, seed):
    q0, box = cfg.lattice_water_box(6, 18.64, seed=seed, jitter_bohr=0.05)
    beta_n = 1.0 / (cfg.K2AU * 298.0) / P
    q = cfg.free_ring_polymer_beads(q0, P, beta_n, seed=seed + 1)
    p = cfg.sample_momenta(216, P, beta_n, seed=seed + 2)
    return q, p, box


def test_every_configuration_matches_oracle(oracle_mod):
    from qtip4pf_water_amd import Water
    P = 4
    q, p, box = _state(P, 31)
    o = oracle_mod.Oracle(216, P, box)
    point, lst, _ = o.all_lists(q)
    Fr, fr = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
    with oracle_mod.exact_energy():
        er, scr = o.force_eval_sc(q, Fr, fr, point, lst)
    n = 0
    with Water(216, P, box) as w:
        w.set_state(q, p)
        for rsw, lanes, mk, reg, small, usplit in itertools.product(
                (4, 2), (16, 64), (1, 0), (0, 1), (0, 1), (1, 2, 4)):
            if not mk and (small or usplit > 1):
                continue
            say("tuning", rsw, lanes, mk, reg, small, usplit)
            w.set_tuning(lanes=lanes, matrix_kspace=mk, rs_registers=reg,
                         kforce_16_waves=small if mk else -1,
                         ksum_blocks_per_chunk=usplit if mk else 0, rs_waves=rsw)
            _, _, F, f, _ = w.get_state()
            s = w.scalars()
            assert s["status"] == 0
            assert rel_atom_err(f, fr) < 1e-10
            assert rel_atom_err(F, Fr) < 1e-9     # odd beads: the FD Hessian term
            assert single(s["energy"], er) and single(s["sc_pot"], scr)
            n += 1
    assert n == 2 * 2 * 2 * (2 * 3 + 1)


def test_autotune_caches_and_runs():
    from qtip4pf_water_amd import Water
    P = 8
    q, p, box = _state(P, 41)
    say("autotune 216 x", P)
    with Water(216, P, box) as w:
        w.set_seed(1)
        w.set_state(q, p)
        r = w.autotune(retime=True)
        assert r["timed"] >= 8 and r["best_ms"] <= r["default_ms"] * 1.001
        assert {k: w.tuning()[k] for k in r["chosen"]} == r["chosen"]
        w.run(20, is_nvt=True)
        s = w.scalars()
        assert s["status"] == 0 and s["steps"] == 20
        q1 = w.get_state()[0]
    with Water(216, P, box) as w:
        w.set_seed(1)
        w.set_state(q, p)
        r2 = w.autotune()
        assert r2["timed"] == 0 and r2["chosen"] == r["chosen"]
        w.run(20, is_nvt=True)
        assert np.array_equal(w.get_state()[0], q1)


@pytest.mark.parametrize("rsw", [4, 2])
def test_realspace_split_is_bitwise_neutral(rsw):
    """Real-space tiles shared between the two streams (rs_split eighths on
    the second): every tile is computed the same way into its own slots and
    partials, so forces, energies and a resident run agree bit for bit with
    the unsplit launch."""
    from qtip4pf_water_amd import Water
    P = 4
    q, p, box = _state(P, 81)
    out = []
    for sp in range(5):
        say("real-space split", rsw, sp)
        with Water(216, P, box) as w:
            w.set_thermostat("pile_l", 100.0)
            w.set_seed(9)
            w.set_tuning(rs_waves=rsw, rs_split=sp)
            w.set_state(q, p)
            _, _, F, f, _ = w.get_state()
            e0 = w.scalars()["energy"]
            w.run(10, is_nvt=True)
            s = w.scalars()
            assert s["status"] == 0
            out.append((F, f, e0, w.get_state()[0], s["energy"]))
    for o in out[1:]:
        assert np.array_equal(o[0], out[0][0]) and np.array_equal(o[1], out[0][1])
        assert o[2] == out[0][2] and o[4] == out[0][4]
        assert np.array_equal(o[3], out[0][3])


def test_vector_only_system_rejects_matrix_kspace():
    """k_max 24: 2 K = 50 z-factor entries, beyond the matrix-core kernels'
    twelve k-steps."""
    from qtip4pf_water_amd import Water, QtpError
    q, p, box = _state(2, 51)
    L = float(box[0])
    ew = {"kappa": 0.3 / cfg.ANGS2BOHR, "rcut": 9.0 * cfg.ANGS2BOHR,
          "k_cut": 2.0 * np.pi * 24.0 / L * (1.0 + 1e-9), "k_max": 24}
    with Water(216, 2, box, ewald=ew) as w:
        assert w.tuning()["matrix_kspace"] == 0
        with pytest.raises(QtpError):
            w.set_tuning(matrix_kspace=1)


@pytest.mark.parametrize("thermo,mode,nvt", [
    ("pile_l", "pimd", True), (None, "pimd", False), ("pile_l", "pacmd", True),
    ("pile_l", "centroid_fixed", True), ("pile_g", "pimd", True)])
def test_fused_propagation_is_the_two_kernel_path(thermo, mode, nvt):
    """k_prop_tile (one LDS-staged kernel per step) against k_prop_fwd +
    k_prop_inv: the same arithmetic in the same order, so positions,
    momenta and the thermostat heat agree bit for bit (with PILE-G's global
    centroid rescale both settings take the two kernels)."""
    from qtip4pf_water_amd import Water
    P = 8
    q, p, box = _state(P, 61)
    say("fused propagation", thermo, mode, nvt)
    out = []
    for fused in (1, 0):
        with Water(216, P, box) as w:
            if thermo:
                w.set_thermostat(thermo, 100.0)
            w.set_pimd_mode(mode)
            w.set_tuning(fused_propagation=fused)
            w.set_seed(4)
            w.set_state(q, p)
            w.run(20, is_nvt=nvt)
            s = w.scalars()
            assert s["status"] == 0
            qq, pp = w.get_state()[:2]
            out.append((qq, pp, s["heat"]))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]
