"""Device-side observables of the resident mode (K_cv, the ring-polymer
conserved quantity) against the oracle's get_kcv / get_conserved
(path_integrals.cpp:29-81) on the downloaded state, and the on-device trace
of qtp_run_observed against sampling step by step."""
import sys

import numpy as np
import pytest

from qtip4pf_water_amd import config as cfg

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _state(n_side, L, P, seed):
    q0, box = cfg.lattice_water_box(n_side, L, seed=seed, jitter_bohr=0.05)
    n = n_side ** 3
    beta_n = 1.0 / (cfg.K2AU * 298.0) / P
    q = cfg.free_ring_polymer_beads(q0, P, beta_n, seed=seed + 1) if P > 1 else q0
    p = cfg.sample_momenta(n, P, beta_n, seed=seed + 2)
    return q, p, box, n


def test_observables_match_oracle(oracle_mod):
    from qtip4pf_water_amd import Water
    P = 8
    say("resident observables 216 P", P)
    q0, p0, box, n = _state(6, 18.64, P, 5)
    o = oracle_mod.Oracle(n, P, box)
    with Water(n, P, box) as w:
        w.set_seed(3)
        w.set_state(q0, p0)
        w.run(3, is_nvt=True)
        ob = w.observables()
        sc = w.scalars()
        q, p, F, fph, _ = w.get_state()
    assert ob["energy"] == sc["energy"] and ob["sc_pot"] == sc["sc_pot"]
    kr = o.get_kcv(q, fph)
    cr = o.get_conserved(q.copy(), p.copy(), sc["energy"])
    assert abs(ob["kcv"] - kr) <= 1e-12 * abs(kr)
    assert abs(ob["conserved"] - cr) <= 1e-12 * abs(cr)


def test_classical_observables():
    from qtip4pf_water_amd import Water
    say("resident observables classical 216")
    q0, p0, box, n = _state(6, 18.64, 1, 9)
    with Water(n, 1, box) as w:
        w.set_state(q0, p0)
        w.run(4)
        ob = w.observables()
        e = w.scalars()["energy"]
        _, p, _, _, _ = w.get_state()
    m = np.tile(np.repeat([cfg.MO, cfg.MH, cfg.MH], 3), n)
    ke = 0.5 * np.sum(p * p / m)
    assert ob["kcv"] == pytest.approx(ke, rel=1e-13)
    assert ob["conserved"] == pytest.approx(e + ke, rel=1e-13)


def test_run_observed_trace_equals_stepwise_samples():
    from qtip4pf_water_amd import Water
    P = 4
    say("run_observed trace 64 P", P)
    q0, p0, box, n = _state(4, 12.43, P, 13)
    with Water(n, P, box) as w:
        w.set_seed(11)
        w.set_state(q0, p0)
        tr = w.run_observed(6, is_nvt=True, every=2)
    assert tr.shape == (3, 4)
    ref = []
    with Water(n, P, box) as w:
        w.set_seed(11)
        w.set_state(q0, p0)
        for _ in range(3):
            w.run(2, is_nvt=True)
            o = w.observables()
            ref.append([o["kcv"], o["conserved"], o["energy"], o["sc_pot"]])
    assert np.array_equal(tr, np.array(ref))
