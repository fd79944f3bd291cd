"""NVE energy drift of the GPU trajectories is no worse than the reference
algorithm's: the same start integrated by the library and by the oracle (the
reference's operation order, on the host), the conserved quantity sampled
along both.  The trajectories part after a few hundred steps (chaos), so the
comparison is of the size of the excursions, not of the values.

64 molecules in the 18.64 A box (the 216-molecule box's edge, so the 9 A LJ
cutoff stays below half the box) keep the oracle's runs to seconds."""
import sys

import numpy as np
import pytest

from qtip4pf_water_amd import config as cfg

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _start(P, seed):
    q0, box = cfg.water_box(64, 18.64, seed=seed, jitter_bohr=0.05)
    beta_n = 1.0 / (cfg.K2AU * 298.0) / P
    q = cfg.replicate_beads(q0, P)
    p = cfg.sample_momenta(64, P, beta_n, seed=seed + 1)
    return np.ascontiguousarray(q), p, box


def _excursion(e):
    e = np.asarray(e)
    return float(np.max(np.abs(e - e[0])))


def test_classical_nve_drift_no_worse_than_reference(oracle_mod):
    from qtip4pf_water_amd import Water
    n_steps, every = 2000, 20
    say("classical NVE drift 64 mols", n_steps, "steps")
    q, p, box = _start(1, 71)
    with Water(64, 1, box) as w:
        w.set_state(q, p)
        c0 = w.observables()["conserved"]
        tr = w.run_observed(n_steps, is_nvt=False, every=every)
    gpu = np.concatenate([[c0], tr[:, 1]])  # the start, then every `every` steps
    o = oracle_mod.Oracle(64, 1, box)
    m = o.mass
    qq, pp = q.copy(), p.copy()
    point, lst, disp = o.all_lists(qq)
    F = np.zeros(o.n_dof)
    e = o.get_forces(qq, F, point, lst, 0)
    ref = [e + 0.5 * np.sum(pp * pp / m)]
    for s in range(1, n_steps + 1):
        e, _ = o.update_class(qq, pp, F, point, lst, disp)
        if s % every == 0:
            ref.append(e + 0.5 * np.sum(pp * pp / m))
    ref = np.array(ref)
    assert np.all(np.isfinite(gpu))
    assert abs(gpu[0] - ref[0]) <= 1e-10 * abs(ref[0])      # same start
    ke = 0.5 * np.sum(p * p / m)
    dg, dr = _excursion(gpu), _excursion(ref)
    say(f"classical excursion gpu {dg:.3e} ref {dr:.3e} Eh (KE {ke:.3e})")
    assert dg <= 1.5 * dr + 1e-6 * ke


def test_sc_pimd_nve_drift_no_worse_than_reference(oracle_mod):
    from qtip4pf_water_amd import Water
    P, n_steps, every = 8, 400, 10
    say("SC-PIMD NVE drift 64 mols P", P, n_steps, "steps")
    q, p, box = _start(P, 73)
    with Water(64, P, box) as w:
        w.set_state(q, p)
        c0 = w.observables()["conserved"]
        tr = w.run_observed(n_steps, is_nvt=False, every=every)
    gpu = np.concatenate([[c0], tr[:, 1]])
    o = oracle_mod.Oracle(64, P, box)
    qq, pp = q.copy(), p.copy()
    point, lst, disp = o.all_lists(qq)
    F, fph = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
    e, _ = o.force_eval_sc(qq, F, fph, point, lst, nthreads=8)
    ref = [o.get_conserved(qq.copy(), pp.copy(), e)]
    h = 0.0
    for s in range(1, n_steps + 1):
        e, h, _, _ = o.update_sc(qq, pp, F, fph, point, lst, disp, False, None, h, nthreads=8)
        if s % every == 0:
            ref.append(o.get_conserved(qq.copy(), pp.copy(), e))
    ref = np.array(ref)
    assert np.all(np.isfinite(gpu))
    assert abs(gpu[0] - ref[0]) <= 1e-10 * abs(ref[0])
    dg, dr = _excursion(gpu), _excursion(ref)
    ke = 0.5 * np.sum(p * p / np.tile(o.mass, P))
    say(f"SC-PIMD excursion gpu {dg:.3e} ref {dr:.3e} Eh (KE {ke:.3e})")
    assert dg <= 1.5 * dr + 1e-6 * ke


def test_long_nvt_run_stays_bounded():
    """216 x 32 SC-PIMD NVT, 20,000 steps (5 ps) of the benchmark workload's
    start: no list overflow, lists rebuilt as molecules move, and the
    observables stay finite and within the band of their own early values
    (K_cv within 10 %, the potential energy within 10 Eh of the SC sum)."""
    from qtip4pf_water_amd import Water
    n_side, L, P = 6, 18.64, 32
    n = n_side ** 3
    say("long NVT run", n, "mols P", P, "20000 steps")
    q0, box = cfg.lattice_water_box(n_side, L, seed=2024)
    beta_n = 1.0 / (cfg.K2AU * 298.0) / P
    q = cfg.free_ring_polymer_beads(q0, P, beta_n, seed=2025)
    p = cfg.sample_momenta(n, P, beta_n, seed=2026)
    with Water(n, P, box) as w:
        w.set_seed(77)
        w.set_state(q, p)
        w.run(2000, is_nvt=True)  # leave the lattice start
        tr = w.run_observed(18000, is_nvt=True, every=1000)
        s = w.scalars()
    assert s["status"] == 0 and s["steps"] == 20000
    assert s["rebuilds"] > 100
    assert np.all(np.isfinite(tr))
    kcv, energy = tr[:, 0], tr[:, 2]
    say(f"K_cv {kcv.min():.4f}..{kcv.max():.4f} Eh, V {energy.min():.3f}..{energy.max():.3f} Eh, "
        f"{s['rebuilds']:.0f} rebuilds")
    assert np.all(np.abs(kcv / kcv[0] - 1.0) < 0.1)
    assert np.all(np.abs(energy - energy[0]) < 10.0)
