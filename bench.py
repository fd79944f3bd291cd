"""Throughput of the qTIP4P/F Suzuki-Chin PIMD step on one MI355X, and the
speed-up of learned-correction CMD over plain PIMD-based CMD.

Workload (default): 216 H2O in a cubic 18.64 A box, 32 beads, 298 K, FP64,
dt = 0.25 fs, PILE-L centroid thermostat (tau = 100 fs; the other normal
modes critically damped, thermostat.cpp:114-129), Ewald alpha = 0.3 / A,
r_c = 9 A, k = 2 pi n / L with |n| <= 7, LJ cutoff 9 A, seeded lattice start
with random orientations, beads = centroid + free-ring-polymer normal-mode
displacements.  One "step" is one full update_sc_nospline (update.cpp:271-401):
64 single-bead force evaluations plus propagation, thermostat and list checks,
state resident in HBM, replayed as one HIP graph per step.

The surrogate_cmd object times, on the same GPU, CMD of 512 H2O at 298 K with
the learned centroid-force correction (classical dynamics on qTIP4P/F plus
the network, dt 0.5 fs) against partially adiabatic CMD of the 32-bead ring
polymer (dt 0.1 fs), both with the Langevin centroid thermostat, and states
the integrated g_OO(r) / g_OH(r) error of a 20 ps surrogate run against the
20 ps plain-CMD g(r) shipped in qtip4pf_water_amd/data/rdf_pacmd_512x32.npz
(made by `python -m qtip4pf_water_amd.cmd validate` on an MI355X).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
    python bench.py --ewald derived                               # forces.cpp:574-584 parameters
    python bench.py --workload batch [--replicas 64] [--steps K]  # batched replicas
    python bench.py --workload c1                                 # classical 216 NVE
    python bench.py --workload c5 --steps 20 --warmup 3           # 4096 H2O x 32 beads

Prints one JSON line.  Needs the built library (python -m qtip4pf_water_amd.build).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from qtip4pf_water_amd import config as cfg  # noqa: E402
from qtip4pf_water_amd import flops  # noqa: E402

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 (vector and matrix), AMD published
KERNELS = ["realspace", "ksum", "sreduce", "kforce", "ljintra", "sites", "finalize", "sc1",
           "sc2", "prop", "prop_a", "prop_b", "check", "nlist"]


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (profiles/pmc_traffic_sc_pimd_216x32.txt: FETCH_SIZE and WRITE_SIZE passes
    of this workload, calibrated against a copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_sc_pimd_216x32.txt")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("TRAFFIC_BYTES_PER_LAUNCH"):
                    kv = dict(t.split("=") for t in line.split()[1:])
                    return float(kv[kernel]) if kernel in kv else None
    except OSError:
        return None
    return None


def workload(name):
    if name == "c2":
        return dict(n_side=6, L_angs=18.64, P=32, T_K=298.0, dtfs=0.25, nvt=True, n_max=7,
                    thermostat=("pile_l", 100.0), rebuild_every=0,
                    label="SC-PIMD NVT 216 H2O x 32 beads (PILE-L, tau 100 fs), 298 K, "
                          "dt 0.25 fs")
    if name == "c1":
        return dict(n_side=6, L_angs=18.64, P=1, T_K=298.0, dtfs=0.25, nvt=False, n_max=7,
                    thermostat=None, rebuild_every=0,
                    label="classical NVE 216 H2O, velocity Verlet, dt 0.25 fs")
    if name == "c5":
        return dict(n_side=16, L_angs=49.72, P=32, T_K=298.0, dtfs=0.25, nvt=True, n_max=18,
                    thermostat=("pile_l", 100.0), rebuild_every=10,
                    label="SC-PIMD NVT 4096 H2O x 32 beads (PILE-L, tau 100 fs), 298 K, "
                          "dt 0.25 fs, cell lists rebuilt every 10 steps")
    raise SystemExit(f"unknown workload {name}")


def ewald_point(wl, box, mode):
    """The stated Ewald point (alpha = 0.3 / A, r_c = 9 A, |n| <= n_max), or
    None for the reference's derivation from eps_ewald (forces.cpp:574-584)."""
    if mode == "derived":
        return None
    L = float(box[0])
    return {"kappa": 0.3 / cfg.ANGS2BOHR, "rcut": 9.0 * cfg.ANGS2BOHR,
            "k_cut": 2.0 * np.pi * wl["n_max"] / L * (1.0 + 1e-9), "k_max": wl["n_max"]}


def initial_state(wl, seed=2024):
    q0, box = cfg.lattice_water_box(wl["n_side"], wl["L_angs"], seed=seed)
    n_mols = wl["n_side"] ** 3
    beta = 1.0 / (cfg.K2AU * wl["T_K"])
    beta_n = beta / wl["P"]
    q = cfg.free_ring_polymer_beads(q0, wl["P"], beta_n, seed=seed + 1)
    p = cfg.sample_momenta(n_mols, wl["P"], beta_n, seed=seed + 2)
    return q, p, box, n_mols


def cpu_baseline(wl, q, p, box, n_mols, seconds, ew):
    """The oracle (our C restatement with the reference's OpenMP-over-beads
    structure, forces.cpp:716; its thermostat is the reference's pile_g) on
    the host cores, bounded sample."""
    import oracle
    oracle.build()
    threads = max(1, min(16, os.cpu_count() or 1))
    o = oracle.Oracle(n_mols, wl["P"], box, T_K=wl["T_K"], dtfs=wl["dtfs"], ewald=ew)
    qq, pp = q.copy(), p.copy()
    point, lst, disp = o.all_lists(qq)
    P = wl["P"]
    F, fph = np.zeros(P * o.n_dof), np.zeros(P * o.n_dof)
    rng = np.random.default_rng(7)
    if P > 1:
        o.force_eval_sc(qq, F, fph, point, lst, nthreads=threads)
    else:
        o.get_forces(qq, F, point, lst, 0)
    n, t0 = 0, time.perf_counter()
    heat = 0.0
    while True:
        if P > 1:
            dev = rng.standard_normal(o.n_deviates()) if wl["nvt"] else None
            if dev is not None:
                dev[2] = 0.5 * rng.chisquare(3 * o.n_atoms - 2)
            _, heat, _, _ = o.update_sc(qq, pp, F, fph, point, lst, disp, wl["nvt"], dev, heat,
                                        nthreads=threads)
        else:
            o.update_class(qq, pp, F, point, lst, disp)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds and n >= 2:
            break
    return {"value": n / el, "unit": "steps/s", "cores": threads if P > 1 else 1,
            "kind": "port",
            "sample": f"{n} consecutive steps of the same workload on the host "
                      f"({'OpenMP over beads, ' + str(threads) + ' threads' if P > 1 else '1 thread'}"
                      f"{', pile_g thermostat' if P > 1 and wl['nvt'] else ''}), {el:.1f} s"}


def run_batch(a):
    """Replica-steps/s of a batch; value = total replica steps per second."""
    from qtip4pf_water_amd import Batch, Water
    reps = cfg.batch_replicas(a.replicas, 2024)
    ws = []
    for r in reps:
        q, p, box = cfg.replica_state(r)
        w = Water(r["n_mols"], r["P"], box, T_K=r["T_K"], dtfs=0.25, eps_ewald=1e-6,
                  lj_rcut=r["lj_rcut"])
        w.set_seed(r["seed"])
        try:
            w.set_state(q, p)
        except Exception:
            print("replica", len(ws), r, q.shape, bool(np.isfinite(q).all()),
                  w.n_beads, w.n_dof, file=sys.stderr, flush=True)
            raise
        ws.append(w)
    tuned = 0
    t_tune = time.perf_counter()
    if not a.no_autotune:
        for w in ws:
            tuned += w.autotune()["timed"] > 0
    t_tune = time.perf_counter() - t_tune
    b = Batch(ws, True)
    b.run(a.warmup)
    b.sync()
    t0 = time.perf_counter()
    b.run(a.steps)
    b.sync()
    el = time.perf_counter() - t0
    status = [w.scalars()["status"] for w in ws]
    b.close()
    bead_mol = sum(r["n_mols"] * r["P"] for r in reps)
    res = {
        "metric": "replica steps/s, 64 SC-PIMD replicas (batched)",
        "value": a.steps * len(reps) / el,
        "unit": "replica-steps/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * el / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded lattice starts, free-ring-polymer beads)",
        "config": {"workload": "batched replicas: N in {64..512} H2O, P in {8..64}, T in 250-350 K",
                   "replicas": len(reps), "molecule_bead_total": bead_mol,
                   "parallelism": "single GPU"},
        "molecule_bead_steps_per_s": bead_mol * a.steps / el,
        "status_max": int(max(status)),
        "autotune": None if a.no_autotune else {"shapes_timed": tuned,
                                                 "seconds": round(t_tune, 2)},
    }
    for w in ws:
        w.close()
    print(json.dumps(res), flush=True)


def other_point(a, wl, q, p, box, n_mols):
    """The same workload with the reference's Ewald derivation from eps_ewald
    = 1e-6 (forces.cpp:574-584: more k-vectors than the stated point), timed
    the same way: steps/s."""
    from qtip4pf_water_amd import Water
    w = Water(n_mols, wl["P"], box, T_K=wl["T_K"], dtfs=wl["dtfs"], eps_ewald=1e-6,
              epsilon_sc=1e-5)
    d = w.derived
    if wl["thermostat"]:
        w.set_thermostat(*wl["thermostat"])
    w.set_seed(12345)
    w.set_state(q, p)
    if not a.no_autotune:
        w.autotune()
    w.run(a.warmup, is_nvt=wl["nvt"])
    w.sync()
    t0 = time.perf_counter()
    w.run(a.steps, is_nvt=wl["nvt"])
    w.sync()
    el = time.perf_counter() - t0
    st = w.scalars()["status"]
    w.close()
    return {"value": a.steps / el, "unit": "steps/s", "steps": a.steps,
            "ewald": {"kappa_per_angs": d["kappa"] * cfg.ANGS2BOHR,
                      "rcut_angs": d["rcut"] / cfg.ANGS2BOHR, "k_max": d["k_max"],
                      "k_groups": d["n_kgroups"]},
            "status": int(st)}


def surrogate_cmd(ps=20.0, pacmd_ps=0.2):
    """Live CMD legs at 512 H2O, 298 K, PILE-L centroid thermostat (tau
    100 fs): the learned correction on classical dynamics (dt 0.5 fs, ps of
    g(r) sampling after 2 ps) and classical MD without it, against
    partially adiabatic CMD of 32 beads (dt 0.1 fs; throughput from a short
    run, g(r) from the shipped 20 ps reference run)."""
    from qtip4pf_water_amd import cmd
    ref_path = os.path.join(cmd.DATA_DIR, "rdf_pacmd_512x32.npz")
    model = cmd.load_model()
    with np.load(ref_path) as z:
        ref = {k: z[k] for k in ("r", "g_oo", "g_oh")}
        ref_meta = {k: z[k].item() for k in ("steps", "dt_fs", "ns_per_day", "n_samples")}
    pa = cmd.cmd_run("pacmd", ps=pacmd_ps, equil_ps=0.02)
    sg = cmd.cmd_run("surrogate", model=model, ps=ps, equil_ps=2.0)
    cl = cmd.cmd_run("classical", ps=ps, equil_ps=2.0)
    err = {k: {"g_oo": cmd.rdf_error(r["g_oo"], ref["g_oo"], ref["r"]),
               "g_oh": cmd.rdf_error(r["g_oh"], ref["g_oh"], ref["r"])}
           for k, r in (("surrogate", sg), ("classical", cl))}
    # the same surrogate run with longer steps (reported beside the 0.5 fs
    # headline: the g(r) criterion holds, the O-H stretch is resolved more coarsely)
    longer = []
    for dt in (0.75, 1.0):
        sl = cmd.cmd_run("surrogate", model=model, ps=ps, equil_ps=2.0, dt_fs=dt)
        longer.append({"dt_fs": dt, "steps_per_s": sl["steps_per_s"],
                       "speedup_ns_per_day": sl["ns_per_day"] / pa["ns_per_day"],
                       "g_oo_error": cmd.rdf_error(sl["g_oo"], ref["g_oo"], ref["r"]),
                       "g_oh_error": cmd.rdf_error(sl["g_oh"], ref["g_oh"], ref["r"]),
                       "status": int(sl["status"])})
    return {
        "metric": "surrogate-CMD speedup at matched g_OO(r)",
        "speedup_ns_per_day": sg["ns_per_day"] / pa["ns_per_day"],
        "unit": "x (ns/day ratio)",
        "surrogate": {"steps_per_s": sg["steps_per_s"], "ns_per_day": sg["ns_per_day"],
                      "dt_fs": sg["dt_fs"], "g_oo_error": err["surrogate"]["g_oo"],
                      "g_oh_error": err["surrogate"]["g_oh"], "sample_ps": ps,
                      "status": int(sg["status"])},
        "pacmd": {"steps_per_s": pa["steps_per_s"], "ns_per_day": pa["ns_per_day"],
                  "dt_fs": pa["dt_fs"], "beads": 32, "timed_steps": int(pa["steps"]),
                  "status": int(pa["status"])},
        "surrogate_longer_steps": longer,
        "classical_without_correction": {"ns_per_day": cl["ns_per_day"],
                                         "g_oo_error": err["classical"]["g_oo"],
                                         "g_oh_error": err["classical"]["g_oh"]},
        "g_error": "integral of |g - g_ref| over integral of g_ref, 0-8 A",
        "g_ref": {"file": "qtip4pf_water_amd/data/rdf_pacmd_512x32.npz",
                  "run": f"PA-CMD 512 H2O x 32 beads, {ref_meta['steps']} steps of "
                         f"{ref_meta['dt_fs']} fs, {ref_meta['n_samples']} samples"},
        "config": "512 H2O, 24.8 A box, 298 K, LJ and Ewald as the reference derives them",
    }


def copy_calibration(mib):
    """A plain device-to-device copy of `mib` MiB (the runtime's copy kernel),
    three times: under rocprofv3 --pmc its FETCH_SIZE / WRITE_SIZE per
    dispatch calibrate the counters against a known byte count."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = int(mib) << 20
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(n)) or hip.hipMalloc(C.byref(b), C.c_size_t(n)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(n))
    for _ in range(3):
        hip.hipMemcpy(b, a, C.c_size_t(n), 3)   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="c2", choices=["c2", "c1", "c5", "batch"])
    ap.add_argument("--ewald", default="stated", choices=["stated", "derived"],
                    help="stated: alpha 0.3/A, r_c 9 A, |n| <= 7 (18 at 4096); "
                         "derived: forces.cpp:574-584 from eps_ewald = 1e-6")
    ap.add_argument("--no-cmd", action="store_true", help="skip the surrogate-CMD legs")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch configuration (no qtp_autotune)")
    ap.add_argument("--replicas", type=int, default=64)
    ap.add_argument("--timing-steps", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="after the run, copy this many MiB device-to-device (counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build shards nothing across GPUs "
              "(each replica or bead set runs on one device); run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    if a.workload == "batch":
        run_batch(a)
        return

    from qtip4pf_water_amd import Water

    wl = workload(a.workload)
    q, p, box, n_mols = initial_state(wl)
    ew = ewald_point(wl, box, a.ewald)
    w = Water(n_mols, wl["P"], box, T_K=wl["T_K"], dtfs=wl["dtfs"], eps_ewald=1e-6,
              epsilon_sc=1e-5, ewald=ew)
    d = w.derived
    if wl["thermostat"]:
        w.set_thermostat(*wl["thermostat"])
    if wl["rebuild_every"]:
        w.set_lists("auto", rebuild_every=wl["rebuild_every"])
    w.set_seed(12345)
    w.set_state(q, p)
    tune = None if a.no_autotune else w.autotune()

    # warm-up (also captures the step graph)
    w.run(a.warmup, is_nvt=wl["nvt"])
    w.sync()
    # timed region: K graph replays, bracketed by device synchronisation
    t0 = time.perf_counter()
    w.run(a.steps, is_nvt=wl["nvt"])
    w.sync()
    el = time.perf_counter() - t0
    sps = a.steps / el
    sc = w.scalars()

    # per-kernel durations: HIP events on the library's stream around every
    # launch, over two further blocks of steps launched without the graph;
    # per kernel the block with the lower average (a host stall between an
    # event and its launch inflates one block, not the kernel)
    kt = {}
    for _ in range(2):
        w.enable_timing(True)
        w.run(max(1, a.timing_steps // 2), is_nvt=wl["nvt"])
        w.sync()
        for k in KERNELS:
            ms, n = w.timing(k)
            if n and (k not in kt or 1e3 * ms / n < kt[k]["us_avg"]):
                kt[k] = {"ms_total": ms, "launches": n, "us_avg": 1e3 * ms / n}
        w.enable_timing(False)

    fl = flops.bead_eval_flops(n_mols, box, d["rcut"], d["k_cut"], d["k_max"])
    evals_per_launch = wl["P"]  # every force-phase kernel launch covers P evaluations
    dom = max((k for k in ("realspace", "ksum", "kforce", "ljintra") if k in kt),
              key=lambda k: kt[k]["ms_total"])
    achieved = fl[dom] * evals_per_launch / (kt[dom]["us_avg"] * 1e-6) / 1e12
    step_flops = 2 * wl["P"] * fl["total"] if wl["P"] > 1 else fl["total"]

    metric = (f"PIMD steps/s {n_mols} H2O x {wl['P']} beads" if wl["P"] > 1
              else f"MD steps/s {n_mols} H2O")
    if a.workload == "c2":
        metric += "; surrogate-CMD speedup at matched g_OO(r)"
    res = {
        "metric": metric,
        "value": sps,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 / sps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lattice start, free-ring-polymer beads)",
        "config": {
            "workload": wl["label"],
            "n_mols": n_mols, "n_beads": wl["P"], "box_angs": wl["L_angs"],
            "dt_fs": wl["dtfs"], "T_K": wl["T_K"], "parallelism": "single GPU",
            "ewald": {"kappa_per_angs": d["kappa"] * cfg.ANGS2BOHR,
                      "rcut_angs": d["rcut"] / cfg.ANGS2BOHR, "k_max": d["k_max"],
                      "k_groups": d["n_kgroups"], "point": a.ewald},
            "lists": w.list_info(), "rebuild_every": wl["rebuild_every"],
            "launch": w.tuning(),
            "autotune": tune and {"best_ms": tune["best_ms"], "default_ms": tune["default_ms"],
                                  "candidates": tune["timed"]},
        },
        "ns_per_day": sps * wl["dtfs"] * 0.0864,
        "algorithmic_tflops": step_flops * sps / 1e12,
        "roofline": {
            # FP64 vector ALU work (no matrix instructions in this kernel);
            # MI355X's FP64 vector and matrix peaks are both 78.6 TFLOP/s
            "bound": "fp64_valu",
            "kernel": dom,
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": measured_traffic(dom) if a.workload == "c2" else None,
            "flops_per_launch": fl[dom] * evals_per_launch,
            "us_per_launch": kt[dom]["us_avg"],
        },
        "kernel_us": {k: round(v["us_avg"], 2) for k, v in kt.items()},
        "state": {"energy": sc["energy"], "rebuilds": sc["rebuilds"], "steps": sc["steps"],
                  "status": sc["status"]},
    }
    w.close()
    if a.copy_calib_mib > 0:
        copy_calibration(a.copy_calib_mib)
    if a.workload in ("c2", "c1") and a.ewald == "stated":
        res["derived_point"] = other_point(a, wl, q, p, box, n_mols)
    if a.workload == "c2" and not a.no_cmd:
        res["surrogate_cmd"] = surrogate_cmd()
    if not a.no_cpu:
        res["cpu_baseline"] = cpu_baseline(wl, q, p, box, n_mols, a.cpu_seconds, ew)
        res["speedup_vs_cpu"] = sps / res["cpu_baseline"]["value"]
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
