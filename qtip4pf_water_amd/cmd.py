"""Centroid molecular dynamics (CMD) of qTIP4P/F water with a learned
centroid-force correction, and plain PIMD-based CMD as its fallback and
judge.  No counterpart in the reference: its CMD tooling ends at the
additive spline hook (forces.cpp:452-466, 634-691) and its README leaves
CMD propagation and g(r) to other programs.

Pipeline (everything on the GPU through the library):
  1. centroid mean forces: a 216 H2O x 32-bead PIMD run (PILE-L, tau = 100
     fs) sampled every `stride` steps; from each sample a run with the
     centroid held fixed (qtp_set_pimd_mode CENTROID_FIXED) averages the force
     on the centroid, (1/P) sum_b F_sc,b, over n_acc steps after n_eq steps;
     the classical qTIP4P/F force at the centroid is subtracted, so the target
     is the correction force F_mean - F_classical;
  2. force matching of the atom-centred BF16 network (qtp_nn_train);
  3. CMD runs at 512 H2O: surrogate CMD = classical dynamics on qTIP4P/F plus
     the correction (dt 0.5 fs); plain CMD = partially adiabatic CMD of the
     32-bead ring polymer (qtp_set_pimd_mode PACMD); classical MD for the
     method without the correction; each samples g_OO(r), g_OH(r) of the
     (centroid) positions, with the same Langevin centroid thermostat.

    python -m qtip4pf_water_amd.cmd build --out DIR [--frames N] ...
    python -m qtip4pf_water_amd.cmd validate --model FILE --out DIR ...
    python -m qtip4pf_water_amd.cmd steps --kind surrogate --steps 200
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import config as cfg

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
MODEL_FILE = os.path.join(DATA_DIR, "cmd_surrogate_216x32.npz")


def say(*a):
    # progress on stderr: stdout stays free for results (bench.py's JSON line)
    print("[cmd]", *a, file=sys.stderr, flush=True)


def box_for(n_mols, rho_gcc=0.997):
    return cfg.cubic_box_angs(n_mols, rho_gcc)


def start_state(n_mols, P, T_K, seed):
    L = box_for(n_mols)
    q0, box = cfg.water_box(n_mols, L, seed=seed)
    beta_n = 1.0 / (cfg.K2AU * T_K) / P
    q = cfg.free_ring_polymer_beads(q0, P, beta_n, seed=seed + 1) if P > 1 else q0
    p = cfg.sample_momenta(n_mols, P, beta_n, seed=seed + 2)
    return q, p, box


def centroid_dataset(n_frames, n_mols=216, P=32, T_K=298.0, dtfs=0.25, stride=80, n_eq=20,
                     n_acc=200, n_equil=4000, tau_fs=100.0, seed=11, log_every=50):
    """Centroid positions, centroid mean forces and classical forces of
    n_frames samples (arrays (n_frames, n_dof), Hartree/bohr)."""
    from .api import Water
    q, p, box = start_state(n_mols, P, T_K, seed)
    main = Water(n_mols, P, box, T_K=T_K, dtfs=dtfs)
    main.set_thermostat("pile_l", tau_fs)
    main.set_seed(seed)
    main.set_state(q, p)
    t0 = time.perf_counter()
    main.run(n_equil, True)
    main.sync()
    say(f"equilibrated {n_equil} steps in {time.perf_counter() - t0:.1f} s")
    cons = Water(n_mols, P, box, T_K=T_K, dtfs=dtfs)
    cons.set_pimd_mode("centroid_fixed")
    cons.set_seed(seed + 1)
    cls = Water(n_mols, 1, box, T_K=T_K, dtfs=dtfs)
    nd = 9 * n_mols
    pos = np.zeros((n_frames, nd))
    fmean = np.zeros((n_frames, nd))
    fcls = np.zeros((n_frames, nd))
    fhalf = np.zeros((n_frames, nd))   # mean over the first half of each run
    zero = np.zeros(nd)
    t0 = time.perf_counter()
    for f in range(n_frames):
        main.run(stride, True)
        qb, pb, _, _, _ = main.get_state()
        cons.set_state(qb, pb)
        cons.run(n_eq, True)
        cons.centroid_force(True)
        cons.run(n_acc // 2, True)
        mh, _, _ = cons.get_centroid_force()
        cons.run(n_acc - n_acc // 2, True)
        mf, qc, n = cons.get_centroid_force()
        cons.centroid_force(False)
        if n != n_acc:
            raise RuntimeError(f"mean force over {n} steps, expected {n_acc}")
        cls.set_state(qc, zero)
        _, _, Fc, _, _ = cls.get_state()
        pos[f], fmean[f], fcls[f], fhalf[f] = qc, mf, Fc, mh
        if (f + 1) % log_every == 0:
            el = time.perf_counter() - t0
            say(f"frame {f + 1}/{n_frames}  {el:.1f} s  "
                f"rms(F_mean - F_cl) = {np.sqrt(np.mean((fmean[:f+1] - fcls[:f+1]) ** 2)):.3e}")
    for w in (main, cons, cls):
        w.close()
    return dict(pos=pos, fmean=fmean, fcls=fcls, fhalf=fhalf, box=np.array(box), n_mols=n_mols, P=P,
                T_K=T_K, dtfs=dtfs, stride=stride, n_eq=n_eq, n_acc=n_acc, tau_fs=tau_fs)


def descriptor_stats(pos, n_mols, box, max_frames=16):
    """Per-element mean and standard deviation of the raw descriptors."""
    from .api import Surrogate
    from .api import NN_PARAMS_PER_ELEMENT as NPE
    with Surrogate(n_mols, box, max_frames=max_frames) as g:
        g.set_model(np.zeros(2 * NPE, np.float32), np.zeros(80), np.ones(80), 1.0)
        _, _, G = g.eval(pos, descriptors=True)
    el = np.tile(np.array([0, 1, 1]), n_mols)
    G = G.reshape(-1, n_mols * 3, 40)
    mean = np.stack([G[:, el == e].reshape(-1, 40).mean(0) for e in (0, 1)])
    std = np.stack([G[:, el == e].reshape(-1, 40).std(0) for e in (0, 1)])
    return mean, np.maximum(std, 1e-3 * np.max(std, axis=1, keepdims=True) + 1e-6)


def init_theta(seed=0, w3_scale=0.1):
    """Seeded initial parameters: W ~ N(0, 1/fan_in), zero biases, small w3."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(2):
        out += [rng.normal(0, 1 / math.sqrt(40), 128 * 40), np.zeros(128),
                rng.normal(0, 1 / math.sqrt(128), 128 * 128), np.zeros(128),
                rng.normal(0, w3_scale / math.sqrt(128), 128), np.zeros(1)]
    return np.concatenate(out).astype(np.float32)


def fit(data, n_train, epochs=200, batch=8, lr=2e-3, lr_end=1e-4, seed=3, log_every=20):
    """Force-match the correction on the first n_train frames; returns the
    model dictionary and the loss trace."""
    from .api import Surrogate
    n_mols, box = int(data["n_mols"]), data["box"]
    target = data["fmean"] - data["fcls"]
    mean, std = descriptor_stats(data["pos"][:n_train], n_mols, box)
    e_scale = float(np.sqrt(np.mean(target[:n_train] ** 2)))
    theta = init_theta(seed)
    trace = []
    with Surrogate(n_mols, box, max_frames=batch) as g:
        g.set_model(theta, mean, std, e_scale)
        g.set_data(data["pos"][:n_train], target[:n_train])
        steps_per_epoch = max(1, n_train // batch)
        t0 = time.perf_counter()
        for ep in range(epochs):
            # exponential decay of the learning rate from lr to lr_end
            r = lr * (lr_end / lr) ** (ep / max(1, epochs - 1))
            tr = g.train(steps_per_epoch, batch, r, seed=seed * 100003 + ep)
            trace.append(float(np.mean(tr)))
            if (ep + 1) % log_every == 0:
                say(f"epoch {ep + 1}/{epochs} lr {r:.2e} loss {trace[-1]:.4e} "
                    f"({time.perf_counter() - t0:.1f} s)")
        theta = g.get_model()
    model = dict(theta=theta, mean=mean.astype(np.float32), std=std.astype(np.float32),
                 e_scale=e_scale)
    return model, np.array(trace)


def force_errors(model, data, frames):
    """RMS errors on the given frames of the centroid force with and without
    the correction (Hartree/bohr), per element and overall."""
    from .api import Surrogate
    n_mols, box = int(data["n_mols"]), data["box"]
    with Surrogate(n_mols, box, max_frames=16) as g:
        g.set_model(model["theta"], model["mean"], model["std"], model["e_scale"])
        _, Fnn = g.eval(data["pos"][frames])
    tgt = data["fmean"][frames] - data["fcls"][frames]
    # sampling noise of each mean force: the two half-run means differ by
    # twice the full-run mean's error, rms(first half - full) estimates it
    noise = data["fhalf"][frames] - data["fmean"][frames]
    el = np.repeat(np.tile(np.array([0, 1, 1]), n_mols), 3)
    out = {}
    for name, sel in (("all", slice(None)), ("O", el == 0), ("H", el == 1)):
        out[name] = {"without": float(np.sqrt(np.mean(tgt[:, sel] ** 2))),
                     "with": float(np.sqrt(np.mean((tgt[:, sel] - Fnn[:, sel]) ** 2))),
                     "target_noise": float(np.sqrt(np.mean(noise[:, sel] ** 2)))}
    return out


def save_model(path, model, meta):
    np.savez(path, theta=model["theta"], mean=model["mean"], std=model["std"],
             e_scale=np.array(model["e_scale"]), meta=np.array(json.dumps(meta)))


def load_model(path=MODEL_FILE):
    with np.load(path, allow_pickle=False) as z:
        return dict(theta=z["theta"], mean=z["mean"], std=z["std"], e_scale=float(z["e_scale"]),
                    meta=json.loads(str(z["meta"])))


# ---------------------------------------------------------------------------
# CMD runs and radial distribution functions
# ---------------------------------------------------------------------------
def rdf_normalise(oo, oh, n_samples, n_mols, box_bohr, r_max_bohr):
    """g_OO and g_OH from the pair counts (O-O pairs i < j; O-H every pair)."""
    nb = len(oo)
    dr = r_max_bohr / nb
    r = (np.arange(nb) + 0.5) * dr
    shell = 4.0 / 3.0 * np.pi * ((np.arange(nb) + 1) ** 3 - np.arange(nb) ** 3) * dr ** 3
    V = float(np.prod(box_bohr))
    g_oo = 2.0 * oo / (n_samples * n_mols * (n_mols / V) * shell)
    g_oh = oh / (n_samples * n_mols * (2 * n_mols / V) * shell)
    return r * cfg.ANGS2BOHR ** -1, g_oo, g_oh


def rdf_error(g, g_ref, r_angs, r_max_angs=8.0):
    """Integrated relative error int |g - g_ref| dr / int g_ref dr over
    0 <= r <= r_max_angs."""
    sel = r_angs <= r_max_angs
    return float(np.sum(np.abs(g[sel] - g_ref[sel])) / np.sum(g_ref[sel]))


def cmd_run(kind, n_mols=512, T_K=298.0, model=None, ps=10.0, equil_ps=2.0, sample_fs=20.0,
            dt_fs=None, P=32, tau_fs=100.0, seed=21, r_max_angs=12.0, n_bins=600):
    """One CMD (or classical) run with g(r) sampling.  kind: "surrogate"
    (classical + learned correction, dt 0.5 fs), "classical" (dt 0.5 fs),
    "pacmd" (32 beads, dt 0.1 fs), "pimd" (32 beads, dt 0.25 fs)."""
    from .api import Surrogate, Water
    dflt = {"surrogate": 0.5, "classical": 0.5, "pacmd": 0.1, "pimd": 0.25}
    dt = dt_fs or dflt[kind]
    beads = 1 if kind in ("surrogate", "classical") else P
    q, p, box = start_state(n_mols, beads, T_K, seed)
    w = Water(n_mols, beads, box, T_K=T_K, dtfs=dt)
    g = None
    if kind == "surrogate":
        g = Surrogate(n_mols, box, max_frames=1)
        g.set_model(model["theta"], model["mean"], model["std"], model["e_scale"])
        w.set_nn(g)
    if kind == "pacmd":
        w.set_pimd_mode("pacmd")
    w.set_thermostat("pile_l", tau_fs)
    w.set_seed(seed)
    w.set_state(q, p)
    n_eq = int(round(equil_ps * 1000 / dt))
    every = max(1, int(round(sample_fs / dt)))
    n_samp = int(round(ps * 1000 / sample_fs))
    t0 = time.perf_counter()
    w.run(n_eq, True)
    w.sync()
    r_max = r_max_angs * cfg.ANGS2BOHR
    w.rdf_reset(r_max, n_bins)
    t1 = time.perf_counter()
    for s in range(n_samp):
        w.run(every, True)
        w.rdf_sample()
        if (s + 1) % max(1, n_samp // 5) == 0:
            w.sync()
            say(f"{kind}: {s + 1}/{n_samp} samples, {time.perf_counter() - t1:.1f} s")
    w.sync()
    el = time.perf_counter() - t1
    oo, oh, n = w.rdf_get()
    st = w.scalars()
    w.close()
    if g is not None:
        g.close()
    sps = n_samp * every / el
    r, g_oo, g_oh = rdf_normalise(oo, oh, n, n_mols, np.array(box), r_max)
    return dict(kind=kind, n_mols=n_mols, beads=beads, dt_fs=dt, steps=n_samp * every,
                seconds=el, steps_per_s=sps, ns_per_day=sps * dt * 0.0864,
                equil_seconds=t1 - t0, r=r, g_oo=g_oo, g_oh=g_oh, oo=oo, oh=oh, n_samples=n,
                status=st["status"])


def steps(kind, n_steps=200, n_mols=512, T_K=298.0, model=None, P=32, tau_fs=100.0, seed=21):
    """n_steps resident NVT steps of one CMD kind (no g(r) sampling), for
    kernel traces: steps/s."""
    from .api import Surrogate, Water
    dflt = {"surrogate": 0.5, "classical": 0.5, "pacmd": 0.1, "pimd": 0.25}
    beads = 1 if kind in ("surrogate", "classical") else P
    q, p, box = start_state(n_mols, beads, T_K, seed)
    w = Water(n_mols, beads, box, T_K=T_K, dtfs=dflt[kind])
    g = None
    if kind == "surrogate":
        g = Surrogate(n_mols, box, max_frames=1)
        g.set_model(model["theta"], model["mean"], model["std"], model["e_scale"])
        w.set_nn(g)
    if kind == "pacmd":
        w.set_pimd_mode("pacmd")
    w.set_thermostat("pile_l", tau_fs)
    w.set_seed(seed)
    w.set_state(q, p)
    w.run(20, True)
    w.sync()
    t0 = time.perf_counter()
    w.run(n_steps, True)
    w.sync()
    el = time.perf_counter() - t0
    st = w.scalars()["status"]
    w.close()
    if g is not None:
        g.close()
    return n_steps / el, st


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("steps", help="time resident steps of one kind (for kernel traces)")
    r.add_argument("--kind", default="surrogate",
                   choices=["surrogate", "classical", "pacmd", "pimd"])
    r.add_argument("--steps", type=int, default=200)
    r.add_argument("--n-mols", type=int, default=512)
    r.add_argument("--model", default=MODEL_FILE)
    b = sub.add_parser("build", help="generate mean forces, train, save the model")
    b.add_argument("--out", required=True)
    b.add_argument("--frames", type=int, default=1500)
    b.add_argument("--val-frames", type=int, default=100)
    b.add_argument("--stride", type=int, default=80)
    b.add_argument("--n-acc", type=int, default=200)
    b.add_argument("--n-eq", type=int, default=20)
    b.add_argument("--n-acc-val", type=int, default=1000)
    b.add_argument("--epochs", type=int, default=200)
    b.add_argument("--batch", type=int, default=8)
    b.add_argument("--lr", type=float, default=2e-3)
    v = sub.add_parser("validate", help="CMD runs at 512 H2O and g(r) errors")
    v.add_argument("--model", default=MODEL_FILE)
    v.add_argument("--out", required=True)
    v.add_argument("--ps", type=float, default=10.0)
    v.add_argument("--equil-ps", type=float, default=2.0)
    v.add_argument("--kinds", default="pacmd,surrogate,classical")
    v.add_argument("--n-mols", type=int, default=512)
    v.add_argument("--ref", default=None,
                   help="g(r) file of an earlier plain-CMD run (rdf_pacmd.npz) instead of a new one")
    a = ap.parse_args(argv)
    if a.cmd == "steps":
        model = load_model(a.model) if a.kind == "surrogate" else None
        sps, st = steps(a.kind, a.steps, a.n_mols, model=model)
        say(f"{a.kind}: {sps:.1f} steps/s, status {st}")
        return 0 if st == 0 else 1
    os.makedirs(a.out, exist_ok=True)
    if a.cmd == "build":
        t0 = time.perf_counter()
        data = centroid_dataset(a.frames, stride=a.stride, n_eq=a.n_eq, n_acc=a.n_acc, seed=11)
        tgen = time.perf_counter() - t0
        say(f"training data: {a.frames} frames in {tgen:.1f} s")
        # validation frames from an independent trajectory, longer averages
        t0 = time.perf_counter()
        val = centroid_dataset(a.val_frames, stride=a.stride, n_eq=a.n_eq, n_acc=a.n_acc_val,
                               seed=97)
        tval = time.perf_counter() - t0
        say(f"validation data: {a.val_frames} frames in {tval:.1f} s")
        # the validation frames only (the training set is ~90 MB)
        np.savez_compressed(os.path.join(a.out, "centroid_val.npz"),
                            **{k: (v.astype(np.float32) if isinstance(v, np.ndarray) and v.ndim == 2
                                   else v) for k, v in val.items()})
        t0 = time.perf_counter()
        model, trace = fit(data, a.frames, epochs=a.epochs, batch=a.batch, lr=a.lr)
        ttr = time.perf_counter() - t0
        errs = force_errors(model, val, np.arange(a.val_frames))
        errs_train = force_errors(model, data, np.arange(0, min(a.frames, 200)))
        meta = dict(frames=a.frames, val_frames=a.val_frames, stride=a.stride, n_acc=a.n_acc,
                    n_acc_val=a.n_acc_val, n_eq=a.n_eq, epochs=a.epochs, batch=a.batch, lr=a.lr,
                    gen_seconds=tgen, val_gen_seconds=tval, train_seconds=ttr,
                    val_force_rms=errs, train_force_rms=errs_train,
                    final_loss=float(trace[-1]),
                    system="216 H2O x 32 beads, 298 K, PILE-L tau 100 fs, dt 0.25 fs")
        save_model(os.path.join(a.out, "cmd_surrogate_216x32.npz"), model, meta)
        np.save(os.path.join(a.out, "loss_trace.npy"), trace)
        with open(os.path.join(a.out, "build.json"), "w") as f:
            json.dump(meta, f, indent=1)
        say(json.dumps(meta))
    else:
        model = load_model(a.model)
        res = {}
        if a.ref:
            with np.load(a.ref) as z:
                res["pacmd"] = {k: (z[k] if z[k].ndim else z[k].item()) for k in z.files}
        for kind in a.kinds.split(","):
            if kind in res:
                continue
            r = cmd_run(kind, n_mols=a.n_mols, model=model, ps=a.ps, equil_ps=a.equil_ps)
            res[kind] = r
            np.savez(os.path.join(a.out, f"rdf_{kind}.npz"),
                     **{k: np.asarray(v) for k, v in r.items() if k != "kind"})
            say(f"{kind}: {r['steps_per_s']:.1f} steps/s, {r['ns_per_day']:.3f} ns/day")
        ref = "pacmd" if "pacmd" in res else None
        summary = {}
        for kind, r in res.items():
            s = {k: r[k] for k in ("n_mols", "beads", "dt_fs", "steps", "seconds", "steps_per_s",
                                   "ns_per_day", "n_samples", "status")}
            if ref and kind != ref:
                s["g_oo_error_vs_" + ref] = rdf_error(r["g_oo"], res[ref]["g_oo"], r["r"])
                s["g_oh_error_vs_" + ref] = rdf_error(r["g_oh"], res[ref]["g_oh"], r["r"])
                s["speedup_ns_per_day_vs_" + ref] = r["ns_per_day"] / res[ref]["ns_per_day"]
            summary[kind] = s
        with open(os.path.join(a.out, "validate.json"), "w") as f:
            json.dump(summary, f, indent=1)
        say(json.dumps(summary))


if __name__ == "__main__":
    sys.exit(main())
