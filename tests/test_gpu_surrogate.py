"""The learned centroid-force correction on the GPU (BF16 matrix cores,
FP32 accumulation) against the FP64 oracle (oracle/surrogate.py, itself
pinned by finite differences in test_surrogate_oracle.py): descriptors,
energies and forces, the force-matching gradient, training, and the
coupling into the classical resident step.  No counterpart in the
reference.

Tolerances (stated with each check) follow from the arithmetic: the
descriptors are FP32 sums (1e-5 relative); the network takes BF16 operands
(8-bit significands, 2^-9 relative rounding), so against an oracle that
rounds the same operands the remaining differences are FP32 accumulation
and the occasional activation that rounds the other way (1e-2 relative);
against the exact FP64 gradient the BF16 gradient agrees to a few percent
in norm and better than 0.995 in direction."""
import sys

import numpy as np
import pytest

from qtip4pf_water_amd import config as cfg

from oracle import surrogate as S

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _frames(n_side, L, n, seed, jitter=0.12):
    out = []
    box = None
    for k in range(n):
        q0, box = cfg.lattice_water_box(n_side, L, seed=seed + k, jitter_bohr=jitter)
        out.append(q0)
    return np.array(out), np.array(box, float), n_side ** 3


def _model_for(frames, box, n_mols, seed, e_scale=2e-3):
    """Random parameters; normalisation from the oracle's descriptors."""
    elem = S.elements(n_mols)
    Gs = [S.descriptors(f.reshape(-1, 3) * S.ANG, elem, box * S.ANG, S.default_desc())[0]
          for f in frames]
    G = np.concatenate(Gs)
    el = np.tile(elem, len(frames))
    mean = np.stack([G[el == e].mean(0) for e in (0, 1)])
    std = np.stack([G[el == e].std(0) + 1e-3 for e in (0, 1)])
    theta = S.init_theta(seed, w3_scale=1.0)
    return S.Model(theta, mean, std, e_scale)


def _gpu_model(m, n_mols, box, max_frames=4):
    from qtip4pf_water_amd import Surrogate
    g = Surrogate(n_mols, box, max_frames=max_frames)
    g.set_model(m.theta.astype(np.float32), m.mean, m.std, m.e_scale)
    return g


def test_descriptors_energies_forces_match_oracle():
    say("surrogate eval 216 x 3 frames")
    X, box, n = _frames(6, 18.64, 3, seed=31)
    m = _model_for(X, box, n, seed=1)
    # the oracle takes the parameters as the GPU holds them (FP32)
    m.theta = m.theta.astype(np.float32).astype(np.float64)
    m.mean = m.mean.astype(np.float32).astype(np.float64)
    m.std = m.std.astype(np.float32).astype(np.float64)
    with _gpu_model(m, n, box, max_frames=2) as g:     # 3 frames: two batches
        E, F, G = g.eval(X, descriptors=True)
    elem = S.elements(n)
    for k in range(3):
        Gr, _ = S.descriptors(X[k].reshape(-1, 3) * S.ANG, elem, box * S.ANG, S.default_desc())
        # FP32 descriptor sums: 1e-5 of each column's scale
        scale = np.max(np.abs(Gr), axis=0) + 1e-6
        assert np.max(np.abs(G[k] - Gr) / scale) < 1e-5, k
        Er, Fr = S.energy_forces(m, X[k], box, n, emulate_bf16=True)
        Fr = Fr.ravel()
        # against BF16-operand emulation: 1e-2 of the force scale, energy 1e-2 relative
        rms = np.sqrt(np.mean(Fr ** 2))
        assert np.sqrt(np.mean((F[k] - Fr) ** 2)) < 1e-2 * rms, k
        assert abs(E[k] - Er) < 1e-2 * abs(Er), (E[k], Er)
        # and against exact FP64 arithmetic: 5e-2 of the force scale
        E64, F64 = S.energy_forces(m, X[k], box, n)
        assert np.sqrt(np.mean((F[k] - F64.ravel()) ** 2)) < 5e-2 * rms, k


def test_loss_gradient_matches_oracle():
    say("surrogate loss gradient 64 x 2 frames")
    X, box, n = _frames(4, 12.43, 2, seed=41)
    m = _model_for(X, box, n, seed=2)
    m.theta = m.theta.astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(3)
    T = rng.normal(0, 2e-3, X.shape)
    L_ref, g_ref = S.loss_and_grad(m, [x.reshape(-1, 3) for x in X],
                                   [t.reshape(-1, 3) for t in T], box, n)
    with _gpu_model(m, n, box, max_frames=2) as g:
        g.set_data(X, T)
        L, grad = g.loss_grad(0, 2)
    assert abs(L - L_ref) < 2e-2 * L_ref, (L, L_ref)
    cos = np.dot(grad, g_ref) / (np.linalg.norm(grad) * np.linalg.norm(g_ref))
    assert cos > 0.995, cos
    assert np.linalg.norm(grad - g_ref) < 0.08 * np.linalg.norm(g_ref)
    # every parameter block carries its share (W1, b1, W2, b2, w3 of O and H)
    NH, NI = S.NH, S.NIN
    bounds = [0, NH * NI, NH * NI + NH, NH * NI + NH + NH * NH, NH * NI + 2 * NH + NH * NH,
              S.NP_ELEM - 1]
    for e in (0, 1):
        for a, b in zip(bounds[:-1], bounds[1:]):
            sl = slice(e * S.NP_ELEM + a, e * S.NP_ELEM + b)
            c = np.dot(grad[sl], g_ref[sl]) / (np.linalg.norm(grad[sl]) * np.linalg.norm(g_ref[sl]))
            assert c > 0.98, (e, a, c)


def test_training_fits_a_teacher():
    """Force matching on the GPU recovers a teacher network's forces: targets
    are the FP64 oracle forces of another parameter set."""
    say("surrogate training 64 x 24 frames")
    X, box, n = _frames(4, 12.43, 24, seed=51, jitter=0.2)
    teacher = _model_for(X, box, n, seed=7, e_scale=2e-3)
    T = np.array([S.energy_forces(teacher, x, box, n)[1].ravel() for x in X])
    student = _model_for(X, box, n, seed=8, e_scale=2e-3)
    with _gpu_model(student, n, box, max_frames=4) as g:
        g.set_data(X[:20], T[:20])
        tr = g.train(300, 4, 2e-3, seed=1)
        _, Fh = g.eval(X[20:])
    assert np.all(np.isfinite(tr))
    assert tr[-20:].mean() < 0.1 * tr[:5].mean(), (tr[:5].mean(), tr[-20:].mean())
    # held-out frames: the fit generalises
    err0 = np.mean(T[20:] ** 2)
    assert np.mean((Fh - T[20:]) ** 2) < 0.2 * err0


def test_correction_enters_the_classical_step(oracle_mod):
    """Attached to a classical context, the correction's forces and energy add
    to qTIP4P/F's (same frame evaluated alone), and an NVE run with it is
    stable."""
    from qtip4pf_water_amd import Water
    say("surrogate + classical step 216")
    X, box, n = _frames(6, 18.64, 1, seed=61, jitter=0.05)
    m = _model_for(X, box, n, seed=3, e_scale=1e-4)
    q = X[0].copy()
    p = cfg.sample_momenta(n, 1, 1.0 / (cfg.K2AU * 298.0), seed=5)
    with _gpu_model(m, n, box, max_frames=1) as g, Water(n, 1, box) as w:
        En, Fn = g.eval(q)
        w.set_state(q, p)
        e0 = w.scalars()["energy"]
        _, _, F0, _, _ = w.get_state()
        w.set_nn(g)
        w.set_state(q, p)
        e1 = w.scalars()["energy"]
        _, _, F1, _, _ = w.get_state()
        assert abs((e1 - e0) - En[0]) < 1e-9 * abs(e0)
        assert np.max(np.abs(F1 - F0 - Fn[0])) < 1e-12 * np.max(np.abs(F0))
        obs = w.run_observed(200, every=20)
        assert np.all(np.isfinite(obs))
        w.set_nn(None)
