"""Pins of the FP64 oracle of the learned centroid-force correction
(oracle/surrogate.py; no counterpart in the reference): finite differences
of the descriptor, of the energy (F = -grad E) and of the loss in the
parameters; invariances; single terms evaluated by hand."""
import math

import numpy as np
import pytest

from qtip4pf_water_amd import config as cfg

from oracle import surrogate as S


def _frame(n_side=2, L=9.0, seed=0, jitter=0.15):
    q0, box = cfg.lattice_water_box(n_side, L, seed=seed, jitter_bohr=jitter)
    return q0.reshape(-1, 3), np.array(box, float), n_side ** 3


def _model(seed=1):
    rng = np.random.default_rng(seed)
    theta = S.init_theta(seed, w3_scale=1.0)
    mean = rng.normal(0, 0.3, (2, S.NIN))
    std = rng.uniform(0.5, 2.0, (2, S.NIN))
    return S.Model(theta, mean, std, e_scale=1e-3)


def test_descriptor_gradient_by_finite_differences():
    x, box, n = _frame()
    xa, boxa = x * S.ANG, box * S.ANG
    elem = S.elements(n)
    G, parts = S.descriptors(xa, elem, boxa, S.default_desc(), with_grad=True)
    rng = np.random.default_rng(3)
    V = rng.normal(size=xa.shape)
    u = S.tangent(parts, V, xa.shape[0])
    h = 1e-5
    Gp, _ = S.descriptors(xa + h * V, elem, boxa, S.default_desc())
    Gm, _ = S.descriptors(xa - h * V, elem, boxa, S.default_desc())
    fd = (Gp - Gm) / (2 * h)
    assert np.max(np.abs(u - fd)) <= 1e-6 * max(1.0, np.max(np.abs(fd)))
    assert np.max(np.abs(u[:, :S.NR])) > 0 and np.max(np.abs(u[:, S.NR:])) > 0


def test_forces_are_minus_energy_gradient():
    x, box, n = _frame(seed=4)
    m = _model()
    E, F = S.energy_forces(m, x, box, n)
    rng = np.random.default_rng(5)
    for _ in range(3):
        V = rng.normal(size=x.shape)
        h = 1e-5
        Ep, _ = S.energy_forces(m, x + h * V, box, n)
        Em, _ = S.energy_forces(m, x - h * V, box, n)
        assert (Ep - Em) / (2 * h) == pytest.approx(-np.sum(F * V), rel=1e-6)
    # no net force (translation invariance of a sum of pair/triplet terms)
    assert np.max(np.abs(F.sum(0))) < 1e-12 * np.max(np.abs(F)) * F.shape[0]


def test_loss_gradient_by_finite_differences():
    x, box, n = _frame(seed=6)
    m = _model(2)
    rng = np.random.default_rng(7)
    frames = [x, x + rng.normal(0, 0.05, x.shape)]
    targets = [rng.normal(0, 1e-3, x.shape) for _ in frames]
    L, g = S.loss_and_grad(m, frames, targets, box, n)
    # directional derivatives along random parameter directions and along the
    # gradient itself, plus single coordinates of every parameter block
    for k in range(3):
        d = rng.normal(size=m.theta.size)
        h = 1e-4 / np.linalg.norm(d) * math.sqrt(d.size)
        Lp, _ = S.loss_and_grad(m, frames, targets, box, n, m.theta + h * d)
        Lm, _ = S.loss_and_grad(m, frames, targets, box, n, m.theta - h * d)
        assert (Lp - Lm) / (2 * h) == pytest.approx(np.dot(g, d), rel=1e-5)
    offs = [0, S.NH * S.NIN, S.NH * S.NIN + S.NH, S.NH * S.NIN + 2 * S.NH + S.NH * S.NH - 5,
            S.NP_ELEM - 2, S.NP_ELEM + 17]
    for o in offs:
        th = m.theta.copy()
        h = 1e-5
        th[o] += h
        Lp, _ = S.loss_and_grad(m, frames, targets, box, n, th)
        th[o] -= 2 * h
        Lm, _ = S.loss_and_grad(m, frames, targets, box, n, th)
        assert (Lp - Lm) / (2 * h) == pytest.approx(g[o], rel=1e-4, abs=1e-12)
    assert g[S.NP_ELEM - 1] == 0.0        # b3 does not move forces


def test_descriptor_invariances():
    x, box, n = _frame(seed=8)
    xa, boxa = x * S.ANG, box * S.ANG
    elem = S.elements(n)
    G, _ = S.descriptors(xa, elem, boxa, S.default_desc())
    # translation by an arbitrary vector and by whole box vectors
    G2, _ = S.descriptors(xa + np.array([0.37, -1.2, 2.9]), elem, boxa, S.default_desc())
    assert np.allclose(G, G2, atol=1e-12)
    xs = xa.copy()
    xs[5] += boxa * np.array([1, -2, 0])
    G3, _ = S.descriptors(xs, elem, boxa, S.default_desc())
    assert np.allclose(G, G3, atol=1e-12)
    # swapping the two hydrogens of a molecule permutes their rows only
    xp = xa.copy()
    xp[[1, 2]] = xp[[2, 1]]
    G4, _ = S.descriptors(xp, elem, boxa, S.default_desc())
    assert np.allclose(G4[[1, 2]], G[[2, 1]], atol=1e-12)
    assert np.allclose(G4[3:], G[3:], atol=1e-12)


def test_rotation_invariance_in_an_open_cluster():
    # a cluster far smaller than the box: no image within the cutoffs, so a
    # rotation about its centre leaves every descriptor unchanged
    x, _, n = _frame(n_side=2, L=6.0, seed=9)
    box = np.array([200.0, 200.0, 200.0])
    xa, boxa = x * S.ANG + 30.0, box * S.ANG
    elem = S.elements(n)
    G, _ = S.descriptors(xa, elem, boxa, S.default_desc())
    th = 0.7
    R = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    c = xa.mean(0)
    Gr, _ = S.descriptors((xa - c) @ R.T + c, elem, boxa, S.default_desc())
    assert np.allclose(G, Gr, atol=1e-11)


def test_single_terms_by_hand():
    # three atoms: O at the origin, H at (1, 0, 0) A, O at (0, 2, 0) A, in a
    # huge box: each descriptor element written out directly
    d = S.default_desc()
    box = np.array([100.0, 100.0, 100.0])
    x = np.array([[0, 0, 0], [1.0, 0, 0], [0, 2.0, 0]]) + 10.0
    elem = np.array([0, 1, 0])
    G, _ = S.descriptors(x, elem, box, d)
    fc = lambda r, R: 0.5 * (math.cos(math.pi * r / R) + 1) if r < R else 0.0
    m = 5
    mu = d["mu0"] + m * d["dmu"]
    want = (d["w"][0][1] * math.exp(-d["eta_r"] * (1 - mu) ** 2) * fc(1, 6)
            + d["w"][0][0] * math.exp(-d["eta_r"] * (2 - mu) ** 2) * fc(2, 6))
    assert G[0, m] == pytest.approx(want, rel=1e-14)
    # angle at atom 0 between H (1,0,0) and O (0,2,0): cos = 0
    for n_ in range(S.NA):
        z, lam, eta = d["zeta"][n_], d["lam"][n_], d["eta_a"][n_]
        t = (2 ** (1 - z) * d["w"][0][1] * d["w"][0][0] * (1 + lam * 0.0) ** z
             * math.exp(-eta * (1 + 4)) * fc(1, 4) * fc(2, 4))
        assert G[0, S.NR + n_] == pytest.approx(t, rel=1e-13)
    # at the H atom: neighbours O (dist 1) and O (dist sqrt 5), cos = 1/sqrt5
    r2 = math.sqrt(5.0)
    c = (-1 * -1 + 0) / (1 * r2)
    z, lam, eta = d["zeta"][3], d["lam"][3], d["eta_a"][3]
    t = 2 ** (1 - z) * d["w"][1][0] ** 2 * (1 + lam * c) ** z * math.exp(-eta * (1 + 5)) * fc(1, 4) * fc(r2, 4)
    assert G[1, S.NR + 3] == pytest.approx(t, rel=1e-13)


def test_rdf_counts_pairs_once():
    x, box, n = _frame(n_side=3, L=9.4, seed=10)
    oo, oh = S.rdf(x, box, n, 4.6, 46)
    assert oo.sum() <= n * (n - 1) // 2 and oh.sum() <= n * 2 * n
    # every intramolecular O-H distance (0.94 A, jittered) falls in 0.6-1.3 A
    assert oh[6:13].sum() >= 2 * n
