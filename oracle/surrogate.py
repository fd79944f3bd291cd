"""FP64 numpy restatement of the learned centroid-force correction.

TEST INFRASTRUCTURE ONLY: the checker the GPU surrogate kernels are compared
against; imported by tests/ only, never by the product package.

There is no counterpart in the reference.  Its centroid-potential hook is an
additive pair correction (the spline path, forces.cpp:452-466, 634-691):
V_total = V_qTIP4P/F + V_corr.  Here V_corr is a sum of atomic energies of an
atom-centred network:

  descriptor of atom i (element e_i, 0 = O, 1 = H), 40 numbers, positions in
  Angstrom, minimum image in the orthorhombic box:
    radial  G_m = sum_j w[e_i][e_j] exp(-eta (r_ij - mu_m)^2) fc(r_ij; Rr),
            mu_m = mu0 + m dmu, m = 0..23
    angular G_24+n = 2^(1-zeta_n) sum_{j<k} w[e_i][e_j] w[e_i][e_k]
                     (1 + lambda_n cos t_jik)^zeta_n exp(-eta_n (r_ij^2 + r_ik^2))
                     fc(r_ij; Ra) fc(r_ik; Ra),   n = 0..15
    fc(r; R) = (cos(pi r / R) + 1) / 2 for r < R, else 0
  normalised g = (G - mean[e]) / std[e]
  E_i = e_scale (w3 . tanh(W2 tanh(W1 g + b1) + b2) + b3)     (per element)

(Behler-Parrinello symmetry functions, J. Chem. Phys. 134, 074106 (2011),
with the element-pair weights of weighted ACSFs, Gastegger et al., J. Chem.
Phys. 148, 241709 (2018).)

Forces F = -dE/dr (Hartree/bohr); the force-matching loss over frames
L = mean (F + F_model - F_target)^2 and its gradient in the parameters by
the tangent/adjoint construction the GPU kernels use, written out here in
plain numpy.  The parameter vector is packed per element as
  W1[128][40], b1[128], W2[128][128], b2[128], w3[128], b3
(21,889 numbers per element, O first).

What pins it (tests/test_surrogate_oracle.py): central finite differences of
G in the positions, of E in the positions (F = -grad E), of the loss in the
parameters, and of J in the tangent; rotation, translation and same-element
permutation invariance; brute-force evaluation of single terms.
"""
from __future__ import annotations

import math

import numpy as np

ANG = 0.529177210903          # Angstrom per bohr
NR, NA, NIN, NH = 24, 16, 40, 128
NP_ELEM = NH * NIN + NH + NH * NH + NH + NH + 1


def default_desc():
    """Descriptor constants (Angstrom): the same numbers as the library's
    qtp_nn_descriptor_params."""
    zeta = [1, 1, 2, 2, 4, 4, 8, 8] * 2
    lam = [1, -1] * 8
    eta_a = [0.02] * 8 + [0.2] * 8
    return dict(r_rad=6.0, r_ang=4.0, mu0=0.8, dmu=(6.0 - 0.8) / 23.0,
                eta_r=0.5 / ((6.0 - 0.8) / 23.0) ** 2,
                zeta=np.array(zeta, float), lam=np.array(lam, float),
                eta_a=np.array(eta_a, float),
                w=np.array([[1.0, 0.5], [1.0, 0.5]]))


def elements(n_mols):
    return np.tile(np.array([0, 1, 1]), n_mols)


def _fc(r, R):
    return np.where(r < R, 0.5 * (np.cos(np.pi * r / R) + 1.0), 0.0)


def _dfc(r, R):
    return np.where(r < R, -0.5 * np.pi / R * np.sin(np.pi * r / R), 0.0)


def _pairs(x, box, rc):
    """All ordered pairs (i, j), i != j, with minimum-image distance < rc.
    x: (N, 3) Angstrom; returns i, j, d = x_j - x_i (min image), r."""
    N = x.shape[0]
    d = x[None, :, :] - x[:, None, :]
    d -= box * np.round(d / box)
    r2 = np.einsum("ijk,ijk->ij", d, d)
    np.fill_diagonal(r2, np.inf)
    i, j = np.nonzero(r2 < rc * rc)
    return i, j, d[i, j], np.sqrt(r2[i, j])


class Model:
    """Parameters theta (packed), normalisation mean/std (2, 40), e_scale."""

    def __init__(self, theta, mean, std, e_scale=1.0, desc=None):
        self.theta = np.asarray(theta, float)
        self.mean = np.asarray(mean, float).reshape(2, NIN)
        self.std = np.asarray(std, float).reshape(2, NIN)
        self.e_scale = float(e_scale)
        self.desc = desc or default_desc()

    def layers(self, e, theta=None):
        t = (self.theta if theta is None else theta)[e * NP_ELEM:(e + 1) * NP_ELEM]
        o = 0
        W1 = t[o:o + NH * NIN].reshape(NH, NIN); o += NH * NIN
        b1 = t[o:o + NH]; o += NH
        W2 = t[o:o + NH * NH].reshape(NH, NH); o += NH * NH
        b2 = t[o:o + NH]; o += NH
        w3 = t[o:o + NH]; o += NH
        b3 = t[o]
        return W1, b1, W2, b2, w3, b3


def init_theta(seed=0, w3_scale=0.1):
    """Seeded initial parameters: W ~ N(0, 1/fan_in), biases 0, small w3."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(2):
        out += [rng.normal(0, 1 / math.sqrt(NIN), NH * NIN), np.zeros(NH),
                rng.normal(0, 1 / math.sqrt(NH), NH * NH), np.zeros(NH),
                rng.normal(0, w3_scale / math.sqrt(NH), NH), np.zeros(1)]
    return np.concatenate(out)


def descriptors(x, elem, box, desc, with_grad=False):
    """G (N, 40) and, with_grad, the derivative pieces for forces/tangents:
    a list of (atom index array, derivative (n, 40, 3)) contributions
    dG[i]/dr[atom] over radial pairs and angular triplets."""
    N = x.shape[0]
    G = np.zeros((N, NIN))
    w = desc["w"]
    # radial
    i, j, d, r = _pairs(x, box, desc["r_rad"])
    mu = desc["mu0"] + desc["dmu"] * np.arange(NR)
    wij = w[elem[i], elem[j]]
    ex = np.exp(-desc["eta_r"] * (r[:, None] - mu[None, :]) ** 2)
    fc = _fc(r, desc["r_rad"])
    phi = wij[:, None] * ex * fc[:, None]
    np.add.at(G, (i, slice(0, NR)), phi)
    parts = []
    if with_grad:
        dphi = wij[:, None] * (ex * (-2 * desc["eta_r"] * (r[:, None] - mu[None, :])) * fc[:, None]
                               + ex * _dfc(r, desc["r_rad"])[:, None])
        u = d / r[:, None]            # unit vector i -> j
        # dr_ij/dr_i = -u, dr_ij/dr_j = +u
        parts.append(("rad", i, j, dphi, u))
    # angular
    ia, ja, da, ra = _pairs(x, box, desc["r_ang"])
    order = np.lexsort((ja, ia))
    ia, ja, da, ra = ia[order], ja[order], da[order], ra[order]
    starts = np.searchsorted(ia, np.arange(N + 1))
    trip_i, trip_j, trip_k, A, B = [], [], [], [], []
    for c in range(N):
        s0, s1 = starts[c], starts[c + 1]
        n = s1 - s0
        if n < 2:
            continue
        jj, kk = np.triu_indices(n, 1)
        trip_i.append(np.full(jj.size, c))
        trip_j.append(ja[s0 + jj])
        trip_k.append(ja[s0 + kk])
        A.append(da[s0 + jj])
        B.append(da[s0 + kk])
    if trip_i:
        ti = np.concatenate(trip_i)
        tj, tk = np.concatenate(trip_j), np.concatenate(trip_k)
        A, B = np.concatenate(A), np.concatenate(B)
        T, dTa, dTb = _angular_terms(A, B, elem[ti], elem[tj], elem[tk], desc, with_grad)
        np.add.at(G, (ti, slice(NR, NIN)), T)
        if with_grad:
            parts.append(("ang", ti, tj, tk, dTa, dTb))
    return G, parts


def _angular_terms(A, B, ei, ej, ek, desc, with_grad):
    w = desc["w"]
    ra = np.linalg.norm(A, axis=1)
    rb = np.linalg.norm(B, axis=1)
    c = np.einsum("ij,ij->i", A, B) / (ra * rb)
    fa, fb = _fc(ra, desc["r_ang"]), _fc(rb, desc["r_ang"])
    wt = w[ei, ej] * w[ei, ek]
    zeta, lam, eta = desc["zeta"], desc["lam"], desc["eta_a"]
    base = 1.0 + lam[None, :] * c[:, None]
    pw = base ** zeta[None, :]
    ex = np.exp(-eta[None, :] * (ra ** 2 + rb ** 2)[:, None])
    pref = 2.0 ** (1.0 - zeta)[None, :] * wt[:, None]
    T = pref * pw * ex * (fa * fb)[:, None]
    if not with_grad:
        return T, None, None
    dpw_dc = zeta[None, :] * base ** (zeta[None, :] - 1) * lam[None, :]
    # dc/dA = B/(ra rb) - c A/ra^2
    dcA = B / (ra * rb)[:, None] - (c / ra ** 2)[:, None] * A
    dcB = A / (ra * rb)[:, None] - (c / rb ** 2)[:, None] * B
    dfa, dfb = _dfc(ra, desc["r_ang"]), _dfc(rb, desc["r_ang"])
    common = pref * ex                                  # (n, 16)
    # dT/dA = common [dpw/dc dc/dA fa fb + pw (-2 eta A) fa fb + pw fa' A/ra fb]
    t1 = (common * dpw_dc * (fa * fb)[:, None])[:, :, None] * dcA[:, None, :]
    t2 = (common * pw * (fa * fb)[:, None] * (-2 * eta[None, :]))[:, :, None] * A[:, None, :]
    t3 = (common * pw * (dfa / ra * fb)[:, None])[:, :, None] * A[:, None, :]
    dTa = t1 + t2 + t3
    u1 = (common * dpw_dc * (fa * fb)[:, None])[:, :, None] * dcB[:, None, :]
    u2 = (common * pw * (fa * fb)[:, None] * (-2 * eta[None, :]))[:, :, None] * B[:, None, :]
    u3 = (common * pw * (fa * dfb / rb)[:, None])[:, :, None] * B[:, None, :]
    dTb = u1 + u2 + u3
    return T, dTa, dTb


def tangent(parts, V, N):
    """u (N, 40): the derivative of G along the displacement field V (N, 3)
    (Angstrom derivative)."""
    u = np.zeros((N, NIN))
    for p in parts:
        if p[0] == "rad":
            _, i, j, dphi, uv = p
            # r_ij = |x_j - x_i|: d r_ij = uv . (V_j - V_i)
            dr = np.einsum("nk,nk->n", uv, V[j] - V[i])
            np.add.at(u, (i, slice(0, NR)), dphi * dr[:, None])
        else:
            _, ti, tj, tk, dTa, dTb = p
            # A = x_j - x_i, B = x_k - x_i
            dA, dB = V[tj] - V[ti], V[tk] - V[ti]
            val = np.einsum("nmk,nk->nm", dTa, dA) + np.einsum("nmk,nk->nm", dTb, dB)
            np.add.at(u, (ti, slice(NR, NIN)), val)
    return u


def grad_positions(parts, dEdG, N):
    """dE/dx (N, 3) (Angstrom) from dE/dG (N, 40) through the descriptor."""
    g = np.zeros((N, 3))
    for p in parts:
        if p[0] == "rad":
            _, i, j, dphi, uv = p
            c = np.einsum("nm,nm->n", dEdG[i, :NR], dphi)
            np.add.at(g, j, c[:, None] * uv)
            np.add.at(g, i, -c[:, None] * uv)
        else:
            _, ti, tj, tk, dTa, dTb = p
            e = dEdG[ti, NR:]
            ga = np.einsum("nm,nmk->nk", e, dTa)
            gb = np.einsum("nm,nmk->nk", e, dTb)
            np.add.at(g, tj, ga)
            np.add.at(g, tk, gb)
            np.add.at(g, ti, -(ga + gb))
    return g


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64).reshape(np.shape(x))


def mlp(model, e, g, theta=None, emulate_bf16=False):
    """Per-atom energies (in e_scale units) and dE/dg for element e.
    emulate_bf16: round every matrix-product operand to BF16 as the GPU's
    matrix cores take them (weights, inputs, activations, adjoints), with
    exact accumulation."""
    W1, b1, W2, b2, w3, b3 = model.layers(e, theta)
    r = bf16 if emulate_bf16 else (lambda x: x)
    W1r, W2r = r(W1), r(W2)
    z1 = r(g) @ W1r.T + b1
    a1 = np.tanh(z1)
    z2 = r(a1) @ W2r.T + b2
    a2 = np.tanh(z2)
    E = a2 @ w3 + b3
    dz2 = w3[None, :] * (1 - a2 * a2)
    da1 = r(dz2) @ W2r
    dz1 = da1 * (1 - a1 * a1)
    dg = r(dz1) @ W1r
    return E, dg


def energy_forces(model, x_bohr, box_bohr, n_mols, theta=None, return_parts=False,
                  emulate_bf16=False):
    """Correction energy (Hartree) and forces (Hartree/bohr, (N, 3)) of one
    frame.  x_bohr: (N, 3) positions, box_bohr: (3,)."""
    x = np.asarray(x_bohr, float).reshape(-1, 3) * ANG
    box = np.asarray(box_bohr, float) * ANG
    elem = elements(n_mols)
    G, parts = descriptors(x, elem, box, model.desc, with_grad=True)
    g = (G - model.mean[elem]) / model.std[elem]
    E = 0.0
    dEdG = np.zeros_like(G)
    for e in (0, 1):
        sel = elem == e
        Ee, dg = mlp(model, e, g[sel], theta, emulate_bf16)
        E += model.e_scale * Ee.sum()
        dEdG[sel] = model.e_scale * dg / model.std[e]
    gx = grad_positions(parts, dEdG, x.shape[0])
    F = -ANG * gx
    if return_parts:
        return E, F, (G, g, parts, elem)
    return E, F


def loss_and_grad(model, frames, targets, box_bohr, n_mols, theta=None):
    """L = mean over frames, atoms, components of (F_model - T)^2 and dL/dtheta
    by the construction of the GPU kernels: R = F_model - T, the tangent u of
    the normalised descriptor along R (times Angstrom/bohr), J = sum_i
    grad_g E_i . u_i, dL/dtheta = -(2/n) dJ/dtheta with u held fixed."""
    th = model.theta if theta is None else np.asarray(theta, float)
    grad = np.zeros_like(th)
    L, n = 0.0, 0
    for X, T in zip(frames, targets):
        E, F, (G, g, parts, elem) = energy_forces(model, X, box_bohr, n_mols, th, True)
        R = F - np.asarray(T, float).reshape(-1, 3)
        L += np.sum(R * R)
        n += R.size
        u = ANG * tangent(parts, R, G.shape[0]) / model.std[elem]
        for e in (0, 1):
            sel = elem == e
            grad[e * NP_ELEM:(e + 1) * NP_ELEM] += model.e_scale * _jvp_param_grad(
                model, e, g[sel], u[sel], th)
    return L / n, -(2.0 / n) * grad


def _jvp_param_grad(model, e, g, u, theta):
    """d/dtheta of sum_i grad_g E_i(g_i) . u_i (E without e_scale)."""
    W1, b1, W2, b2, w3, b3 = model.layers(e, theta)
    z1 = g @ W1.T + b1
    a1 = np.tanh(z1)
    s1 = 1 - a1 * a1
    zd1 = u @ W1.T
    ad1 = s1 * zd1
    z2 = a1 @ W2.T + b2
    a2 = np.tanh(z2)
    s2 = 1 - a2 * a2
    zd2 = ad1 @ W2.T
    ad2 = s2 * zd2
    # J = sum_i w3 . ad2_i
    gw3 = ad2.sum(0)
    d_ad2 = np.broadcast_to(w3, ad2.shape)
    d_zd2 = s2 * d_ad2
    d_s2 = zd2 * d_ad2
    d_a2 = -2 * a2 * d_s2
    d_z2 = s2 * d_a2
    gW2 = d_zd2.T @ ad1 + d_z2.T @ a1
    gb2 = d_z2.sum(0)
    d_ad1 = d_zd2 @ W2
    d_a1 = d_z2 @ W2
    d_zd1 = s1 * d_ad1
    d_s1 = zd1 * d_ad1
    d_a1 = d_a1 - 2 * a1 * d_s1
    d_z1 = s1 * d_a1
    gW1 = d_zd1.T @ u + d_z1.T @ g
    gb1 = d_z1.sum(0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [0.0]])


def rdf(x_bohr, box_bohr, n_mols, r_max_angs, n_bins):
    """Pair-distance histograms O-O and O-H (counts per bin of minimum-image
    distances in Angstrom, O-H counting every (O, H) pair once)."""
    x = np.asarray(x_bohr, float).reshape(-1, 3) * ANG
    box = np.asarray(box_bohr, float) * ANG
    O = x[0::3]
    H = np.concatenate([x[1::3], x[2::3]])
    edges = np.linspace(0.0, r_max_angs, n_bins + 1)

    def hist(a, b, same):
        d = b[None, :, :] - a[:, None, :]
        d -= box * np.round(d / box)
        r = np.sqrt(np.einsum("ijk,ijk->ij", d, d))
        if same:
            r = r[np.triu_indices(len(a), 1)]
        else:
            r = r.ravel()
        return np.histogram(r, bins=edges)[0]

    return hist(O, O, True), hist(O, H, False)
