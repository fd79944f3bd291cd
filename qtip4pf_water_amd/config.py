"""Host-side set-up: reference constants, derived system parameters, seeded
water boxes, momentum sampling and the reference's text formats.

The derived parameters restate water.cpp:74-157 for a box given directly
(the reference derives the box from an fcc/orth lattice, water.cpp:90-120;
a cubic box of arbitrary molecule count cannot be expressed that way, so the
caller supplies it).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# qtip4pfparams.h:43-133 (pi is the reference's truncated value, :45)
PI_REF = 3.14159265359
ANGS2BOHR = 1.0 / 0.529177210903
ME = 9.10938356e-31
KB = 1.38064852e-23
NA = 6.0221409e23
EH = 4.3597447222071e-18
KCAL2AU = 1.5936e-3
MHGMOL = 1.0080
MOGMOL = 15.9994
MO = 0.001 * MOGMOL / (NA * ME)
MH = 0.001 * MHGMOL / (NA * ME)
K2AU = KB / EH
HBAR_SI = 1.054571817e-34
AU2FS = HBAR_SI * 1e15 / EH
FS2AU = 1.0 / AU2FS
MH2O = (MHGMOL * 2.0 + MOGMOL) / NA
R_CUT = 9.0 * ANGS2BOHR
R_CUT_SQ = R_CUT * R_CUT
R_SKIN = 1.30 * R_CUT
R_SKIN_SQ = R_SKIN * R_SKIN
GAMMA_VS = 0.73612
EPSLJ = 0.1852 * KCAL2AU
SIGMALJ = 3.1589 * ANGS2BOHR
ROHEQ = 0.9419 * ANGS2BOHR
THETA_EQ = 107.4 * PI_REF / 180.0
QO = -1.1128
QH = -0.5 * QO


@dataclass
class System:
    """The reference's process globals (water.cpp:29-38) for one system."""
    n_mols: int
    n_beads: int
    boxl: tuple
    T_K: float = 298.0
    dtfs: float = 0.25
    eps: float = 1e-6
    epsilonsc: float = 1e-5
    n_atoms: int = field(init=False)
    n_dof: int = field(init=False)

    def __post_init__(self):
        if self.n_beads != 1 and self.n_beads % 2 == 1:
            raise ValueError("n_beads should be even for Suzuki-Chin calculations, "
                             "or 1 for classical (water.cpp:69-72)")
        self.boxl = tuple(float(x) for x in self.boxl)
        self.n_atoms = 3 * self.n_mols
        self.n_dof = 3 * self.n_atoms

    @property
    def vol(self):
        return self.boxl[0] * self.boxl[1] * self.boxl[2]


def cubic_box_angs(n_mols: int, rho_gcc: float = 0.997) -> float:
    """Edge (Angstrom) of a cubic box holding n_mols waters at rho_gcc."""
    return 1e8 * (n_mols * MH2O / rho_gcc) ** (1.0 / 3.0)


def _random_rotations(n: int, rng: np.random.Generator) -> np.ndarray:
    """Uniform rotations on SO(3) from uniform unit quaternions (Shoemake)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    qw = a * np.sin(2 * np.pi * u2)
    qx = a * np.cos(2 * np.pi * u2)
    qy = b * np.sin(2 * np.pi * u3)
    qz = b * np.cos(2 * np.pi * u3)
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (qy * qy + qz * qz)
    R[:, 0, 1] = 2 * (qx * qy - qz * qw)
    R[:, 0, 2] = 2 * (qx * qz + qy * qw)
    R[:, 1, 0] = 2 * (qx * qy + qz * qw)
    R[:, 1, 1] = 1 - 2 * (qx * qx + qz * qz)
    R[:, 1, 2] = 2 * (qy * qz - qx * qw)
    R[:, 2, 0] = 2 * (qx * qz - qy * qw)
    R[:, 2, 1] = 2 * (qy * qz + qx * qw)
    R[:, 2, 2] = 1 - 2 * (qx * qx + qy * qy)
    return R


def lattice_water_box(n_side, L_angs, seed=0, jitter_bohr=0.0):
    """n_side^3 rigid waters (r_OH = r_eq, HOH = theta_eq) with O on a cubic
    grid of spacing L/n_side and uniformly random orientations.  Returns the
    bead-0 position array in reference layout (9 per molecule, bohr) and the
    box edge in bohr.  n_side may also be a tuple (nx, ny, nz)."""
    if np.isscalar(n_side):
        n_side = (int(n_side),) * 3
    nx, ny, nz = n_side
    L = L_angs * ANGS2BOHR if np.isscalar(L_angs) else None
    Ls = np.array([L, L, L]) if L is not None else np.asarray(L_angs, float) * ANGS2BOHR
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    gx, gy, gz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    O = np.stack([(gx.ravel() + 0.5) * Ls[0] / nx,
                  (gy.ravel() + 0.5) * Ls[1] / ny,
                  (gz.ravel() + 0.5) * Ls[2] / nz], axis=1)
    half = 0.5 * THETA_EQ
    h1 = ROHEQ * np.array([math.sin(half), 0.0, math.cos(half)])
    h2 = ROHEQ * np.array([-math.sin(half), 0.0, math.cos(half)])
    R = _random_rotations(n, rng)
    H1 = O + R @ h1
    H2 = O + R @ h2
    if jitter_bohr > 0:
        O = O + rng.normal(0, jitter_bohr, O.shape)
        H1 = H1 + rng.normal(0, jitter_bohr, H1.shape)
        H2 = H2 + rng.normal(0, jitter_bohr, H2.shape)
    q = np.stack([O, H1, H2], axis=1).reshape(-1)
    return np.ascontiguousarray(q), tuple(float(x) for x in Ls)


def water_box(n_mols, L_angs, seed=0, jitter_bohr=0.0):
    """n_mols rigid waters in a cubic box of edge L_angs: a full cubic
    lattice when n_mols is a cube, otherwise n_mols sites of the smallest
    cubic lattice that holds them, chosen at random (seeded).  Returns the
    reference-layout positions and the box edge in bohr."""
    n_side = 1
    while n_side ** 3 < n_mols:
        n_side += 1
    q, box = lattice_water_box(n_side, L_angs, seed=seed, jitter_bohr=jitter_bohr)
    if n_side ** 3 == n_mols:
        return q, box
    keep = np.sort(np.random.default_rng(seed + 7919).permutation(n_side ** 3)[:n_mols])
    q = q.reshape(-1, 9)[keep].reshape(-1)
    return np.ascontiguousarray(q), box


def replicate_beads(q0: np.ndarray, n_beads: int) -> np.ndarray:
    """Every bead a copy of bead 0 (initialise.cpp:128-132)."""
    return np.ascontiguousarray(np.tile(q0, n_beads))


def charges_masses(n_mols: int):
    """simulation.cpp:811-822: M-site charge in the O slot; per-dof masses."""
    charge = np.tile([QO, QH, QH], n_mols).astype(np.float64)
    mass = np.tile(np.repeat([MO, MH, MH], 3), n_mols).astype(np.float64)
    return charge, mass


def sample_momenta(n_mols, n_beads, beta_n, seed=0):
    """Maxwell-Boltzmann momenta at beta_n with the centre-of-mass momentum of
    the whole extended system removed (simulation.cpp:848-871).  Uses numpy's
    generator (the reference's libstdc++ stream is machine dependent,
    water.cpp:161-164)."""
    rng = np.random.default_rng(seed)
    sd = np.repeat([math.sqrt(MO / beta_n), math.sqrt(MH / beta_n), math.sqrt(MH / beta_n)], 3)
    p = rng.standard_normal((n_beads, n_mols, 9)) * sd
    n_atoms = 3 * n_mols
    pcom = p.reshape(n_beads, n_atoms, 3).sum(axis=(0, 1)) / (n_atoms * n_beads)
    p = p.reshape(n_beads, n_atoms, 3) - pcom
    return np.ascontiguousarray(p.reshape(-1))


def write_positions(path: str, q: np.ndarray) -> None:
    """initialise_fromfile format (initialise.cpp:100-135): x y z per atom."""
    with open(path, "w") as f:
        for row in np.asarray(q).reshape(-1, 3):
            f.write(f"{row[0]!r} {row[1]!r} {row[2]!r}\n")


def read_positions(path: str, n_beads: int = 1) -> np.ndarray:
    """Read the initialise_fromfile format; beads 1..P-1 become copies of
    bead 0 as the reference does (initialise.cpp:128-132)."""
    rows = []
    with open(path) as f:
        for line in f:
            if line.strip():
                rows.append([float(x) for x in line.split()[:3]])
    q0 = np.asarray(rows, dtype=np.float64).reshape(-1)
    return replicate_beads(q0, n_beads)


def read_config(path: str) -> dict:
    """Positional config.txt reader (water.cpp:44-66): names are ignored."""
    keys = ["n_beads", "T_K", "epsilon_sc", "n_cell_x", "n_cell_y", "n_cell_z", "rho_inp",
            "dtfs", "eps_ewald", "n_equil", "n_samp", "is_spline", "cell_type", "from_file"]
    casts = [int, float, float, int, int, int, float, float, float, int, int, int, str, str]
    toks = open(path).read().split()
    out, k = {}, 0
    for key, cast in zip(keys, casts):
        out[key] = cast(toks[k + 1])
        k += 2
    if out["from_file"] == "yes":
        out["filename"] = toks[k]
    return out


def free_ring_polymer_beads(q0, n_beads, beta_n, seed=0):
    """Beads = centroid copies plus Gaussian free-ring-polymer normal-mode
    displacements: for mode k >= 1 (frequency w_k = 2 sin(k pi/P)/beta_n) the
    normal-mode coordinate is N(0, 1/sqrt(beta_n m w_k^2)); the centroid is
    unchanged.  Modes use the orthonormal real transform of fourier.cpp.
    (The reference starts every bead as a copy, initialise.cpp:128-132.)"""
    P = int(n_beads)
    q0 = np.asarray(q0, dtype=np.float64)
    if P == 1:
        return q0.copy()
    n_mols = q0.size // 9
    mass = np.tile(np.repeat([MO, MH, MH], 3), n_mols)
    rng = np.random.default_rng(seed)
    j = np.arange(P)
    C = np.zeros((P, P))
    C[0] = 1.0 / math.sqrt(P)
    for k in range(1, (P + 1) // 2):
        C[k] = math.sqrt(2.0 / P) * np.cos(2 * np.pi * j * k / P)
        C[P - k] = -math.sqrt(2.0 / P) * np.sin(2 * np.pi * j * k / P)
    if P % 2 == 0:
        C[P // 2] = (-1.0) ** j / math.sqrt(P)
    u = np.zeros((P, q0.size))
    u[0] = math.sqrt(P) * q0
    for k in range(1, P):
        wk = 2.0 * math.sin(k * math.pi / P) / beta_n
        u[k] = rng.standard_normal(q0.size) / np.sqrt(beta_n * mass * wk * wk)
    return np.ascontiguousarray((C.T @ u).reshape(-1))


# Batched replicas: molecule counts from {64, 128, 216, 343, 512} at 0.997
# g/cm^3 (cubic boxes), bead counts from {8, 16, 32, 64}, temperatures
# uniform in 250-350 K, one master seed.  Boxes narrower than twice the 9 A
# LJ cutoff (64 and 128 molecules) take the cutoff L/2.
BATCH_SIZES = {64: 12.43, 128: 15.66, 216: 18.64, 343: 21.75, 512: 24.86}


def lj_rcut_for(L_angs):
    """LJ cutoff (bohr) of a cubic box: 9 A, or L/2 when the box is narrower."""
    return min(9.0, 0.5 * L_angs) * ANGS2BOHR


def batch_replicas(n_rep, seed):
    rng = np.random.default_rng(seed)
    reps = []
    for r in range(n_rep):
        n = int(rng.choice(sorted(BATCH_SIZES)))
        L = BATCH_SIZES[n]
        reps.append(dict(n_mols=n, L_angs=L, P=int(rng.choice([8, 16, 32, 64])),
                         T_K=float(rng.uniform(250.0, 350.0)), seed=int(seed * 1000 + r),
                         lj_rcut=lj_rcut_for(L)))
    return reps


def replica_state(r):
    """Seeded start of one replica: lattice molecules, free-ring-polymer beads,
    Maxwell-Boltzmann momenta at its bead temperature."""
    q0, box = water_box(r["n_mols"], r["L_angs"], seed=r["seed"])
    beta_n = 1.0 / (K2AU * r["T_K"]) / r["P"]
    q = free_ring_polymer_beads(q0, r["P"], beta_n, seed=r["seed"] + 1)
    p = sample_momenta(r["n_mols"], r["P"], beta_n, seed=r["seed"] + 2)
    return q, p, box
