"""Host-side starts and position files of the library (no GPU):
initialise_fcc / initialise_orth (initialise.cpp:20-98, :334-541), the initial
momenta (simulation.cpp:826-871) and the position file (initialise.cpp:107-135,
water.cpp:226-238).

Pinned three ways: bit for bit against the reference's own initialise.cpp
compiled into oracle/_ref (same seed, same mt19937 stream), by the geometry
the lattices must have (bond length, angle, ice rules, zero dipole, density),
and for the momenta by the zero centre-of-mass momentum and the stream
position after the lattice draws."""
import numpy as np
import pytest

import oracle
from qtip4pf_water_amd import HostRng, QtpError, initialise_fcc, read_positions, write_positions

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")

ANGS2BOHR = 1.0 / 0.529177210903
ROHEQ = 0.9419 * ANGS2BOHR
THETA_EQ = 107.4 * 3.14159265359 / 180.0
MH2O_G = (1.0080 * 2.0 + 15.9994) / 6.0221409e23  # grams per molecule


def _mols(pos, n_beads=1):
    return pos.reshape(n_beads, -1, 3, 3)


def _mi(r, box):
    L = np.asarray(box)
    return r - L * np.round(r / L)


def _density(n_mols, box):
    vol_cm3 = np.prod(np.asarray(box) / ANGS2BOHR * 1e-8)
    return n_mols * MH2O_G / vol_cm3


@needs_ref
@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 3, 4), (3, 3, 3)])
@pytest.mark.parametrize("rho", [1.0, 0.997])
def test_fcc_bit_exact_vs_reference(dims, rho):
    got, box = initialise_fcc(*dims, rho, n_beads=3)
    want, wbox = oracle.ref_lattice(False, *dims, rho, n_beads=3)
    assert np.array_equal(got, want)
    assert box == wbox


def test_fcc_geometry():
    nx, ny, nz = 3, 2, 4
    pos, box = initialise_fcc(nx, ny, nz, 0.997, n_beads=2)
    m = _mols(pos, 2)
    assert np.array_equal(m[0], m[1])
    n = 4 * nx * ny * nz
    assert m.shape[1] == n
    assert _density(n, box) == pytest.approx(0.997, rel=1e-12)
    r1, r2 = m[0, :, 1] - m[0, :, 0], m[0, :, 2] - m[0, :, 0]
    assert np.allclose(np.linalg.norm(r1, axis=1), ROHEQ, rtol=1e-14)
    assert np.allclose(np.linalg.norm(r2, axis=1), ROHEQ, rtol=1e-14)
    cosang = np.sum(r1 * r2, axis=1) / ROHEQ ** 2
    assert np.allclose(cosang, np.cos(THETA_EQ), atol=1e-14)
    # oxygens: the fcc lattice, 12 nearest neighbours at a / sqrt(2)
    a = box[0] / nx
    O = m[0, :, 0]
    D = np.linalg.norm(_mi(O[:, None] - O[None], box), axis=2)
    np.fill_diagonal(D, np.inf)
    assert np.allclose(np.sort(D, axis=1)[:, :12], a / np.sqrt(2.0), rtol=1e-12)
    assert np.all(np.sort(D, axis=1)[:, 12] > a / np.sqrt(2.0) * 1.1)


@needs_ref
@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 1, 1), (1, 1, 2), (2, 2, 2), (3, 2, 2)])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_orth_bit_exact_vs_reference(dims, seed):
    want, wbox = oracle.ref_lattice(True, *dims, 0.92, n_beads=2, seed=seed)
    got, box, _ = HostRng(seed).initialise_orth(*dims, 0.92, n_beads=2)
    assert np.array_equal(got, want)
    assert box == wbox


@needs_ref
def test_draws_after_orth_continue_the_reference_stream():
    """The lattice consumes exactly the reference's draws: momenta drawn from
    the same generator afterwards agree bit for bit."""
    dims, seed, P, T = (2, 2, 1), 11, 4, 273.0
    _, _ = oracle.ref_lattice(True, *dims, 0.92, n_beads=P, seed=seed)
    n = 8 * 4
    R = oracle.ref()
    # beta_n as water.cpp:77-80 derives it (K2au = kb / eh)
    K2au = 1.38064852e-23 / 4.3597447222071e-18
    R.ref_set_beta_n(1.0 / (K2au * T) / P)
    want = np.zeros(P * 9 * n)
    R.ref_draw_momenta(oracle._ptr(want))
    rng = HostRng(seed)
    rng.initialise_orth(*dims, 0.92, n_beads=P)
    got = rng.draw_momenta(n, P, T)
    assert np.array_equal(got, want)


def test_orth_geometry_and_ice_rules():
    nx, ny, nz = 2, 2, 2
    pos, box, n_outer = HostRng(5).initialise_orth(nx, ny, nz, 0.92)
    assert n_outer >= 1
    m = _mols(pos)[0]
    n = 8 * nx * ny * nz
    assert m.shape[0] == n
    assert _density(n, box) == pytest.approx(0.92, rel=1e-12)
    O = m[:, 0]
    D = np.linalg.norm(_mi(O[:, None] - O[None], box), axis=2)
    np.fill_diagonal(D, np.inf)
    rnn = np.sort(D, axis=1)[0, 0]
    nbr = np.argsort(D, axis=1)[:, :4]
    assert np.allclose(np.sort(D, axis=1)[:, :4], rnn, rtol=1e-12)  # tetrahedral network
    # each hydrogen at roheq from its oxygen, on the line to a neighbouring oxygen
    bonds = set()
    for i in range(n):
        for h in (1, 2):
            r = _mi(m[i, h] - O[i], box)
            assert np.linalg.norm(r) == pytest.approx(ROHEQ, rel=1e-12)
            u = r / np.linalg.norm(r)
            cos_to = [np.dot(u, _mi(O[j] - O[i], box)) / rnn for j in nbr[i]]
            k = int(np.argmax(cos_to))
            assert cos_to[k] == pytest.approx(1.0, abs=1e-12)
            bonds.add(frozenset((i, int(nbr[i][k]))))
    # ice rules: 2n hydrogens on 2n distinct O-O bonds, one hydrogen per bond
    assert len(bonds) == 2 * n
    # zero net dipole: sum of unit H-O-H bisectors
    bis = _mi(m[:, 1] - O, box) + _mi(m[:, 2] - O, box)
    assert np.all(np.abs((bis / np.linalg.norm(bis, axis=1)[:, None]).sum(0)) < 1e-10)


def test_orth_outer_bound_reports_error():
    # seed 1 on 3x2x2 needs more than one fresh assignment (see the reference
    # comparison above); a bound of one attempt must fail, not loop
    with pytest.raises(QtpError):
        HostRng(1).initialise_orth(3, 2, 2, 0.92, max_outer=1)


def test_momenta_statistics():
    n, P, T = 64, 8, 300.0
    p = HostRng(9).draw_momenta(n, P, T).reshape(P, n, 3, 3)
    assert np.allclose(p.sum(axis=(0, 1, 2)), 0.0, atol=1e-9)
    # variance m / beta_n per component (bead temperature P*T)
    K2au = 1.38064852e-23 / 4.3597447222071e-18
    beta_n = 1.0 / (K2au * T) / P
    me = 9.10938356e-31
    mo = 0.001 * 15.9994 / (6.0221409e23 * me)
    mh = 0.001 * 1.0080 / (6.0221409e23 * me)
    vo = p[:, :, 0].var() * beta_n / mo
    vh = p[:, :, 1:].var() * beta_n / mh
    assert vo == pytest.approx(1.0, rel=0.1)
    assert vh == pytest.approx(1.0, rel=0.1)


@needs_ref
def test_position_file_round_trip_and_reference_reader(tmp_path):
    n, P = 16, 4
    rng = np.random.default_rng(3)
    pos = rng.uniform(-20.0, 20.0, P * 9 * n)
    f_all = tmp_path / "all.txt"
    write_positions(f_all, pos, n, P)
    back, nb = read_positions(f_all, n, P)
    assert nb == P and np.array_equal(back, pos)
    # one bead: copied to every bead, as initialise_fromfile does
    f_one = tmp_path / "one.txt"
    write_positions(f_one, pos[: 9 * n], n, 1)
    one, nb = read_positions(f_one, n, P)
    assert nb == 1
    assert np.array_equal(one.reshape(P, -1), np.tile(pos[: 9 * n], (P, 1)))
    # the reference reader on the same files: equal on one bead; on every
    # bead it keeps only bead 0 (initialise.cpp:128-132)
    R = oracle.ref()
    box = np.zeros(3)
    R.ref_set_lattice(0, 2, 1, 2, 1.0, P, oracle._ptr(box))  # 16 molecules
    ref_one = np.zeros(P * 9 * n)
    R.ref_initialise_fromfile(str(f_one).encode(), oracle._ptr(ref_one))
    assert np.array_equal(ref_one, one)
    ref_all = np.zeros(P * 9 * n)
    R.ref_initialise_fromfile(str(f_all).encode(), oracle._ptr(ref_all))
    assert np.array_equal(ref_all.reshape(P, -1), np.tile(pos[: 9 * n], (P, 1)))


def test_position_file_rejects_wrong_length(tmp_path):
    f = tmp_path / "bad.txt"
    f.write_text("0 0 0\n1 1 1\n")
    with pytest.raises(QtpError):
        read_positions(f, 1, 2)
    f.write_text("0 0 x\n" * 3)
    with pytest.raises(QtpError):
        read_positions(f, 1, 1)
