"""ctypes binding of the CPU oracle (qtp_oracle.c) and of the reference build
in oracle/_ref (the reference's neighbour.cpp and thermostat.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  Never imported by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libqtp_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libqtp_ref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OqSys(C.Structure):
    _fields_ = [
        ("n_mols", C.c_int), ("n_atoms", C.c_int), ("n_dof", C.c_int),
        ("n_beads", C.c_int), ("list_size", C.c_int),
        ("boxl", C.c_double * 3), ("vol", C.c_double), ("min_box", C.c_double),
        ("max_box", C.c_double),
        ("T_K", C.c_double), ("beta", C.c_double), ("beta_n", C.c_double),
        ("piovern", C.c_double), ("omega_p", C.c_double), ("omega_psq", C.c_double),
        ("dt", C.c_double), ("halfdt", C.c_double), ("epsilonsc", C.c_double),
        ("eps", C.c_double), ("acc", C.c_double), ("rho", C.c_double),
        ("kwald_prefactor", C.c_double),
        ("sd_ox", C.c_double), ("sd_h", C.c_double),
        ("use_spline", C.c_int),
        ("r_cut", C.c_double), ("r_cut_sq", C.c_double), ("r_skin", C.c_double),
        ("r_skin_sq", C.c_double), ("e_cut", C.c_double),
        ("ew_kappa", C.c_double), ("ew_rcut", C.c_double), ("ew_kcut", C.c_double),
        ("ew_kmax", C.c_int),
    ]


class OqEwald(C.Structure):
    _fields_ = [("kappa", C.c_double), ("rcut", C.c_double), ("rcut_sq", C.c_double),
                ("k_cut", C.c_double), ("k_cut_sq", C.c_double), ("prefactor", C.c_double),
                ("k_max", C.c_int)]


_CONST_FIELDS = ["pi", "angs2bohr", "me", "kb", "Na", "eh", "kcal2au", "mo", "mh", "K2au",
                 "au2fs", "fs2au", "r_cut", "r_cut_sq", "r_skin", "r_skin_sq", "gamma_vs",
                 "epslj", "sigmalj", "sigmalj6", "sigmalj12", "e_cut", "roheq", "ar", "ar2",
                 "ar3", "ar4fac", "Dr", "theta_eq", "ktheta", "qo", "qh", "mh2o"]


class OqConsts(C.Structure):
    _fields_ = [(n, C.c_double) for n in _CONST_FIELDS]


def build(force: bool = False) -> None:
    """Compile the oracle (make rebuilds it when its sources are newer)."""
    subprocess.check_call(["make", "-s", "-C", HERE] + (["-B"] if force else []) + ["oracle"])


def build_ref() -> None:
    subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        L.oq_sys_init.argtypes = [C.POINTER(OqSys), C.c_int, C.c_int, _dp, C.c_double,
                                  C.c_double, C.c_double, C.c_double]
        L.oq_ewald_params.argtypes = [C.POINTER(OqSys), C.POINTER(OqEwald)]
        L.oq_sys_set_lj_rcut.argtypes = [C.POINTER(OqSys), C.c_double]
        L.oq_get_consts.argtypes = [C.POINTER(OqConsts)]
        L.oq_charges_masses.argtypes = [C.POINTER(OqSys), _dp, _dp]
        L.oq_lj_oxygen.restype = C.c_double
        L.oq_lj_oxygen.argtypes = [_dp, C.c_double, _dp, C.c_int, C.c_int]
        L.oq_intra_morse_bond.restype = C.c_double
        L.oq_intra_morse_bond.argtypes = [_dp, C.c_double, _dp, C.c_int, C.c_int]
        L.oq_intra_harm_angle.restype = C.c_double
        L.oq_intra_harm_angle.argtypes = [_dp, _dp, C.c_double, C.c_double, _dp, C.c_int]
        L.oq_get_forces.restype = C.c_double
        L.oq_get_forces.argtypes = [C.POINTER(OqSys), _dp, _dp, _dp, _ip, _ip, C.c_int]
        L.oq_get_forces_parts.restype = C.c_double
        L.oq_get_forces_parts.argtypes = [C.POINTER(OqSys), _dp, _dp, _dp, _ip, _ip, C.c_int, _dp]
        L.oq_set_exact_energy.restype = None
        L.oq_set_exact_energy.argtypes = [C.c_int]
        L.oq_force_eval_sc.restype = C.c_double
        L.oq_force_eval_sc.argtypes = [C.POINTER(OqSys), _dp, _dp, _dp, _dp, _ip, _ip, _dp, _dp,
                                       C.c_int]
        L.oq_neighbour_list.restype = C.c_int
        L.oq_neighbour_list.argtypes = [C.POINTER(OqSys), _dp, _ip, _ip, _dp, C.c_int]
        L.oq_ftransform.argtypes = [C.POINTER(OqSys), _dp, C.c_int, C.c_int]
        L.oq_pile_g.restype = C.c_double
        L.oq_pile_g.argtypes = [C.POINTER(OqSys), _dp, _dp, _dp]
        L.oq_update_sc.restype = C.c_double
        L.oq_update_sc.argtypes = [C.POINTER(OqSys), _dp, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _dp,
                                   C.c_int, _dp, _dp, _dp, C.c_int, _ip]
        L.oq_update_class.restype = C.c_double
        L.oq_update_class.argtypes = [C.POINTER(OqSys), _dp, _dp, _dp, _dp, _dp, _ip, _ip, _dp, _ip]
        L.oq_spline_set.restype = C.c_int
        L.oq_spline_set.argtypes = [_dp, C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_double,
                                    C.c_double]
        L.oq_spline_fit.restype = C.c_int
        L.oq_spline_fit.argtypes = [_dp, C.c_int, C.c_double, _dp]
        L.oq_spline_eval.restype = C.c_double
        L.oq_spline_eval.argtypes = [_dp, C.c_int, C.c_double, C.c_double, C.c_double, _dp]
        L.oq_get_kcv.restype = C.c_double
        L.oq_get_kcv.argtypes = [C.POINTER(OqSys), _dp, _dp]
        L.oq_get_conserved.restype = C.c_double
        L.oq_get_conserved.argtypes = [C.POINTER(OqSys), _dp, _dp, _dp, C.c_double]
        _lib = L
    return _lib


_exact = [0]


class exact_energy:
    """Context manager: inside it every oracle bead energy is summed with an
    80-bit accumulator (oq_set_exact_energy) instead of the reference's FP64
    running sum; the previous setting comes back on exit.  Forces are
    unchanged."""

    def __enter__(self):
        self._prev = _exact[0]
        _exact[0] = 1
        lib().oq_set_exact_energy(1)
        return self

    def __exit__(self, *exc):
        _exact[0] = self._prev
        lib().oq_set_exact_energy(self._prev)
        return False


def spline_set(oo, oh, hh, a=0.0, h=0.01) -> None:
    """Load the O-O, O-H, H-H spline tables (the reference's spline files:
    values at a + j h) for force calls of oracles with use_spline set."""
    t = [np.ascontiguousarray(x, dtype=np.float64) for x in (oo, oh, hh)]
    rc = lib().oq_spline_set(_ptr(t[0]), t[0].size, _ptr(t[1]), t[1].size, _ptr(t[2]), t[2].size,
                             float(a), float(h))
    if rc != 0:
        raise ValueError("spline tables need at least 5 values each")


def spline_fit(f, h):
    f = np.ascontiguousarray(f, dtype=np.float64)
    beta = np.zeros(f.size + 2)
    if lib().oq_spline_fit(_ptr(f), f.size, float(h), _ptr(beta)) != 0:
        raise ValueError("spline fit needs at least 5 values")
    return beta


def spline_eval(beta, n, a, h, x):
    """(value, derivative) of a fitted spline."""
    d = C.c_double()
    v = lib().oq_spline_eval(_ptr(np.ascontiguousarray(beta, dtype=np.float64)), int(n), float(a),
                             float(h), float(x), C.byref(d))
    return v, d.value


def consts() -> dict:
    c = OqConsts()
    lib().oq_get_consts(C.byref(c))
    return {n: getattr(c, n) for n in _CONST_FIELDS}


class Oracle:
    """Reference-shaped CPU restatement for one system (the reference's globals)."""

    def __init__(self, n_mols, n_beads, boxl, T_K=298.0, dtfs=0.25, eps=1e-6, epsilonsc=1e-5,
                 lj_rcut=None, ewald=None):
        self.sys = OqSys()
        b = np.ascontiguousarray(np.asarray(boxl, dtype=np.float64).reshape(3))
        lib().oq_sys_init(C.byref(self.sys), int(n_mols), int(n_beads), _ptr(b), float(T_K),
                          float(dtfs), float(eps), float(epsilonsc))
        if lj_rcut is not None:
            lib().oq_sys_set_lj_rcut(C.byref(self.sys), float(lj_rcut))
        if ewald is not None:   # {"kappa", "rcut", "k_cut", "k_max"} instead of derived
            self.sys.ew_kappa, self.sys.ew_rcut = float(ewald["kappa"]), float(ewald["rcut"])
            self.sys.ew_kcut, self.sys.ew_kmax = float(ewald["k_cut"]), int(ewald["k_max"])
        s = self.sys
        self.n_mols, self.n_atoms, self.n_dof = s.n_mols, s.n_atoms, s.n_dof
        self.n_beads, self.list_size = s.n_beads, s.list_size
        self.charge = np.zeros(self.n_atoms)
        self.mass = np.zeros(self.n_dof)
        lib().oq_charges_masses(C.byref(s), _ptr(self.charge), _ptr(self.mass))

    @property
    def use_spline(self) -> bool:
        return bool(self.sys.use_spline)

    @use_spline.setter
    def use_spline(self, on: bool):
        self.sys.use_spline = 1 if on else 0

    def ewald(self) -> dict:
        e = OqEwald()
        lib().oq_ewald_params(C.byref(self.sys), C.byref(e))
        return {f: getattr(e, f) for f, _ in OqEwald._fields_}

    # -- lists -------------------------------------------------------------
    def new_lists(self):
        return (np.zeros(self.n_beads * self.n_mols, np.int32),
                np.zeros(self.n_beads * self.list_size, np.int32),
                np.zeros(self.n_beads * 3 * self.n_mols))

    def neighbour_list(self, q, point, lst, disp, bead=0) -> int:
        return lib().oq_neighbour_list(C.byref(self.sys), _ptr(q), _ptr(point, _ip),
                                       _ptr(lst, _ip), _ptr(disp), int(bead))

    def all_lists(self, q):
        point, lst, disp = self.new_lists()
        for b in range(self.n_beads):
            self.neighbour_list(q, point, lst, disp, b)
        return point, lst, disp

    # -- forces ------------------------------------------------------------
    def get_forces(self, q, F, point, lst, bead=0, parts=False):
        if parts:
            out = np.zeros(6)
            e = lib().oq_get_forces_parts(C.byref(self.sys), _ptr(q), _ptr(F), _ptr(self.charge),
                                          _ptr(point, _ip), _ptr(lst, _ip), int(bead), _ptr(out))
            return e, out
        return lib().oq_get_forces(C.byref(self.sys), _ptr(q), _ptr(F), _ptr(self.charge),
                                   _ptr(point, _ip), _ptr(lst, _ip), int(bead))

    def force_eval_sc(self, q, Fsc, fphys, point, lst, nthreads=1):
        sc_pot = C.c_double(0.0)
        e = lib().oq_force_eval_sc(C.byref(self.sys), _ptr(q), _ptr(Fsc), _ptr(self.charge),
                                   _ptr(self.mass), _ptr(point, _ip), _ptr(lst, _ip),
                                   C.byref(sc_pot), _ptr(fphys), int(nthreads))
        return e, sc_pot.value

    # -- normal modes / thermostat / steps --------------------------------
    def ftransform(self, a, length, forwards):
        lib().oq_ftransform(C.byref(self.sys), _ptr(a), int(length), 1 if forwards else 0)

    def pile_g(self, p, dev):
        return lib().oq_pile_g(C.byref(self.sys), _ptr(p), _ptr(self.mass), _ptr(dev))

    def n_deviates(self):
        return 3 + (self.n_beads - 1) * self.n_dof

    def update_sc(self, q, p, F, fphys, point, lst, disp, is_nvt, dev, heat, nthreads=1):
        """Returns (energy, heat, sc_pot, n_rebuilds)."""
        h = C.c_double(heat)
        sc = C.c_double(0.0)
        nr = C.c_int(0)
        if dev is None:
            dev = np.zeros(1)
        e = lib().oq_update_sc(C.byref(self.sys), _ptr(q), _ptr(p), _ptr(self.mass), _ptr(F),
                               _ptr(fphys), _ptr(self.charge), _ptr(point, _ip), _ptr(lst, _ip),
                               _ptr(disp), 1 if is_nvt else 0, _ptr(dev), C.byref(h),
                               C.byref(sc), int(nthreads), C.byref(nr))
        return e, h.value, sc.value, nr.value

    def update_class(self, q, p, F, point, lst, disp):
        nr = C.c_int(0)
        e = lib().oq_update_class(C.byref(self.sys), _ptr(q), _ptr(p), _ptr(self.mass), _ptr(F),
                                  _ptr(self.charge), _ptr(point, _ip), _ptr(lst, _ip), _ptr(disp),
                                  C.byref(nr))
        return e, nr.value

    def get_kcv(self, q, f):
        return lib().oq_get_kcv(C.byref(self.sys), _ptr(q), _ptr(f))

    def get_conserved(self, q, p, energy):
        return lib().oq_get_conserved(C.byref(self.sys), _ptr(q), _ptr(p), _ptr(self.mass),
                                      float(energy))


# -- the reference build (oracle/_ref) ----------------------------------------
_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        # lazy binding: get_conserved's ftransform (fourier.cpp, not built) is
        # never called and stays unresolved
        L = C.CDLL(REF_SO, mode=os.RTLD_LAZY)
        L.ref_get_kcv.restype = C.c_double
        L.ref_get_kcv.argtypes = [_dp, _dp]
        L.ref_set_globals.argtypes = [C.c_int, C.c_int, _dp, C.c_int, C.c_double, C.c_double,
                                      C.c_double]
        L.ref_seed.argtypes = [C.c_uint]
        L.ref_neighbour_list.argtypes = [_dp, _ip, _ip, _dp, C.c_int]
        L.ref_pile_g.restype = C.c_double
        L.ref_pile_g.argtypes = [_dp, _dp]
        L.ref_bussi.restype = C.c_double
        L.ref_bussi.argtypes = [_dp, _dp]
        L.ref_draw_pile_deviates.argtypes = [_dp]
        L.ref_set_lattice.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, _dp]
        L.ref_initialise.argtypes = [C.c_int, _dp]
        L.ref_initialise_fromfile.argtypes = [C.c_char_p, _dp]
        L.ref_draw_momenta.argtypes = [_dp]
        L.ref_set_beta_n.argtypes = [C.c_double]
        _ref = L
    return _ref


def ref_set_from(o: Oracle) -> None:
    s = o.sys
    b = np.array([s.boxl[0], s.boxl[1], s.boxl[2]])
    ref().ref_set_globals(s.n_mols, s.n_beads, _ptr(b), s.list_size, s.beta, s.beta_n, s.dt)


def ref_lattice(orth: bool, nx: int, ny: int, nz: int, rho: float, n_beads: int = 1,
                seed: int | None = None):
    """The reference's own initialise_fcc / initialise_orth
    (initialise.cpp:20-98, :334-541) with water.cpp's lattice globals; orth
    draws from the harness mt19937 seeded with `seed`.  Returns (positions,
    box)."""
    R = ref()
    box = np.zeros(3)
    R.ref_set_lattice(1 if orth else 0, nx, ny, nz, float(rho), n_beads, _ptr(box))
    if seed is not None:
        R.ref_seed(seed)
    n = n_beads * (72 if orth else 36) * nx * ny * nz
    out = np.zeros(n)
    R.ref_initialise(1 if orth else 0, _ptr(out))
    return out, tuple(box)
