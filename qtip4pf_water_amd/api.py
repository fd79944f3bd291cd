"""Reference-shaped Python interface over the gfx950 library.

`Water` holds what the reference keeps in process globals (water.cpp:29-38)
and exposes the hot-path routines under the reference's names with the same
argument meaning and array layout (indexing.h:6-24):

    get_forces_nospline        forces.cpp:468   (forces.h:16)
    force_eval_sc_nospline     forces.cpp:698   (forces.h:20)
    neighbour_list             neighbour.cpp:12 (neighbour.h:4)
    ftransform                 fourier.cpp:12   (fourier.h:4)
    pile_g                     thermostat.cpp:64 (thermostat.h:6)
    update_sc_nospline         update.cpp:271   (update.h:7)
    update_trajectory_class_nospline  update.cpp:78 (update.h:5)
    get_kcv / get_conserved    path_integrals.cpp:29, :52

Arrays are numpy float64 / int32, modified in place like the reference's.
Random numbers of pile_g are passed in (`dev`) in the order the reference
draws them; `HostRng` draws them from std::mt19937 with libstdc++.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import QtpDerived, QtpError, QtpParams, ListOverflow, dp, ip, ptr


def _f64(a, n=0, name="array"):
    """`a` as passed to the library, which reads or writes n elements of it."""
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
        raise TypeError(f"{name}: expected a C-contiguous float64 numpy array")
    if a.size < n:
        raise ValueError(f"{name}: {a.size} elements, the call uses {n}")
    return a


def _i32(a, n=0, name="array"):
    if not (isinstance(a, np.ndarray) and a.dtype == np.int32 and a.flags.c_contiguous):
        raise TypeError(f"{name}: expected a C-contiguous int32 numpy array")
    if a.size < n:
        raise ValueError(f"{name}: {a.size} elements, the call uses {n}")
    return a


def voigt_to_matrix(v):
    """(xx, yy, zz, xy, xz, yz) -> symmetric 3x3."""
    xx, yy, zz, xy, xz, yz = (float(x) for x in v)
    return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])


class HostRng:
    """std::mt19937 owned by the library; draws pile_g's deviates in the
    reference's order (thermostat.cpp:71, 88-98, 124)."""

    def __init__(self, seed: int):
        self._L = _lib.load()
        h = C.c_void_p()
        rc = self._L.qtp_host_rng_create(C.c_uint32(seed & 0xFFFFFFFF), C.byref(h))
        if rc != 0:
            raise QtpError(rc, "host rng")
        self._h = h

    def draw_pile(self, n_atoms: int, n_beads: int) -> np.ndarray:
        out = np.zeros(3 + (n_beads - 1) * 3 * n_atoms)
        rc = self._L.qtp_host_rng_draw_pile(self._h, int(n_atoms), int(n_beads), ptr(out))
        if rc != 0:
            raise QtpError(rc, "host rng draw")
        return out

    def draw_momenta(self, n_mols: int, n_beads: int, T_K: float) -> np.ndarray:
        """Initial momenta of every bead (simulation.cpp:826-871)."""
        out = np.zeros(n_beads * 9 * n_mols)
        rc = self._L.qtp_host_rng_draw_momenta(self._h, int(n_mols), int(n_beads), float(T_K),
                                               ptr(out))
        if rc != 0:
            raise QtpError(rc, "host rng draw_momenta")
        return out

    def initialise_orth(self, nx: int, ny: int, nz: int, rho_inp: float = 0.92,
                        n_beads: int = 1, max_outer: int = 0):
        """Proton-disordered ice, ice rules and zero dipole, drawn from this
        generator (initialise.cpp:334-541).  Returns (positions, box, number
        of hydrogen assignments tried)."""
        pos = np.zeros(n_beads * 72 * nx * ny * nz)
        box = np.zeros(3)
        n_outer = C.c_long()
        rc = self._L.qtp_initialise_orth(self._h, int(nx), int(ny), int(nz), float(rho_inp),
                                         int(n_beads), int(max_outer), ptr(pos), ptr(box),
                                         C.byref(n_outer))
        if rc != 0:
            raise QtpError(rc, "initialise_orth")
        return pos, tuple(box), n_outer.value

    def __del__(self):
        try:
            if self._h:
                self._L.qtp_host_rng_destroy(self._h)
        except Exception:
            pass


def philox4x32_10(ctr, key):
    L = _lib.load()
    c = np.asarray(ctr, dtype=np.uint32).copy()
    k = np.asarray(key, dtype=np.uint32).copy()
    o = np.zeros(4, dtype=np.uint32)
    rc = L.qtp_philox4x32_10(ptr(c, _lib.u32p), ptr(k, _lib.u32p), ptr(o, _lib.u32p))
    if rc != 0:
        raise QtpError(rc, "philox")
    return o


def initialise_fcc(nx: int, ny: int, nz: int, rho_inp: float = 1.0, n_beads: int = 1):
    """Molecules on an fcc lattice (initialise.cpp:20-98); returns (positions,
    box in bohr) for 4*nx*ny*nz molecules, every bead a copy."""
    pos = np.zeros(n_beads * 36 * nx * ny * nz)
    box = np.zeros(3)
    rc = _lib.load().qtp_initialise_fcc(int(nx), int(ny), int(nz), float(rho_inp), int(n_beads),
                                        ptr(pos), ptr(box))
    if rc != 0:
        raise QtpError(rc, "initialise_fcc")
    return pos, tuple(box)


def read_positions(path, n_mols: int, n_beads: int):
    """Positions in the reference's file format (initialise.cpp:107-135);
    returns (positions, beads read: 1 or n_beads)."""
    pos = np.zeros(n_beads * 9 * n_mols)
    nb = C.c_int()
    rc = _lib.load().qtp_read_positions(os.fsencode(path), int(n_mols), int(n_beads), ptr(pos),
                                        C.byref(nb))
    if rc != 0:
        raise QtpError(rc, f"read_positions: {path}")
    return pos, nb.value


def write_positions(path, pos, n_mols: int, n_beads: int):
    """Every bead, one atom per line, exact on reading back."""
    rc = _lib.load().qtp_write_positions(os.fsencode(path), int(n_mols), int(n_beads),
                                         ptr(_f64(pos, n_beads * 9 * n_mols, "pos")))
    if rc != 0:
        raise QtpError(rc, f"write_positions: {path}")


class Water:
    """One qTIP4P/F system on the GPU (the reference's globals + device state)."""

    def __init__(self, n_mols, n_beads, boxl, T_K=298.0, dtfs=0.25, eps_ewald=1e-6,
                 epsilon_sc=1e-5, list_size=0, ewald=None, lj_rcut=None):
        self._L = _lib.load()
        p = QtpParams()
        p.n_mols = int(n_mols)
        p.n_beads = int(n_beads)
        for k in range(3):
            p.boxl[k] = float(boxl[k])
        p.T_K, p.dtfs, p.eps_ewald, p.epsilon_sc = float(T_K), float(dtfs), float(eps_ewald), float(epsilon_sc)
        p.list_size = int(list_size)
        if ewald is not None:
            p.ewald_kappa, p.ewald_rcut, p.ewald_kcut, p.ewald_kmax = (
                float(ewald["kappa"]), float(ewald["rcut"]), float(ewald["k_cut"]), int(ewald["k_max"]))
        if lj_rcut is not None:
            p.lj_rcut = float(lj_rcut)
        h = C.c_void_p()
        rc = self._L.qtp_create(C.byref(p), C.byref(h))
        self._h = h
        if rc != 0:
            msg = self._L.qtp_last_error(h).decode() if h else "qtp_create"
            self.close()
            raise QtpError(rc, msg)
        d = QtpDerived()
        self._chk(self._L.qtp_get_derived(h, C.byref(d)))
        self.derived = {f: (list(getattr(d, f)) if f == "boxl" else getattr(d, f))
                        for f, _ in QtpDerived._fields_}
        self.n_mols, self.n_atoms, self.n_dof = d.n_mols, d.n_atoms, d.n_dof
        self.n_beads, self.list_size = d.n_beads, d.list_size

    # ------------------------------------------- argument arrays and sizes
    # Every array the library reads or writes is checked for its length
    # first: the C calls take bare pointers.
    def _q(self, a, name="q"):
        """A whole P * n_dof array (positions, momenta, forces)."""
        return ptr(_f64(a, self.n_beads * self.n_dof, name))

    def _opt_q(self, a, name):
        return self._q(a, name) if a is not None else None

    def _charge(self, a):
        return ptr(_f64(a, self.n_atoms, "charge")) if a is not None else None

    def _mass(self, a):
        return ptr(_f64(a, self.n_dof, "mass")) if a is not None else None

    def _point(self, a):
        return ptr(_i32(a, self.n_beads * self.n_mols, "point"), ip)

    def _lst(self, a):
        return ptr(_i32(a, self.n_beads * self.list_size, "list"), ip)

    def _disp(self, a):
        return ptr(_f64(a, self.n_beads * 3 * self.n_mols, "disp"))

    def _dev(self, a):
        return ptr(_f64(a, 3 + (self.n_beads - 1) * self.n_dof, "dev")) if a is not None else None

    # ---------------------------------------------------------------- misc
    def _chk(self, rc):
        if rc != 0:
            msg = self._L.qtp_last_error(self._h).decode()
            if rc == -3:
                raise ListOverflow(rc, msg)
            raise QtpError(rc, msg)

    def close(self):
        if getattr(self, "_h", None):
            self._L.qtp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def stream(self) -> int:
        s = C.c_void_p()
        self._chk(self._L.qtp_get_stream(self._h, C.byref(s)))
        return s.value or 0

    def new_lists(self):
        return (np.zeros(self.n_beads * self.n_mols, np.int32),
                np.zeros(self.n_beads * self.list_size, np.int32),
                np.zeros(self.n_beads * 3 * self.n_mols))

    # ------------------------------------------------ reference-shaped calls
    def get_forces_nospline(self, q, F, charge, point, lst, bead=0) -> float:
        e = C.c_double()
        self._chk(self._L.qtp_get_forces_nospline(self._h, self._q(q), self._q(F, "F"),
                                                  self._charge(charge),
                                                  self._point(point), self._lst(lst),
                                                  int(bead), C.byref(e)))
        return e.value

    def get_forces_virial(self, q, F, charge, point, lst, bead=0):
        """get_forces_nospline plus the bead's virial tensor W = -dV/d(strain)
        at fixed Ewald parameters.  Returns (energy, W as a 3x3 array)."""
        e, v = C.c_double(), np.zeros(6)
        self._chk(self._L.qtp_get_forces_virial(self._h, self._q(q), self._q(F, "F"),
                                                self._charge(charge),
                                                self._point(point), self._lst(lst),
                                                int(bead), C.byref(e), ptr(v)))
        return e.value, voigt_to_matrix(v)

    def force_eval_sc_nospline(self, q, F_sc, charge, mass, point, lst, fphys=None):
        """Returns (tot_energy, sc_pot) as the reference's return value and
        sc_pot reference argument."""
        e, sc = C.c_double(), C.c_double()
        self._chk(self._L.qtp_force_eval_sc_nospline(
            self._h, self._q(q), self._q(F_sc, "F_sc"),
            self._charge(charge),
            self._mass(mass),
            self._point(point), self._lst(lst), C.byref(sc),
            self._opt_q(fphys, "fphys"), C.byref(e)))
        return e.value, sc.value

    # ---------------------------------------------- splined pair potential
    def set_splines(self, oo, oh, hh, start=0.0, step=0.01):
        """Load the fast-CMD spline tables (values at start + j*step) for the
        *_wspline calls and the resident mode (set_state must follow)."""
        t = [np.ascontiguousarray(x, dtype=np.float64) for x in (oo, oh, hh)]
        self._chk(self._L.qtp_set_splines(self._h, ptr(t[0]), t[0].size, ptr(t[1]), t[1].size,
                                          ptr(t[2]), t[2].size, float(start), float(step)))

    def clear_splines(self):
        self._chk(self._L.qtp_clear_splines(self._h))

    def get_forces_wspline(self, q, F, charge, point, lst, bead=0) -> float:
        e = C.c_double()
        self._chk(self._L.qtp_get_forces_wspline(self._h, self._q(q), self._q(F, "F"),
                                                 self._charge(charge),
                                                 self._point(point), self._lst(lst),
                                                 int(bead), C.byref(e)))
        return e.value

    def force_eval_sc_wspline(self, q, F_sc, charge, mass, point, lst, fphys=None):
        e, sc = C.c_double(), C.c_double()
        self._chk(self._L.qtp_force_eval_sc_wspline(
            self._h, self._q(q), self._q(F_sc, "F_sc"),
            self._charge(charge),
            self._mass(mass),
            self._point(point), self._lst(lst), C.byref(sc),
            self._opt_q(fphys, "fphys"), C.byref(e)))
        return e.value, sc.value

    def update_sc_wspline(self, q, p, mass, F, fphys, charge, point, lst, disp, dev, is_nvt, heat):
        h, sc, e = C.c_double(heat), C.c_double(), C.c_double()
        self._chk(self._L.qtp_update_sc_wspline(
            self._h, self._q(q), self._q(p, "p"),
            self._mass(mass), self._q(F, "F"),
            self._opt_q(fphys, "fphys"),
            self._charge(charge),
            self._point(point), self._lst(lst), self._disp(disp),
            self._dev(dev), 1 if is_nvt else 0,
            C.byref(h), C.byref(sc), C.byref(e)))
        return e.value, h.value, sc.value

    def update_trajectory_class_wspline(self, q, p, mass, F, charge, point, lst, disp) -> float:
        e = C.c_double()
        self._chk(self._L.qtp_update_trajectory_class_wspline(
            self._h, self._q(q), self._q(p, "p"),
            self._mass(mass), self._q(F, "F"),
            self._charge(charge),
            self._point(point), self._lst(lst), self._disp(disp), C.byref(e)))
        return e.value

    def neighbour_list(self, q, point, lst, disp, bead=0):
        self._chk(self._L.qtp_neighbour_list(self._h, self._q(q), self._point(point),
                                             self._lst(lst),
                                             self._disp(disp) if disp is not None else None,
                                             int(bead)))

    def ftransform(self, a, length, is_forwards):
        self._chk(self._L.qtp_ftransform(self._h, ptr(_f64(a, self.n_beads * int(length), "a")),
                                         int(length), 1 if is_forwards else 0))

    def pile_g(self, p_nm, mass, dev) -> float:
        h = C.c_double()
        self._chk(self._L.qtp_pile_g(self._h, self._q(p_nm, "p_nm"),
                                     self._mass(mass),
                                     self._dev(dev), C.byref(h)))
        return h.value

    def update_sc_nospline(self, q, p, mass, F, fphys, charge, point, lst, disp, dev, is_nvt,
                           heat):
        """One SC step.  Returns (energy, heat, sc_pot)."""
        h, sc, e = C.c_double(heat), C.c_double(), C.c_double()
        self._chk(self._L.qtp_update_sc_nospline(
            self._h, self._q(q), self._q(p, "p"),
            self._mass(mass), self._q(F, "F"),
            self._opt_q(fphys, "fphys"),
            self._charge(charge),
            self._point(point), self._lst(lst), self._disp(disp),
            self._dev(dev), 1 if is_nvt else 0,
            C.byref(h), C.byref(sc), C.byref(e)))
        return e.value, h.value, sc.value

    def update_trajectory_class_nospline(self, q, p, mass, F, charge, point, lst, disp) -> float:
        e = C.c_double()
        self._chk(self._L.qtp_update_trajectory_class_nospline(
            self._h, self._q(q), self._q(p, "p"),
            self._mass(mass), self._q(F, "F"),
            self._charge(charge),
            self._point(point), self._lst(lst), self._disp(disp), C.byref(e)))
        return e.value

    def get_kcv(self, q, fphys) -> float:
        v = C.c_double()
        self._chk(self._L.qtp_get_kcv(self._h, self._q(q), self._q(fphys, "fphys"), C.byref(v)))
        return v.value

    def get_conserved(self, q, p, mass, energy) -> float:
        v = C.c_double()
        self._chk(self._L.qtp_get_conserved(self._h, self._q(q), self._q(p, "p"),
                                            self._mass(mass),
                                            float(energy), C.byref(v)))
        return v.value

    # ------------------------------------------------------- resident mode
    def set_state(self, q, p, charge=None, mass=None):
        self._chk(self._L.qtp_set_state(self._h, self._q(q), self._q(p, "p"),
                                        self._charge(charge),
                                        self._mass(mass)))

    def get_state(self):
        n = self.n_beads * self.n_dof
        q, p, F, fph = (np.zeros(n) for _ in range(4))
        disp = np.zeros(self.n_beads * 3 * self.n_mols)
        self._chk(self._L.qtp_get_state(self._h, ptr(q), ptr(p), ptr(F), ptr(fph), ptr(disp)))
        return q, p, F, fph, disp

    def get_lists(self):
        point = np.zeros(self.n_beads * self.n_mols, np.int32)
        lst = np.zeros(self.n_beads * self.list_size, np.int32)
        self._chk(self._L.qtp_get_lists(self._h, ptr(point, ip), ptr(lst, ip)))
        return point, lst

    def virial(self):
        """Virial tensor (P x 3 x 3) and energy (P) of every resident bead."""
        v, e = np.zeros(6 * self.n_beads), np.zeros(self.n_beads)
        self._chk(self._L.qtp_get_virial(self._h, ptr(v), ptr(e)))
        return np.stack([voigt_to_matrix(v[6 * b:6 * b + 6]) for b in range(self.n_beads)]), e

    def save_checkpoint(self, path):
        self._chk(self._L.qtp_save_checkpoint(self._h, os.fsencode(path)))

    def load_checkpoint(self, path, charge=None, mass=None):
        self._chk(self._L.qtp_load_checkpoint(self._h, os.fsencode(path),
                                              self._charge(charge),
                                              self._mass(mass)))

    def set_seed(self, seed: int):
        self._chk(self._L.qtp_set_seed(self._h, C.c_uint64(seed)))

    def run(self, n_steps: int, is_nvt: bool = False):
        self._chk(self._L.qtp_run(self._h, int(n_steps), 1 if is_nvt else 0))

    def sync(self):
        self._chk(self._L.qtp_sync(self._h))

    def observables(self) -> dict:
        """K_cv, conserved quantity, SC potential and sc_pot of the resident
        state, computed on the device."""
        o = np.zeros(4)
        self._chk(self._L.qtp_get_observables(self._h, ptr(o)))
        return {"kcv": o[0], "conserved": o[1], "energy": o[2], "sc_pot": o[3]}

    def run_observed(self, n_steps: int, is_nvt: bool = False, every: int = 1):
        """Resident steps with observables recorded every `every` steps on the
        device; returns an (n_samples, 4) array (kcv, conserved, energy, sc_pot)."""
        tr = np.zeros(4 * (int(n_steps) // int(every) + 1))
        n = C.c_int()
        self._chk(self._L.qtp_run_observed(self._h, int(n_steps), 1 if is_nvt else 0, int(every),
                                           ptr(tr), C.byref(n)))
        return tr[:4 * n.value].reshape(-1, 4)

    def scalars(self) -> dict:
        o = np.zeros(6)
        self._chk(self._L.qtp_get_scalars(self._h, ptr(o)))
        return {"energy": o[0], "sc_pot": o[1], "heat": o[2], "rebuilds": int(o[3]),
                "steps": int(o[4]), "status": int(o[5])}

    # ------------------------------ launch configuration and autotuner
    TUNING_KEYS = ("lanes", "matrix_kspace", "rs_registers", "kforce_16_waves",
                   "ksum_blocks_per_chunk", "fused_propagation", "rs_waves", "rs_split",
                   "ksum_chunks_per_block")

    def tuning(self) -> dict:
        out = np.zeros(len(self.TUNING_KEYS), np.int32)
        self._chk(self._L.qtp_get_tuning(self._h, ptr(out, ip)))
        return dict(zip(self.TUNING_KEYS, (int(v) for v in out)))

    def set_tuning(self, **kw):
        """Set launch choices (keys of tuning()); forces of a set state are
        recomputed."""
        cur = self.tuning()
        cur.update(kw)
        arr = np.array([cur[k] for k in self.TUNING_KEYS], np.int32)
        self._chk(self._L.qtp_set_tuning(self._h, ptr(arr, ip)))

    def autotune(self, retime: bool = False) -> dict:
        """Time every candidate launch configuration on the resident state's
        force evaluation and keep the fastest (cached per system shape)."""
        r = np.zeros(11)
        self._chk(self._L.qtp_autotune(self._h, 1 if retime else 0, ptr(r)))
        chosen = dict(zip(self.TUNING_KEYS[:5], (int(v) for v in r[2:7])))
        chosen["rs_waves"] = int(r[8])
        chosen["rs_split"] = int(r[9])
        chosen["ksum_chunks_per_block"] = int(r[10])
        return {"best_ms": float(r[0]), "default_ms": float(r[1]), "chosen": chosen,
                "timed": int(r[7])}

    # ------------------------------ Verlet-list build
    def set_lists(self, method: str = "auto", rebuild_every: int = 0):
        """How lists are built: "auto" (cell-binned above 1024 molecules when
        the box holds the grid), "allpairs" (neighbour.cpp's loop) or "cells";
        rebuild_every > 0 also rebuilds every bead's lists on each such step
        of the resident runs.  The lists are identical whichever builds them."""
        m = {"auto": 0, "allpairs": 1, "cells": 2}[method]
        self._chk(self._L.qtp_set_lists(self._h, m, int(rebuild_every)))

    def list_info(self) -> dict:
        """{"cells": in use, "grid": cells per edge, "S": stencil half-width}."""
        out = np.zeros(5, np.int32)
        self._chk(self._L.qtp_get_list_info(self._h, ptr(out, ip)))
        return {"cells": bool(out[0]), "grid": tuple(int(v) for v in out[1:4]), "S": int(out[4])}

    # ------------------------------ centroid dynamics, thermostats, g(r)
    def set_thermostat(self, kind: str = "pile_g", tau_fs: float = 0.0):
        """Centroid thermostat of NVT resident steps: "pile_g" (Bussi rescale,
        time constant tau_fs; 0 = the reference's pile_g) or "pile_l"
        (Langevin, friction 1/tau_fs).  Classical contexts: "pile_l" only."""
        k = {"pile_g": 0, "pile_l": 1}[kind]
        self._chk(self._L.qtp_set_thermostat(self._h, k, float(tau_fs)))

    def set_pimd_mode(self, mode: str = "pimd", omega_ad: float = 0.0):
        """"pimd", "pacmd" (partially adiabatic CMD at omega_ad, a.u.; 0 =
        P^(P/(P-1))/(beta hbar)) or "centroid_fixed"."""
        m = {"pimd": 0, "pacmd": 1, "centroid_fixed": 2}[mode]
        self._chk(self._L.qtp_set_pimd_mode(self._h, m, float(omega_ad)))

    def centroid_force(self, on: bool = True):
        """Start (zeroing the sum) or stop accumulating the force on the
        centroid, (1/P) sum_b F_sc,b, after every resident step."""
        self._chk(self._L.qtp_centroid_force(self._h, 1 if on else 0))

    def get_centroid_force(self):
        """(mean force since centroid_force(True), centroid positions,
        number of steps); arrays in reference layout (n_dof)."""
        f, c, n = np.zeros(self.n_dof), np.zeros(self.n_dof), C.c_int()
        self._chk(self._L.qtp_get_centroid_force(self._h, ptr(f), ptr(c), C.byref(n)))
        return f, c, n.value

    def rdf_reset(self, r_max_bohr: float, n_bins: int):
        self._chk(self._L.qtp_rdf_reset(self._h, float(r_max_bohr), int(n_bins)))
        self._rdf_bins = int(n_bins)

    def rdf_sample(self):
        self._chk(self._L.qtp_rdf_sample(self._h))

    def rdf_get(self):
        """(O-O counts, O-H counts, number of samples)."""
        n = self._rdf_bins
        oo, oh, k = np.zeros(n), np.zeros(n), C.c_int()
        self._chk(self._L.qtp_rdf_get(self._h, ptr(oo), ptr(oh), C.byref(k)))
        return oo, oh, k.value

    def set_nn(self, model):
        """Attach a Surrogate (or None) to the classical resident steps."""
        self._chk(self._L.qtp_set_nn(self._h, model._h if model is not None else None))
        self._nn = model

    def enable_timing(self, on=True):
        self._chk(self._L.qtp_enable_timing(self._h, 1 if on else 0))

    def timing(self, kernel: str):
        ms, n = C.c_double(), C.c_int()
        self._chk(self._L.qtp_get_timing(self._h, kernel.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value


class Batch:
    """Independent replicas (Water objects with resident state) advanced
    together (qtp_batch_*): every replica's step graph is queued on its own
    streams so the GPU overlaps them.  The Water objects must outlive the
    batch; call sync() before touching them again."""

    def __init__(self, waters, is_nvt=None):
        self._L = _lib.load()
        self.waters = list(waters)
        n = len(self.waters)
        arr = (C.c_void_p * n)(*[w._h for w in self.waters])
        nvt = None
        if is_nvt is not None:
            flags = [int(bool(is_nvt))] * n if np.isscalar(is_nvt) else [int(bool(x)) for x in is_nvt]
            nvt = np.ascontiguousarray(flags, dtype=np.int32)
        h = C.c_void_p()
        rc = self._L.qtp_batch_create(arr, ptr(nvt, ip) if nvt is not None else None, n, C.byref(h))
        self._h = h
        if rc != 0:
            msg = self._L.qtp_batch_last_error(h).decode() if h else "qtp_batch_create"
            self.close()
            raise QtpError(rc, msg)

    def _chk(self, rc):
        if rc != 0:
            raise QtpError(rc, self._L.qtp_batch_last_error(self._h).decode())

    def run(self, n_steps: int):
        self._chk(self._L.qtp_batch_run(self._h, int(n_steps)))

    def sync(self):
        self._chk(self._L.qtp_batch_sync(self._h))

    def close(self):
        if getattr(self, "_h", None):
            self._L.qtp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


NN_PARAMS_PER_ELEMENT = 128 * 40 + 128 + 128 * 128 + 128 + 128 + 1


def nn_descriptor_params() -> dict:
    """The descriptor constants compiled into the library (Angstrom)."""
    L = _lib.load()
    n = L.qtp_nn_descriptor_params(None, 0)
    v = np.zeros(n)
    L.qtp_nn_descriptor_params(ptr(v), n)
    return dict(r_rad=v[0], r_ang=v[1], mu0=v[2], dmu=v[3], eta_r=v[4], zeta=v[5:21].copy(),
                lam=v[21:37].copy(), eta_a=v[37:53].copy(), w=v[53:57].reshape(2, 2).copy())


class Surrogate:
    """The learned centroid-force correction on the GPU (qtp_nn_*): an
    atom-centred BF16 network whose energies add to qTIP4P/F's, trained by
    force matching on centroid mean forces."""

    def __init__(self, n_mols, boxl, max_frames=16):
        self._L = _lib.load()
        h = C.c_void_p()
        b = np.ascontiguousarray(boxl, dtype=np.float64)
        rc = self._L.qtp_nn_create(int(n_mols), ptr(b), int(max_frames), C.byref(h))
        self._h = h
        if rc != 0:
            msg = self._L.qtp_nn_last_error(h).decode() if h else "qtp_nn_create"
            self.close()
            raise QtpError(rc, msg)
        self.n_mols, self.n_atoms, self.n_dof = int(n_mols), 3 * int(n_mols), 9 * int(n_mols)
        self.max_frames = int(max_frames)

    def _chk(self, rc):
        if rc != 0:
            raise QtpError(rc, self._L.qtp_nn_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._L.qtp_nn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_model(self, theta, mean, std, e_scale):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        m = np.ascontiguousarray(mean, dtype=np.float32).ravel()
        s = np.ascontiguousarray(std, dtype=np.float32).ravel()
        if t.size != 2 * NN_PARAMS_PER_ELEMENT or m.size != 80 or s.size != 80:
            raise ValueError("theta: 2 x 21889 values; mean, std: 2 x 40")
        self._chk(self._L.qtp_nn_set_model(self._h, ptr(t, _lib.fp), ptr(m, _lib.fp),
                                           ptr(s, _lib.fp), float(e_scale)))

    def get_model(self):
        t = np.zeros(2 * NN_PARAMS_PER_ELEMENT, np.float32)
        self._chk(self._L.qtp_nn_get_model(self._h, ptr(t, _lib.fp)))
        return t

    def eval(self, pos, descriptors=False):
        """(energies (n_frames), forces (n_frames, n_dof)[, raw descriptors
        (n_frames, n_atoms, 40)]) of frames in reference layout."""
        x = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, self.n_dof)
        nf = x.shape[0]
        e, f = np.zeros(nf), np.zeros((nf, self.n_dof))
        g = np.zeros((nf, self.n_atoms, 40), np.float32) if descriptors else None
        self._chk(self._L.qtp_nn_eval(self._h, nf, ptr(x), ptr(e), ptr(f),
                                      ptr(g, _lib.fp) if g is not None else None))
        return (e, f, g) if descriptors else (e, f)

    def set_data(self, pos, target):
        x = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, self.n_dof)
        t = np.ascontiguousarray(target, dtype=np.float64).reshape(-1, self.n_dof)
        if x.shape != t.shape:
            raise ValueError("positions and targets differ in shape")
        self._chk(self._L.qtp_nn_set_data(self._h, x.shape[0], ptr(x), ptr(t)))
        self.n_data = x.shape[0]

    def loss_grad(self, frame0, n_frames):
        loss = C.c_double()
        g = np.zeros(2 * NN_PARAMS_PER_ELEMENT, np.float32)
        self._chk(self._L.qtp_nn_loss_grad(self._h, int(frame0), int(n_frames), C.byref(loss),
                                           ptr(g, _lib.fp)))
        return loss.value, g

    def train(self, n_steps, batch_frames, lr, seed=0):
        tr = np.zeros(max(1, int(n_steps)))
        self._chk(self._L.qtp_nn_train(self._h, int(n_steps), int(batch_frames), float(lr),
                                       C.c_uint64(seed), ptr(tr)))
        return tr[:int(n_steps)]
