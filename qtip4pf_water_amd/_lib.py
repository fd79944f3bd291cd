"""ctypes loader for libqtip4pf.so (the gfx950 HIP library).

Fails loudly: there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libqtip4pf.so")

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
fp = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)

QTP_OK = 0
STATUS = {-1: "QTP_ERR_ARG", -2: "QTP_ERR_HIP", -3: "QTP_ERR_LIST_OVERFLOW",
          -4: "QTP_ERR_NONFINITE", -5: "QTP_ERR_STATE"}


class QtpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class ListOverflow(QtpError):
    pass


class QtpParams(C.Structure):
    _fields_ = [("n_mols", C.c_int), ("n_beads", C.c_int), ("boxl", C.c_double * 3),
                ("T_K", C.c_double), ("dtfs", C.c_double), ("eps_ewald", C.c_double),
                ("epsilon_sc", C.c_double), ("list_size", C.c_int),
                ("ewald_kappa", C.c_double), ("ewald_rcut", C.c_double),
                ("ewald_kcut", C.c_double), ("ewald_kmax", C.c_int),
                ("lj_rcut", C.c_double)]


class QtpDerived(C.Structure):
    _fields_ = [("n_mols", C.c_int), ("n_atoms", C.c_int), ("n_dof", C.c_int),
                ("n_beads", C.c_int), ("list_size", C.c_int),
                ("boxl", C.c_double * 3), ("vol", C.c_double), ("min_box", C.c_double),
                ("max_box", C.c_double),
                ("beta", C.c_double), ("beta_n", C.c_double), ("piovern", C.c_double),
                ("omega_p", C.c_double), ("omega_psq", C.c_double), ("dt", C.c_double),
                ("halfdt", C.c_double), ("epsilonsc", C.c_double),
                ("acc", C.c_double), ("rho", C.c_double), ("kwald_prefactor", C.c_double),
                ("sd_ox", C.c_double), ("sd_h", C.c_double),
                ("kappa", C.c_double), ("rcut", C.c_double), ("k_cut", C.c_double),
                ("k_max", C.c_int), ("n_kgroups", C.c_int),
                ("lj_rcut", C.c_double), ("r_skin", C.c_double)]


# (name, restype, argtypes) for every entry point of include/qtip4pf.h
_vp = C.c_void_p
SIGNATURES = [
    ("qtp_create", C.c_int, [C.POINTER(QtpParams), C.POINTER(_vp)]),
    ("qtp_destroy", None, [_vp]),
    ("qtp_last_error", C.c_char_p, [_vp]),
    ("qtp_get_derived", C.c_int, [_vp, C.POINTER(QtpDerived)]),
    ("qtp_get_stream", C.c_int, [_vp, C.POINTER(_vp)]),
    ("qtp_get_forces_nospline", C.c_int, [_vp, dp, dp, dp, ip, ip, C.c_int, dp]),
    ("qtp_get_forces_virial", C.c_int, [_vp, dp, dp, dp, ip, ip, C.c_int, dp, dp]),
    ("qtp_force_eval_sc_nospline", C.c_int, [_vp, dp, dp, dp, dp, ip, ip, dp, dp, dp]),
    ("qtp_neighbour_list", C.c_int, [_vp, dp, ip, ip, dp, C.c_int]),
    ("qtp_ftransform", C.c_int, [_vp, dp, C.c_int, C.c_int]),
    ("qtp_pile_g", C.c_int, [_vp, dp, dp, dp, dp]),
    ("qtp_update_sc_nospline", C.c_int,
     [_vp, dp, dp, dp, dp, dp, dp, ip, ip, dp, dp, C.c_int, dp, dp, dp]),
    ("qtp_update_trajectory_class_nospline", C.c_int, [_vp, dp, dp, dp, dp, dp, ip, ip, dp, dp]),
    ("qtp_get_kcv", C.c_int, [_vp, dp, dp, dp]),
    ("qtp_get_conserved", C.c_int, [_vp, dp, dp, dp, C.c_double, dp]),
    ("qtp_get_kcv_noalloc", C.c_int, [C.c_int, C.c_int, C.c_double, _vp, _vp, _vp, _vp]),
    ("qtp_set_splines", C.c_int, [_vp, dp, C.c_int, dp, C.c_int, dp, C.c_int, C.c_double,
                                  C.c_double]),
    ("qtp_clear_splines", C.c_int, [_vp]),
    ("qtp_get_forces_wspline", C.c_int, [_vp, dp, dp, dp, ip, ip, C.c_int, dp]),
    ("qtp_force_eval_sc_wspline", C.c_int, [_vp, dp, dp, dp, dp, ip, ip, dp, dp, dp]),
    ("qtp_update_sc_wspline", C.c_int,
     [_vp, dp, dp, dp, dp, dp, dp, ip, ip, dp, dp, C.c_int, dp, dp, dp]),
    ("qtp_update_trajectory_class_wspline", C.c_int, [_vp, dp, dp, dp, dp, dp, ip, ip, dp, dp]),
    ("qtp_host_rng_create", C.c_int, [C.c_uint32, C.POINTER(_vp)]),
    ("qtp_host_rng_destroy", None, [_vp]),
    ("qtp_host_rng_draw_pile", C.c_int, [_vp, C.c_int, C.c_int, dp]),
    ("qtp_host_rng_draw_momenta", C.c_int, [_vp, C.c_int, C.c_int, C.c_double, dp]),
    ("qtp_initialise_fcc", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, dp, dp]),
    ("qtp_initialise_orth", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                      C.c_long, dp, dp, C.POINTER(C.c_long)]),
    ("qtp_read_positions", C.c_int, [C.c_char_p, C.c_int, C.c_int, dp, C.POINTER(C.c_int)]),
    ("qtp_write_positions", C.c_int, [C.c_char_p, C.c_int, C.c_int, dp]),
    ("qtp_set_state", C.c_int, [_vp, dp, dp, dp, dp]),
    ("qtp_get_state", C.c_int, [_vp, dp, dp, dp, dp, dp]),
    ("qtp_get_lists", C.c_int, [_vp, ip, ip]),
    ("qtp_save_checkpoint", C.c_int, [_vp, C.c_char_p]),
    ("qtp_load_checkpoint", C.c_int, [_vp, C.c_char_p, dp, dp]),
    ("qtp_get_virial", C.c_int, [_vp, dp, dp]),
    ("qtp_set_seed", C.c_int, [_vp, C.c_uint64]),
    ("qtp_run", C.c_int, [_vp, C.c_int, C.c_int]),
    ("qtp_sync", C.c_int, [_vp]),
    ("qtp_get_observables", C.c_int, [_vp, dp]),
    ("qtp_run_observed", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, dp, ip]),
    ("qtp_batch_create", C.c_int, [C.POINTER(_vp), ip, C.c_int, C.POINTER(_vp)]),
    ("qtp_batch_run", C.c_int, [_vp, C.c_int]),
    ("qtp_batch_sync", C.c_int, [_vp]),
    ("qtp_batch_last_error", C.c_char_p, [_vp]),
    ("qtp_batch_destroy", None, [_vp]),
    ("qtp_get_scalars", C.c_int, [_vp, dp]),
    ("qtp_get_tuning", C.c_int, [_vp, ip]),
    ("qtp_set_tuning", C.c_int, [_vp, ip]),
    ("qtp_autotune", C.c_int, [_vp, C.c_int, dp]),
    ("qtp_set_lists", C.c_int, [_vp, C.c_int, C.c_int]),
    ("qtp_get_list_info", C.c_int, [_vp, ip]),
    ("qtp_set_thermostat", C.c_int, [_vp, C.c_int, C.c_double]),
    ("qtp_set_pimd_mode", C.c_int, [_vp, C.c_int, C.c_double]),
    ("qtp_centroid_force", C.c_int, [_vp, C.c_int]),
    ("qtp_get_centroid_force", C.c_int, [_vp, dp, dp, ip]),
    ("qtp_rdf_reset", C.c_int, [_vp, C.c_double, C.c_int]),
    ("qtp_rdf_sample", C.c_int, [_vp]),
    ("qtp_rdf_get", C.c_int, [_vp, dp, dp, ip]),
    ("qtp_nn_descriptor_params", C.c_int, [dp, C.c_int]),
    ("qtp_nn_create", C.c_int, [C.c_int, dp, C.c_int, C.POINTER(_vp)]),
    ("qtp_nn_destroy", None, [_vp]),
    ("qtp_nn_last_error", C.c_char_p, [_vp]),
    ("qtp_nn_set_model", C.c_int, [_vp, fp, fp, fp, C.c_double]),
    ("qtp_nn_get_model", C.c_int, [_vp, fp]),
    ("qtp_nn_eval", C.c_int, [_vp, C.c_int, dp, dp, dp, fp]),
    ("qtp_nn_set_data", C.c_int, [_vp, C.c_int, dp, dp]),
    ("qtp_nn_loss_grad", C.c_int, [_vp, C.c_int, C.c_int, dp, fp]),
    ("qtp_nn_train", C.c_int, [_vp, C.c_int, C.c_int, C.c_double, C.c_uint64, dp]),
    ("qtp_set_nn", C.c_int, [_vp, _vp]),
    ("qtp_enable_timing", C.c_int, [_vp, C.c_int]),
    ("qtp_get_timing", C.c_int, [_vp, C.c_char_p, dp, ip]),
    ("qtp_philox4x32_10", C.c_int, [u32p, u32p, u32p]),
]

_lib = None


def load():
    """Load the in-tree library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build it with "
                               "`python -m qtip4pf_water_amd.build` (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a, t=dp):
    if a is None:
        return None
    return a.ctypes.data_as(t)
