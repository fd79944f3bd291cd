"""Write golden vectors from the reference's own compiled routines.

oracle/_ref/libqtp_ref.so is built by oracle/Makefile from the reference's
neighbour.cpp and thermostat.cpp (unmodified) plus our harness.  This script
runs them on seeded inputs and stores inputs + outputs as .npz fixtures:

  nlist_216.npz   neighbour_list (neighbour.cpp:12-62), 216 waters, 2 beads
  pile_g_27x8.npz pile_g (thermostat.cpp:64-133) with std::mt19937(seed),
                  27 waters, 8 beads, two consecutive calls

Run in the container that holds the reference tree:
    make -C oracle && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from qtip4pf_water_amd import config as cfg  # noqa: E402


def nlist():
    P = 2
    q0, box = cfg.lattice_water_box(6, 18.64, seed=2024, jitter_bohr=0.4)
    q = cfg.replicate_beads(q0, P) + np.random.default_rng(7).normal(0, 0.2, P * q0.size)
    o = oracle.Oracle(216, P, box)
    oracle.ref_set_from(o)
    point, lst, disp = o.new_lists()
    for b in range(P):
        oracle.ref().ref_neighbour_list(oracle._ptr(q), oracle._ptr(point, oracle._ip),
                                        oracle._ptr(lst, oracle._ip), oracle._ptr(disp), b)
    np.savez_compressed(os.path.join(HERE, "nlist_216.npz"), q=q, boxl=np.array(box),
                        n_beads=P, list_size=o.list_size, point=point, list=lst)


def pile():
    P, seed = 8, 4242
    q0, box = cfg.lattice_water_box(3, 9.4, seed=5)
    o = oracle.Oracle(27, P, box)
    oracle.ref_set_from(o)
    p_in = np.random.default_rng(11).normal(0, 25.0, P * o.n_dof)
    R = oracle.ref()
    R.ref_seed(seed)
    p1 = p_in.copy()
    h1 = R.ref_pile_g(oracle._ptr(p1), oracle._ptr(o.mass))
    p2 = p1.copy()
    h2 = R.ref_pile_g(oracle._ptr(p2), oracle._ptr(o.mass))
    np.savez_compressed(os.path.join(HERE, "pile_g_27x8.npz"), boxl=np.array(box), n_beads=P,
                        seed=seed, mass=o.mass, p_in=p_in, p_out1=p1, heat1=h1, p_out2=p2,
                        heat2=h2, T_K=298.0, dtfs=0.25)


if __name__ == "__main__":
    if not oracle.ref_available():
        oracle.build_ref()
    nlist()
    pile()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")
