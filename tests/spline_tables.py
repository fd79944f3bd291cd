"""Synthetic spline files for the fast-CMD path tests.  The reference ships
no spline files (simulation.cpp:36-116 reads oo_spline.txt, oh_spline.txt,
hh_spline.txt from the working directory), so these are smooth pair
corrections of a plausible size tabulated like them: start 0, step 0.01 bohr."""
import numpy as np

START, STEP, N = 0.0, 0.01, 901   # range end 9.0 bohr


def tables():
    r = START + STEP * np.arange(N)
    oo = 2e-3 * np.exp(-(r - 5.3) ** 2 / 0.8) - 1e-3 * np.exp(-(r - 7.0) ** 2)
    oh = -1.5e-3 * np.exp(-(r - 3.6) ** 2 / 0.5)
    hh = 8e-4 * np.exp(-(r - 4.4) ** 2 / 0.7)
    return oo, oh, hh
