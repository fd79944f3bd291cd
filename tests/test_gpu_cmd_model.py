"""The trained correction (qtip4pf_water_amd/data/cmd_surrogate_216x32.npz,
made by `python -m qtip4pf_water_amd.cmd build` on the GPU) on held-out
centroid mean forces of an independent 216 H2O x 32-bead trajectory
(tests/golden/cmd_val_216x32.npz): what the network gains over classical
qTIP4P/F forces at the centroid, against the sampling noise of the targets
themselves.  No counterpart in the reference."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def test_trained_correction_on_held_out_mean_forces():
    from qtip4pf_water_amd import Surrogate, cmd
    say("trained correction, 20 held-out frames")
    m = cmd.load_model()
    z = np.load(os.path.join(ROOT, "tests", "golden", "cmd_val_216x32.npz"))
    pos, tgt, noise = z["pos"], z["target"].astype(np.float64), z["noise"].astype(np.float64)
    with Surrogate(216, z["box"], max_frames=8) as g:
        g.set_model(m["theta"], m["mean"], m["std"], m["e_scale"])
        _, F = g.eval(pos)
    without = np.sqrt(np.mean(tgt ** 2))
    with_ = np.sqrt(np.mean((tgt - F) ** 2))
    floor = np.sqrt(np.mean(noise ** 2))
    say(f"rms centroid-force error: without {without:.3e}, with {with_:.3e}, noise {floor:.3e}")
    # the classical force misses the centroid mean force by `without`; the
    # correction leaves less than 40 % of it, within 30 % of the targets' own noise
    assert with_ < 0.4 * without
    assert with_ < 1.3 * floor
