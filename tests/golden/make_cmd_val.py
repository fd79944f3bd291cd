"""Writes tests/golden/cmd_val_216x32.npz: 20 of the validation frames of a
`python -m qtip4pf_water_amd.cmd build --out DIR` run (centroid positions of
a 216 H2O x 32-bead PILE-L PIMD trajectory, and for each the centroid mean
force over 1000 centroid-fixed steps minus the classical force: the target
of the learned correction; and the first-half mean minus the full mean: its
sampling noise).  These are the GPU pipeline's own outputs, not the
reference's (the reference has no CMD).

    python tests/golden/make_cmd_val.py DIR/centroid_val.npz
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(path):
    z = np.load(path)
    sel = np.arange(0, len(z["pos"]), 5)[:20]
    np.savez_compressed(os.path.join(HERE, "cmd_val_216x32.npz"),
                        pos=z["pos"][sel].astype(np.float64),
                        target=(z["fmean"][sel].astype(np.float64)
                                - z["fcls"][sel].astype(np.float64)).astype(np.float32),
                        noise=(z["fhalf"][sel].astype(np.float64)
                               - z["fmean"][sel].astype(np.float64)).astype(np.float32),
                        box=z["box"], n_acc=z["n_acc"])


if __name__ == "__main__":
    main(sys.argv[1])
