"""Energy tolerances of the GPU parity tests.

Single evaluations: the oracle sums the bead energy with an 80-bit
accumulator (oracle.exact_energy), and the GPU energy is held to 1e-12 of
|V| against it.  The reference's own FP64 running sum is not that accurate
on large boxes (at 4096 waters it is off by ~1e-11 relative, the rounding of
millions of real-space pair terms), so it is compared at that scale
separately.

Trajectories: after k chaotic steps the GPU and oracle positions differ by
dq (measured); V then differs by -F.dq to first order, so the bound is
1e-12 |V| + 2 sum |F_i| |dq_i|, with F and dq taken after the step."""
import numpy as np


def single(e_gpu, e_exact, rel=1e-12):
    return abs(e_gpu - e_exact) <= rel * abs(e_exact)


def trajectory_bound(e_exact, F, dq, rel=1e-12):
    return rel * abs(e_exact) + 2.0 * float(np.sum(np.abs(F) * np.abs(dq)))
