"""Energy tolerances of the GPU parity tests.

The oracle sums every bead energy with an 80-bit accumulator (the GPU tests
run inside oracle.exact_energy; tests/conftest.py), and the GPU energy is
held to 1e-12 of |V| against it, plus a floor of 16 machine epsilons of
the sum of the magnitudes of V's terms (|LJ| + |intra| + |real| + |recip|
+ |self|): V is a difference of Ewald terms up to a thousand times larger
than itself, and no FP64 evaluation of those terms, in any order, is
closer to the exact value than a few epsilons of their size.  The
reference's own FP64 running sum is not that accurate on large boxes
(at 4096 waters it is off by ~1e-11 relative), so it is compared at the
scale of the self energy where it is compared at all.

Trajectories: after k chaotic steps the GPU and oracle positions differ by
dq (measured); V then differs by -F.dq to first order, so the bound adds
2 sum |F_i| |dq_i|, with F and dq taken after the step."""
import numpy as np

EPS = float(np.finfo(np.float64).eps)


def term_scale(o, q, point, lst, bead=0):
    """Sum of the magnitudes of the bead energy's terms (oracle parts)."""
    F = np.zeros(o.n_beads * o.n_dof)
    _, parts = o.get_forces(q, F, point, lst, bead, parts=True)
    return float(np.sum(np.abs(parts[:5])))


def single(e_gpu, e_exact, rel=1e-12, scale=0.0):
    return abs(e_gpu - e_exact) <= rel * abs(e_exact) + 16.0 * EPS * scale


def trajectory_bound(e_exact, F, dq, rel=1e-12, scale=0.0):
    return (rel * abs(e_exact) + 16.0 * EPS * scale +
            2.0 * float(np.sum(np.abs(F) * np.abs(dq))))
