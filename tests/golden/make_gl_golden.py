"""Writes GL-trajectory fixtures computed by the FP64 oracle (oracle/liboracle.so)
on the seeded workloads of pfdd_amd.configs:

  gl_c0.npz   c0 (32^3, NS=3, 500 GL steps): final complex xi (all points),
              gamma and gamma_a for every step.
  gl_c1.npz   c1 (128^3, NS=12, 1000 GL steps): final complex xi at 40000
              seeded sample points on the planes i = 0, 37, 64, 101, j = 5 and
              k = 77 (c1_sample_points()), per-field sums and sums of squares of
              Re xi and Im xi, gamma and gamma_a for every step.

xi is complex on purpose: frec (G.c:2698-2741) puts the Nyquist index at -1/2
on every axis, so B(k) != B(-k) on planes where only some components are at
Nyquist, and the reference's C2C loop drives Im xi to O(1e-3).

Run from the repo root (c1 takes tens of minutes on 8 cores):
    python tests/golden/make_gl_golden.py c0 c1
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oraclepy as O  # noqa: E402
from pfdd_amd import configs as C  # noqa: E402

C1_PLANES_I = [0, 37, 64, 101]


def c1_sample_points(n=40000, N=128, ns=12, seed=7):
    """(field, i, j, k) of the c1 sample: a plane is drawn among i in C1_PLANES_I,
    j = 5 and k = 77, then the two free coordinates and the field."""
    rng = np.random.default_rng(seed)
    plane = rng.integers(0, 6, n)
    a, b = rng.integers(0, N, n), rng.integers(0, N, n)
    f = rng.integers(0, ns, n)
    i = np.where(plane < 4, np.array(C1_PLANES_I)[np.minimum(plane, 3)], a)
    j = np.where(plane == 4, 5, np.where(plane < 4, a, b))
    k = np.where(plane == 5, 77, b)
    return f, i, j, k


def run(w, steps):
    p = O.make_params(w.shape, w.eps, C.C11, C.C12, C.C44, [0.1] * w.ns, C.asf2_reference(w.ns), [1.0] * w.ns)
    o = O.Oracle(p)
    o.set_xi(w.xi)
    o.set_tau(w.tau)
    o.prologue(True)
    t = time.time()
    g, ga = o.gl_steps(steps)
    dt = time.time() - t
    return o.get_xi(), g, ga, dt


def main(which):
    if "c0" in which:
        w = C.c0()
        xi, g, ga, dt = run(w, w.steps)
        np.savez_compressed(os.path.join(HERE, "gl_c0.npz"), xi=xi, gamma=g, gamma_a=ga, steps=w.steps,
                            seconds=dt)
        print("c0", dt, "s")
    if "c1" in which:
        w = C.c1()
        xi, g, ga, dt = run(w, w.steps)
        flat = xi.reshape(w.ns, -1)
        f, i, j, k = c1_sample_points()
        np.savez_compressed(os.path.join(HERE, "gl_c1.npz"), sample=xi[f, i, j, k],
                            sums_re=flat.real.sum(1), sumsq_re=(flat.real ** 2).sum(1),
                            sums_im=flat.imag.sum(1), sumsq_im=(flat.imag ** 2).sum(1),
                            gamma=g, gamma_a=ga, steps=w.steps, seconds=dt)
        print("c1", dt, "s")


if __name__ == "__main__":
    main(sys.argv[1:] or ["c0"])
