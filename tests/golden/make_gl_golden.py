"""Writes GL-trajectory fixtures computed by the FP64 oracle (oracle/liboracle.so)
on the seeded workloads of pfdd_amd.configs:

  gl_c0.npz   c0 (32^3, NS=3, 500 GL steps): final Re xi (all points), max |Im xi|,
              gamma and gamma_a for every step.
  gl_c1.npz   c1 (128^3, NS=12, 1000 GL steps): final Re xi on the planes
              i = 0, 37, 64, 101 and the slabs j = 5, k = 77, per-field sums and
              sums of squares of Re xi, gamma and gamma_a for every step.

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
        np.savez_compressed(os.path.join(HERE, "gl_c0.npz"), re=xi.real, im_max=np.abs(xi.imag).max(),
                            gamma=g, gamma_a=ga, steps=w.steps, seconds=dt)
        print("c0", dt, "s")
    if "c1" in which:
        w = C.c1()
        xi, g, ga, dt = run(w, w.steps)
        re = xi.real
        np.savez_compressed(os.path.join(HERE, "gl_c1.npz"), planes_i=re[:, C1_PLANES_I], plane_j5=re[:, :, 5, :],
                            plane_k77=re[:, :, :, 77], i_idx=np.array(C1_PLANES_I),
                            sums=re.reshape(w.ns, -1).sum(1), sumsq=(re.reshape(w.ns, -1) ** 2).sum(1),
                            im_max=np.abs(xi.imag).max(), gamma=g, gamma_a=ga, steps=w.steps, seconds=dt)
        print("c1", dt, "s")


if __name__ == "__main__":
    main(sys.argv[1:] or ["c0"])
