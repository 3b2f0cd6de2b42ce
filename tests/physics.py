"""Shared test inputs: normalised isotropic material and small eigenstrain sets."""
import numpy as np

NU = 0.33
MU = 1.0
LAM = 2 * NU / (1 - 2 * NU)
C11, C12, C44 = LAM + 2 * MU, LAM, MU
ASF_CU = 0.0064659518668288441        # Asf2[0], Asf2[1] of G.c:4281-4282
ASF_DEFAULT = 0.0262                  # Asf2[is] of G.c:4279
PI_REF = 3.141592654                  # G.c:95


def asf2_reference(ns):
    """Asf2 as main() sets it (G.c:4278-4285)."""
    a = [ASF_DEFAULT] * ns
    for i in range(min(2, ns)):
        a[i] = ASF_CU
    if ns == 4:
        a[2] = a[3] = 0.013307048604153731
    return a


def theta_nz():
    """Rotation whose third row is (111)/sqrt3: maps the (111) normal onto z."""
    return np.array([np.array([-1, 1, 0]) / np.sqrt(2), np.array([-1, -1, 2]) / np.sqrt(6),
                     np.array([1, 1, 1]) / np.sqrt(3)])


def fcc_eps(systems, theta=None):
    """Unsymmetrised eigenstrains theta (b x n) theta^T / dslip of the FCC set
    (setfcc G.c:1812-1909: float s3, b = <110>/2)."""
    s3 = float(np.float32(np.sqrt(3)))
    ns_ = [(1, 1, 1)] * 3 + [(-1, 1, 1)] * 3 + [(1, -1, 1)] * 3 + [(1, 1, -1)] * 3
    bs_ = [(-1, 1, 0), (0, -1, 1), (1, 0, -1), (-1, -1, 0), (1, 0, 1), (0, -1, 1),
           (-1, -1, 0), (1, 0, -1), (0, -1, -1), (-1, 1, 0), (1, 0, 1), (0, -1, -1)]
    th = np.eye(3) if theta is None else theta
    out = []
    for a in systems:
        n = np.array(ns_[a]) / s3
        b = np.array(bs_[a]) / 2.0
        out.append(th @ np.outer(b, n) @ th.T)
    return np.array(out)
