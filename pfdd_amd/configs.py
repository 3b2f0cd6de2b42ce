"""Seeded benchmark workloads (PFDD GL loop instances), built with the
library's own setup helpers so every consumer sees bit-identical inputs.

Shared setup: normalised isotropic elasticity (mu = C44 = 1, nu = 0.33,
lambda = C12 = 2 nu/(1-2 nu), C11 = lambda + 2 mu), unit grid spacing,
CD2 = 0.1, dslip = 1, uq2 = 1, pi = 3.141592654, Asf2 as main() sets it
(G.c:4278-4285).  Random draws come from SplitMix64 with base seed 20261202
(+ config index); per loop the draws are radius, system, centre x, y, z, and
the applied stress takes six more draws for (11, 12, 13, 22, 23, 33).
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import FP64, MIXED, B_RECOMPUTE, B_STORED, load, make_params

BASE_SEED = 20261202
NU = 0.33
MU = 1.0
LAM = 2 * NU / (1 - 2 * NU)
C11, C12, C44 = LAM + 2 * MU, LAM, MU
PI_REF = 3.141592654


def asf2_reference(ns):
    a = [0.0262] * ns
    for i in range(min(2, ns)):
        a[i] = 0.0064659518668288441
    if ns == 4:
        a[2] = a[3] = 0.013307048604153731
    return a


def setfcc():
    xn = np.zeros((12, 3))
    xb = np.zeros((12, 3))
    load().pfdd_setfcc(xn.ctypes.data, xb.ctypes.data)
    return xn, xb


def setorient(rows):
    r = np.ascontiguousarray(rows, dtype=np.float64)
    th = np.zeros((3, 3))
    load().pfdd_setorient(r.ctypes.data, th.ctypes.data)
    return th


def seteps(xn, xb, dslip, theta):
    ns = len(xn)
    xn = np.ascontiguousarray(xn, dtype=np.float64)
    xb = np.ascontiguousarray(xb, dtype=np.float64)
    ds = np.ascontiguousarray(dslip, dtype=np.float64)
    th = np.ascontiguousarray(theta, dtype=np.float64)
    eps = np.zeros((ns, 3, 3))
    load().pfdd_seteps(ns, xn.ctypes.data, xb.ctypes.data, ds.ctypes.data, th.ctypes.data, eps.ctypes.data)
    return eps


def resolve_tau(sigma, eps):
    s = np.ascontiguousarray(sigma, dtype=np.float64)
    e = np.ascontiguousarray(eps, dtype=np.float64)
    tau = np.zeros(len(e))
    load().pfdd_resolve_tau(len(e), s.ctypes.data, e.ctypes.data, tau.ctypes.data)
    return tau


class Rng:
    def __init__(self, seed):
        self.s = ctypes.c_ulonglong(seed)

    def uniform(self):
        return load().pfdd_uniform01(ctypes.byref(self.s))

    def next64(self):
        return load().pfdd_splitmix64(ctypes.byref(self.s))


def add_loop(xi, sys, normal, center, radius):
    ns_, n1, n2, n3 = xi.shape
    assert xi.dtype == np.complex128 and xi.flags.c_contiguous
    nrm = np.ascontiguousarray(normal, dtype=np.float64)
    c = np.ascontiguousarray(center, dtype=np.float64)
    return load().pfdd_add_loop(xi.ctypes.data, n1, n2, n3, sys, nrm.ctypes.data, c.ctypes.data, float(radius))


# Interface normal and penetration point of main (G.c:4200-4203, 4234-4236).
INTERFACE_N = (np.sqrt(3) / 3.0, 0.0, np.sqrt(6) / 3.0)


def interface_ppoint(shape):
    n1, _, n3 = shape
    return (n1 // 2, 0, n3 * 3 // 4)


def line_points(shape, normal=INTERFACE_N, ppoint=None):
    """pfdd_line_points: flat indices of the interface line of avestrain."""
    n1, n2, n3 = shape
    pp = np.ascontiguousarray(interface_ppoint(shape) if ppoint is None else ppoint, dtype=np.int32)
    nrm = np.ascontiguousarray(normal, dtype=np.float64)
    cnt = load().pfdd_line_points(n1, n2, n3, nrm.ctypes.data, pp.ctypes.data, None, 0)
    out = np.zeros(max(cnt, 1), dtype=np.int64)
    load().pfdd_line_points(n1, n2, n3, nrm.ctypes.data, pp.ctypes.data, out.ctypes.data, cnt)
    return out[:cnt]


# Orientation whose rows are [-110]/sqrt2, [-1-12]/sqrt6, [111]/sqrt3: the
# (111) normal maps onto z, so (111) loops lie in grid planes k = const.
ROWS_111_TO_Z = [[-1 / np.sqrt(2), 1 / np.sqrt(2), 0.0],
                 [-1 / np.sqrt(6), -1 / np.sqrt(6), 2 / np.sqrt(6)],
                 [1 / np.sqrt(3), 1 / np.sqrt(3), 1 / np.sqrt(3)]]


@dataclass
class Workload:
    name: str
    shape: tuple
    eps: np.ndarray          # [ns,3,3]
    normals: np.ndarray      # [ns,3] slip-plane normals in grid coordinates
    xi: np.ndarray           # complex [ns, N1, N2, N3]
    tau: np.ndarray          # [ns]
    sigma: np.ndarray        # [3,3] applied stress / mu
    steps: int
    precision: int = FP64
    bmode: int = B_RECOMPUTE
    loops: list = field(default_factory=list)

    @property
    def ns(self):
        return self.eps.shape[0]

    @property
    def P(self):
        return int(np.prod(self.shape))

    def params(self, precision=None, bmode=None, plan=0):
        ns = self.ns
        return make_params(self.shape, self.eps, C11, C12, C44, [0.1] * ns, asf2_reference(ns), [1.0] * ns,
                           precision=self.precision if precision is None else precision,
                           bmode=self.bmode if bmode is None else bmode, pi_const=PI_REF, plan=plan)

    def s_tensor_avepst(self, sigma=None):
        """S:sigma*mu, the xi-independent part of avepst (avestrain G.c:714-729)."""
        sig = self.sigma if sigma is None else sigma
        S11 = 1. / 3. * (1 / (C11 + 2 * C12) + 2 / (C11 - C12))
        S12 = 1. / 3. * (1 / (C11 + 2 * C12) - 1 / (C11 - C12))
        S44 = 1 / C44
        d = np.eye(3)
        out = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                v = 0.0
                for k in range(3):
                    for l in range(3):
                        d4 = 1.0 if i == j == k == l else 0.0
                        Sijkl = ((S11 + 2 * S12) / 3.0 * d[i, j] * d[k, l]
                                 + S44 / 4.0 * (d[i, k] * d[j, l] + d[i, l] * d[j, k])
                                 - S44 / 2.0 * d4 + (S11 - S12) * (d4 - d[i, j] * d[k, l] / 3.0))
                        v += Sijkl * sig[k, l] * MU
                out[i, j] = v
        return out


def _fcc_systems(systems, theta):
    xn, xb = setfcc()
    xn, xb = xn[systems], xb[systems]
    eps = seteps(xn, xb, [1.0] * len(systems), theta)
    normals = (theta @ xn.T).T
    return eps, normals


def _random_loops(rng, xi, normals, nloops, rmin, rmax):
    """pfdd_seed_loops: per loop r, system, centre (x, y, z) from the stream."""
    ns, n1, n2, n3 = xi.shape
    assert xi.dtype == np.complex128 and xi.flags.c_contiguous
    nrm = np.ascontiguousarray(normals, dtype=np.float64)
    out = np.zeros((max(nloops, 1), 6))
    rc = load().pfdd_seed_loops(xi.ctypes.data, n1, n2, n3, ns, nrm.ctypes.data, nloops, float(rmin), float(rmax),
                                ctypes.byref(rng.s), out.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"pfdd_seed_loops failed: {rc}")
    return [(int(o[0]), (o[1], o[2], o[3]), o[4], int(o[5])) for o in out[:nloops]]


def _random_sigma(rng, amp):
    s = np.zeros((3, 3))
    load().pfdd_seed_sigma(float(amp), ctypes.byref(rng.s), s.ctypes.data)
    return s


def c0():
    """32^3, NS=3: the three a/2<110> systems of one (111) plane, crystal
    rotated so the plane is z = const; one loop of radius 8 on plane z = 16
    (centre (16,16,16)) on system 0; sigma gives tau_0 = 0.01 mu."""
    shape = (32, 32, 32)
    theta = setorient(ROWS_111_TO_Z)
    eps, normals = _fcc_systems([0, 1, 2], theta)
    xi = np.zeros((3,) + shape, dtype=np.complex128)
    cnt = add_loop(xi, 0, normals[0], (16.0, 16.0, 16.0), 8.0)
    e0 = eps[0]
    se = 0.5 * (e0 + e0.T)
    sigma = 0.01 * se / np.sum(se * e0)
    return Workload("c0", shape, eps, normals, xi, resolve_tau(sigma, eps), sigma, 500,
                    loops=[(0, (16.0, 16.0, 16.0), 8.0, cnt)])


def _fcc_random(name, shape, nloops, seed, steps, precision, bmode):
    theta = setorient(np.eye(3))
    eps, normals = _fcc_systems(list(range(12)), theta)
    xi = np.zeros((12,) + tuple(shape), dtype=np.complex128)
    rng = Rng(seed)
    loops = _random_loops(rng, xi, normals, nloops, 5.0, 20.0)
    sigma = _random_sigma(rng, 0.02)
    return Workload(name, tuple(shape), eps, normals, xi, resolve_tau(sigma, eps), sigma, steps, precision, bmode,
                    loops)


def c1(bmode=B_RECOMPUTE):
    """128^3, NS=12 (FCC), 16 seeded loops r ~ U[5,20], sigma_ij ~ U[-0.02,0.02]."""
    return _fcc_random("c1", (128, 128, 128), 16, BASE_SEED + 1, 1000, FP64, bmode)


def c2(bmode=B_STORED):
    """256^3, NS=12 (FCC), 64 seeded loops, mixed precision."""
    return _fcc_random("c2", (256, 256, 256), 64, BASE_SEED + 2, 2000, MIXED, bmode)


def c4(bmode=B_RECOMPUTE, shape=(512, 512, 512), steps=5000):
    """512^3, NS=3: the three systems of one (111) plane (rotated so the plane
    is z = const), 8 seeded loops r ~ U[20,60]; the applied stress
    sigma = s_hat * 0.03 * t/(steps-1), s_hat = sym(e_0)/(sym(e_0):e_0), ramps the
    resolved shear of system 0 linearly from 0 to 0.03 mu over the run.
    Workload.tau holds tau(0) = 0 and Workload.tau_rate the per-step increment."""
    theta = setorient(ROWS_111_TO_Z)
    eps, normals = _fcc_systems([0, 1, 2], theta)
    xi = np.zeros((3,) + tuple(shape), dtype=np.complex128)
    rng = Rng(BASE_SEED + 4)
    loops = _random_loops(rng, xi, normals, 8, 20.0, 60.0)
    e0 = eps[0]
    se = 0.5 * (e0 + e0.T)
    s_hat = se / np.sum(se * e0)
    w = Workload("c4", tuple(shape), eps, normals, xi, np.zeros(3), 0.03 * s_hat, steps, MIXED, bmode, loops)
    w.sigma_unit = s_hat
    w.tau_rate = resolve_tau(s_hat, eps) * 0.03 / (steps - 1)
    return w


def small_fcc(shape, nloops=4, seed=BASE_SEED + 100, ns=12, rmin=2.0, rmax=6.0):
    """A reduced FCC instance with the c1 generator, for parity tests."""
    theta = setorient(np.eye(3))
    eps, normals = _fcc_systems(list(range(ns)), theta)
    xi = np.zeros((ns,) + tuple(shape), dtype=np.complex128)
    rng = Rng(seed)
    loops = _random_loops(rng, xi, normals, nloops, rmin, rmax)
    sigma = _random_sigma(rng, 0.02)
    return Workload(f"fcc{shape}", tuple(shape), eps, normals, xi, resolve_tau(sigma, eps), sigma, 10, loops=loops)


C3_SIZES = (32, 48, 64, 96)


@dataclass
class Member:
    """One instance of an ensemble workload."""
    index: int
    shape: tuple
    xi: np.ndarray           # complex [ns, N, N, N]
    tau: np.ndarray          # [ns]
    sigma: np.ndarray        # [3,3]
    loops: list

    @property
    def P(self):
        return int(np.prod(self.shape))


class EnsembleWorkload:
    """Independent NS=3 instances (the systems of one (111) plane, as c0).
    Instance i draws from SplitMix64 seeded BASE_SEED + 1000 + i: the grid
    edge from `sizes` (one draw), the loop count 1..4 (one draw), the loops
    (r ~ U[4,15], system, centre), then the resolved shear tau ~ U[0, 0.05]
    of system 0, applied as sigma = tau * sym(e_0)/(sym(e_0):e_0).
    Members are generated on demand, so the whole ensemble never sits in
    host memory at once."""

    def __init__(self, name, count, steps, sizes=C3_SIZES, rmin=4.0, rmax=15.0, tau_max=0.05, first=0):
        self.name, self.count, self.steps = name, count, steps
        self.sizes, self.rmin, self.rmax, self.tau_max, self.first = tuple(sizes), rmin, rmax, tau_max, first
        theta = setorient(ROWS_111_TO_Z)
        self.eps, self.normals = _fcc_systems([0, 1, 2], theta)
        e0 = self.eps[0]
        se = 0.5 * (e0 + e0.T)
        self.s_hat = se / np.sum(se * e0)
        self._edges = [self._edge(Rng(self.seed(i))) for i in range(count)]

    def seed(self, i):
        return BASE_SEED + 1000 + self.first + i

    def _edge(self, rng):
        return self.sizes[min(int(rng.uniform() * len(self.sizes)), len(self.sizes) - 1)]

    @property
    def ns(self):
        return 3

    def shape(self, i):
        n = self._edges[i]
        return (n, n, n)

    def total_points(self):
        return sum(n ** 3 for n in self._edges)

    def member(self, i):
        rng = Rng(self.seed(i))
        n = self._edge(rng)
        nloops = 1 + min(int(rng.uniform() * 4), 3)
        xi = np.zeros((3, n, n, n), dtype=np.complex128)
        loops = _random_loops(rng, xi, self.normals, nloops, self.rmin, self.rmax)
        tau0 = self.tau_max * rng.uniform()
        sigma = tau0 * self.s_hat
        return Member(i, (n, n, n), xi, resolve_tau(sigma, self.eps), sigma, loops)

    def params(self, i, precision=MIXED, bmode=B_RECOMPUTE, plan=0):
        return make_params(self.shape(i), self.eps, C11, C12, C44, [0.1] * 3, asf2_reference(3), [1.0] * 3,
                           precision=precision, bmode=bmode, pi_const=PI_REF, plan=plan)

    def params_list(self, precision=MIXED, bmode=B_RECOMPUTE, plan=0):
        return [self.params(i, precision, bmode, plan) for i in range(self.count)]


def c3(count=512, steps=500, sizes=C3_SIZES):
    """512 independent NS=3 instances on grids drawn from {32,48,64,96}^3,
    1-4 loops of radius 4-15 each, tau ~ U[0,0.05] mu; 500 GL steps."""
    return EnsembleWorkload("c3", count, steps, sizes)


BY_NAME = {"c0": c0, "c1": c1, "c2": c2, "c4": c4}
