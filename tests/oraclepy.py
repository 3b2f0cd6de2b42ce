"""ctypes view of oracle/liboracle.so (the FP64 CPU checker).

Test infrastructure: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py only.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("n1", ctypes.c_int), ("n2", ctypes.c_int), ("n3", ctypes.c_int), ("ns", ctypes.c_int),
        ("d1", ctypes.c_double), ("d2", ctypes.c_double), ("d3", ctypes.c_double),
        ("C11", ctypes.c_double), ("C12", ctypes.c_double), ("C44", ctypes.c_double),
        ("cd2", ctypes.c_double * 64), ("asf2", ctypes.c_double * 64), ("dslip", ctypes.c_double * 64),
        ("uq2", ctypes.c_double), ("pi_const", ctypes.c_double),
        ("eps", (ctypes.c_double * 9) * 64),
        ("nthreads", ctypes.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.orc_fft3d.argtypes = [vp, ci, ci, ci, ci, ci]
        L.orc_fft3d.restype = ci
        L.orc_frec1.argtypes = [ci, ci, cd]
        L.orc_frec1.restype = cd
        L.orc_cijkl.argtypes = [vp, cd, cd, cd]
        for nm in ("orc_bmatrix_literal_at", "orc_bmatrix_at"):
            getattr(L, nm).argtypes = [vp, ci, vp, ci, vp, ci, cd, cd, cd, vp]
        L.orc_ddsum_at.argtypes = [vp, ci, vp, ci, cd, cd, cd, vp]
        L.orc_fmatrix_at.argtypes = [vp, ci, vp, ci, cd, cd, cd, vp]
        L.orc_create.argtypes = [ctypes.POINTER(OrcParams)]
        L.orc_create.restype = vp
        L.orc_destroy.argtypes = [vp]
        L.orc_set_xi.argtypes = [vp, vp, vp]
        L.orc_get_xi.argtypes = [vp, vp]
        L.orc_set_tau.argtypes = [vp, vp, ci]
        L.orc_set_tau_ramp.argtypes = [vp, vp, vp]
        L.orc_prologue.argtypes = [vp, ci]
        L.orc_gl_steps.argtypes = [vp, ci, vp, vp]
        L.orc_plasticevolv.argtypes = [ci, ctypes.c_long, vp, vp, vp, vp, vp, vp, vp, ci, cd, cd, vp]
        L.orc_plasticevolv.restype = cd
        L.orc_strain_stress.argtypes = [vp, vp, vp, vp, vp, vp]
        L.orc_line_points.argtypes = [ci, ci, ci, vp, vp, vp, ci]
        L.orc_line_points.restype = ci
        L.orc_line_energy.argtypes = [vp, vp, cd, cd, cd, cd, vp, ci, vp]
        L.orc_line_energy.restype = cd
        L.orc_residual_energy.argtypes = [vp, vp, ci, cd, cd, cd, cd]
        L.orc_residual_energy.restype = cd
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def fft3d(x, sign):
    """Unnormalised batched 3D DFT of a complex array [..., n1, n2, n3]."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    n1, n2, n3 = a.shape[-3:]
    batch = int(np.prod(a.shape[:-3])) if a.ndim > 3 else 1
    rc = lib().orc_fft3d(a.ctypes.data, n1, n2, n3, batch, sign)
    if rc != 0:
        raise ValueError("unsupported FFT size")
    return a


def _kvec(fk):
    fk = np.ascontiguousarray(fk, dtype=np.float64).reshape(-1, 3)
    return fk, fk.shape[0]


def bmatrix_at(fk, eps_a, eps_b, C11, C12, C44, literal=False):
    fk, n = _kvec(fk)
    ea = np.ascontiguousarray(eps_a, dtype=np.float64).reshape(-1, 9)
    eb = np.ascontiguousarray(eps_b, dtype=np.float64).reshape(-1, 9)
    out = np.zeros((n, ea.shape[0], eb.shape[0]))
    fn = lib().orc_bmatrix_literal_at if literal else lib().orc_bmatrix_at
    fn(fk.ctypes.data, n, ea.ctypes.data, ea.shape[0], eb.ctypes.data, eb.shape[0], C11, C12, C44, out.ctypes.data)
    return out


def ddsum_at(fk, eps_b, C11, C12, C44):
    fk, n = _kvec(fk)
    eb = np.ascontiguousarray(eps_b, dtype=np.float64).reshape(-1, 9)
    out = np.zeros((n, eb.shape[0]))
    lib().orc_ddsum_at(fk.ctypes.data, n, eb.ctypes.data, eb.shape[0], C11, C12, C44, out.ctypes.data)
    return out


def fmatrix_at(fk, eps_a, C11, C12, C44):
    fk, n = _kvec(fk)
    ea = np.ascontiguousarray(eps_a, dtype=np.float64).reshape(-1, 9)
    out = np.zeros((n, ea.shape[0], 3, 3))
    lib().orc_fmatrix_at(fk.ctypes.data, n, ea.ctypes.data, ea.shape[0], C11, C12, C44, out.ctypes.data)
    return out


def frec1(i, n, d=1.0):
    return lib().orc_frec1(i, n, d)


def kgrid(n1, n2, n3, d=(1.0, 1.0, 1.0)):
    """fx, fy, fz on the grid as [n1,n2,n3] arrays (frec)."""
    f = [np.array([frec1(i, n, dd) for i in range(n)]) for n, dd in zip((n1, n2, n3), d)]
    FX, FY, FZ = np.meshgrid(f[0], f[1], f[2], indexing="ij")
    return FX, FY, FZ


def make_params(n, eps, C11, C12, C44, cd2, asf2, dslip, uq2=1.0, pi_const=3.141592654, nthreads=0):
    p = OrcParams()
    p.n1, p.n2, p.n3 = n
    ns = len(eps)
    p.ns = ns
    p.d1 = p.d2 = p.d3 = 1.0
    p.C11, p.C12, p.C44 = C11, C12, C44
    for a in range(ns):
        p.cd2[a] = cd2[a]
        p.asf2[a] = asf2[a]
        p.dslip[a] = dslip[a]
        for t in range(9):
            p.eps[a][t] = float(np.asarray(eps[a]).reshape(9)[t])
    p.uq2 = uq2
    p.pi_const = pi_const
    p.nthreads = nthreads
    return p


class Oracle:
    """One FP64 GL-loop instance (orc_state)."""

    def __init__(self, params):
        self.p = params
        self.ns = params.ns
        self.shape = (params.n1, params.n2, params.n3)
        self.P = params.n1 * params.n2 * params.n3
        self.h = lib().orc_create(ctypes.byref(params))

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_xi(self, xi, xi_bc=None):
        """xi: complex [ns, n1, n2, n3] (or float re/im interleaved [ns,P,2])."""
        a = np.ascontiguousarray(xi, dtype=np.complex128).reshape(self.ns, self.P)
        b = None
        if xi_bc is not None:
            b = np.ascontiguousarray(xi_bc, dtype=np.complex128).reshape(self.ns, self.P)
        lib().orc_set_xi(self.h, a.ctypes.data, _ptr(b))

    def get_xi(self):
        a = np.zeros((self.ns,) + self.shape, dtype=np.complex128)
        lib().orc_get_xi(self.h, a.ctypes.data)
        return a

    def set_tau(self, tau, per_point=False):
        t = np.ascontiguousarray(tau, dtype=np.float64)
        lib().orc_set_tau(self.h, t.ctypes.data, 1 if per_point else 0)

    def set_tau_ramp(self, tau0, rate):
        t = np.ascontiguousarray(tau0, dtype=np.float64)
        r = np.ascontiguousarray(rate, dtype=np.float64)
        lib().orc_set_tau_ramp(self.h, t.ctypes.data, r.ctypes.data)

    def prologue(self, quirk=True):
        lib().orc_prologue(self.h, 1 if quirk else 0)

    def gl_steps(self, n):
        g = np.zeros(n)
        ga = np.zeros((n, self.ns))
        lib().orc_gl_steps(self.h, n, g.ctypes.data, ga.ctypes.data)
        return g, ga

    def line_energy(self, sigma, S11, S12, S44, mu, pts):
        """(energy_stressTimePf, energy_in4 increment) of avestrain on the line points."""
        sg = np.ascontiguousarray(sigma, dtype=np.float64).reshape(9)
        p = np.ascontiguousarray(pts, dtype=np.int64)
        e4 = ctypes.c_double()
        e3 = lib().orc_line_energy(self.h, sg.ctypes.data, S11, S12, S44, mu, p.ctypes.data, len(p),
                                   ctypes.byref(e4))
        return e3, e4.value

    def residual_energy(self, ppoint, ns1, D00, D01, D10, D11):
        pp = np.ascontiguousarray(ppoint, dtype=np.int32)
        return lib().orc_residual_energy(self.h, pp.ctypes.data, ns1, D00, D01, D10, D11)

    def strain_stress(self, avepst_uniform, fields=False):
        ap = np.ascontiguousarray(avepst_uniform, dtype=np.float64).reshape(9)
        aveps = np.zeros(9)
        avesig = np.zeros(9)
        ef = sf = None
        if fields:
            ef = np.zeros((9, self.P))
            sf = np.zeros((9, self.P))
        lib().orc_strain_stress(self.h, ap.ctypes.data, aveps.ctypes.data, avesig.ctypes.data, _ptr(ef), _ptr(sf))
        if fields:
            return aveps.reshape(3, 3), avesig.reshape(3, 3), ef.reshape((3, 3) + self.shape), sf.reshape((3, 3) + self.shape)
        return aveps.reshape(3, 3), avesig.reshape(3, 3)


def line_points(shape, normal, ppoint):
    """Point indices of the interface line of the energy diagnostics."""
    n = np.ascontiguousarray(normal, dtype=np.float64)
    pp = np.ascontiguousarray(ppoint, dtype=np.int32)
    cnt = lib().orc_line_points(*shape, n.ctypes.data, pp.ctypes.data, None, 0)
    pts = np.zeros(max(cnt, 1), dtype=np.int64)
    lib().orc_line_points(*shape, n.ctypes.data, pp.ctypes.data, pts.ctypes.data, cnt)
    return pts[:cnt]


def plasticevolv(xi, d2, cd2, asf2, dslip, tau, xi_bc=None, tau_per_point=False, uq2=1.0, pi_const=3.141592654):
    """One update on copies; returns (xi_new, gamma, gamma_a)."""
    ns = xi.shape[0]
    P = int(np.prod(xi.shape[1:]))
    x = np.ascontiguousarray(xi, dtype=np.complex128).copy()
    d = np.ascontiguousarray(d2, dtype=np.complex128)
    bc = None if xi_bc is None else np.ascontiguousarray(xi_bc, dtype=np.complex128)
    ga = np.zeros(ns)
    arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (cd2, asf2, dslip, tau)]
    g = lib().orc_plasticevolv(ns, P, x.ctypes.data, _ptr(bc), d.ctypes.data, arrs[0].ctypes.data,
                               arrs[1].ctypes.data, arrs[2].ctypes.data, arrs[3].ctypes.data,
                               1 if tau_per_point else 0, uq2, pi_const, ga.ctypes.data)
    return x, g, ga
