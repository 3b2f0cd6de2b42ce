"""ctypes view of oracle/_ref/libpfdd_ref.so: the reference PFDD_Final_GPU.c
compiled from its own source (FOURN backend, main renamed) by oracle/Makefile.

The library carries the reference's compile-time grid: N1=32, N2=2, N3=32,
NS=2 (NMAT=2 x NS1=1), NV=9.  Only tests/ use it, to pin the oracle.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_ref", "libpfdd_ref.so")

N1, N2, N3, NS, NV, NMAT = 32, 2, 32, 2, 9, 2
NSV = NS + NV
P = N1 * N2 * N3


class Material(ctypes.Structure):
    _fields_ = [("C11", ctypes.c_double), ("C12", ctypes.c_double), ("C44", ctypes.c_double)]


_lib = None


def available():
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIB_PATH)
        vp, ci, cd, cf = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_float
        L.fourn.argtypes = [vp, vp, ci, ci]
        L.frec.argtypes = [vp, vp, vp, cd, cd, cd]
        L.Bmatrix.argtypes = [vp, vp, vp, vp, vp, vp, cd, cd, cd, vp, vp]
        L.Fmatrix.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, cd, cd, cd, vp, vp, vp]
        L.plasticevolv.argtypes = [vp, vp, vp, cf, vp, vp, vp, vp, vp, vp, vp, ci, cd, cd, cd, ci]
        L.plasticevolv.restype = cd
        L.virtualevolv.argtypes = [vp, vp, vp, vp, vp, vp, cd, vp, ci, cd, cd, cd, cd, vp, ci, ci, ci,
                                   vp, vp, vp, vp, vp]
        L.strain.argtypes = [vp, vp, vp, vp, vp, vp, cd, cd, cd, cd, vp, ci, ci, vp]
        L.stress.argtypes = [vp, vp, vp, vp, vp, vp, vp, cd, cd, cd, vp, ci, ci, vp, vp, vp, vp, vp, vp,
                             ci, ci, ci, ci, ci]
        L.seteps.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.setorient.argtypes = [vp]
        L.set3D1sys.argtypes = [vp, vp]
        L.resolSS.argtypes = [vp, vp, vp, vp, vp, vp]
        L.avestrain.argtypes = [vp, vp, vp, vp, vp, ci, vp, cd, cd, cd, cd, vp, vp, ci, vp, vp]
        L.avestrain.restype = cf
        L.ResidualEnergy.argtypes = [vp, vp, ci, ci, ci, cd, cd, cd, cd]
        L.ResidualEnergy.restype = cf
        L.Energy_calculation.argtypes = [vp, vp, vp, vp, vp, cd, cd, cd, vp, vp, ci, ci, ci]
        L.Energy_calculation.restype = cf
        _lib = L
    return _lib


def fourn3d(x, isign, nn=None):
    """The reference fourn on a complex64 array (in place on a copy)."""
    a = np.ascontiguousarray(x, dtype=np.complex64).copy()
    shape = a.shape if nn is None else nn
    nnarr = (ctypes.c_int * 4)(0, *shape)
    # fourn takes data-1 (1-based indexing, G.c:102-107)
    lib().fourn(a.ctypes.data - 4, nnarr, len(shape), isign)
    return a


def frec():
    fx, fy, fz = (np.zeros(P) for _ in range(3))
    lib().frec(fx.ctypes.data, fy.ctypes.data, fz.ctypes.data, 1.0, 1.0, 1.0)
    return fx, fy, fz


def epsv_basis():
    """Virtual-strain bases of seteps (G.c:1917-1948): e_v = xn1 (x) xb1."""
    ev = np.zeros((NV, 3, 3))
    for v in range(NV):
        ev[v, v // 3, v % 3] = 1.0
    return ev


def bmatrix(eps, epsv, C11, C12, C44):
    """BB[kb][ka][P] for ka < NS, kb < NSV (layout of G.c:58)."""
    fx, fy, fz = frec()
    BB = np.zeros(NS * NSV * P)
    e = np.ascontiguousarray(eps, dtype=np.float64)
    ev = np.ascontiguousarray(epsv, dtype=np.float64)
    mc = Material(C11, C12, C44)
    xi_o = np.zeros(P, dtype=np.int32)
    lib().Bmatrix(BB.ctypes.data, fx.ctypes.data, fy.ctypes.data, fz.ctypes.data, e.ctypes.data,
                  ev.ctypes.data, 1.0, 1.0, 1.0, ctypes.byref(mc), xi_o.ctypes.data)
    return BB.reshape(NSV, NS, P)


def fmatrix(eps, epsv, C11, C12, C44):
    """(FF [3(j)][3(i)][NS][P], DD [3][3][NSV][P]) as laid out by G.c:53-56."""
    fx, fy, fz = frec()
    FF = np.zeros(NS * 9 * P)
    DD = np.zeros(NSV * 9 * P)
    FFv = np.zeros(NV * 9 * P)
    e = np.ascontiguousarray(eps, dtype=np.float64)
    ev = np.ascontiguousarray(epsv, dtype=np.float64)
    mc = Material(C11, C12, C44)
    xi_o = np.zeros(P, dtype=np.int32)
    theta = np.zeros((NMAT, 3, 3))
    theta[:] = np.eye(3)
    lib().Fmatrix(FF.ctypes.data, DD.ctypes.data, FFv.ctypes.data, fx.ctypes.data, fy.ctypes.data,
                  fz.ctypes.data, e.ctypes.data, ev.ctypes.data, 1.0, 1.0, 1.0, ctypes.byref(mc),
                  xi_o.ctypes.data, theta.ctypes.data)
    return FF.reshape(3, 3, NS, P), DD.reshape(3, 3, NSV, P)


def plasticevolv(xi, xi_bc, data2, cd2, asf2, tau_pt, dslip, uq2=1.0):
    """Reference plasticevolv on copies.  xi, xi_bc: complex128 [NSV,P];
    data2: complex64 [NSV,P]; tau_pt: [P, NS]."""
    x = np.ascontiguousarray(xi, dtype=np.complex128).copy()
    bc = np.ascontiguousarray(xi_bc, dtype=np.complex128)
    d2 = np.ascontiguousarray(data2, dtype=np.complex64)
    dg = np.zeros_like(d2)
    dat = np.zeros_like(d2)
    c = np.ascontiguousarray(cd2, dtype=np.float64)
    a = np.ascontiguousarray(asf2, dtype=np.float64)
    t = np.ascontiguousarray(tau_pt, dtype=np.float64)
    s = np.ascontiguousarray(dslip, dtype=np.float64)
    g1 = np.zeros(NS)
    g = lib().plasticevolv(bc.ctypes.data, x.ctypes.data, c.ctypes.data, uq2, d2.ctypes.data, a.ctypes.data,
                           t.ctypes.data, s.ctypes.data, dg.ctypes.data, dat.ctypes.data, g1.ctypes.data,
                           P, 0.0, 0.0, 1.0, 0)
    return x, g, g1


def virtualevolv(data, data2, DD, xi, xi_bc):
    """virtualevolv with it past the virtual-update window (it=5 >= NT-NTD),
    so only its data2 side effect and inverse FFT run (G.c:1426-1447)."""
    d = np.ascontiguousarray(data, dtype=np.complex64).copy()
    d2 = np.ascontiguousarray(data2, dtype=np.complex64).copy()
    x = np.ascontiguousarray(xi, dtype=np.complex128).copy()
    bc = np.ascontiguousarray(xi_bc, dtype=np.complex128)
    DDc = np.ascontiguousarray(DD, dtype=np.float64)
    sigmav = np.zeros(NV * P, dtype=np.complex64)
    sigmal = np.zeros(9 * P)
    II = np.zeros(9 * P)
    xi_o = np.zeros(P, dtype=np.int32)
    zl = np.zeros((NMAT, 6), dtype=np.int32)
    prop = np.zeros((NMAT, 6), dtype=np.int32)
    dS = np.zeros((NMAT, 6, 6))
    lib().virtualevolv(d.ctypes.data, d2.ctypes.data, sigmav.ctypes.data, DDc.ctypes.data, x.ctypes.data,
                       bc.ctypes.data, 7e-5, sigmal.ctypes.data, 0, 1.0, 1.0, 1.0, 1.0, None, 5, 0, 0,
                       II.ctypes.data, xi_o.ctypes.data, zl.ctypes.data, prop.ctypes.data, dS.ctypes.data)
    return d, d2


def strain_stress(data, FF, xi, eps, epsv, avepst_pt, C11, C12, C44):
    """strain() then stress() with printing switched off (it=5).  data is the
    forward transform of xi (complex64 [NSV,P]); avepst_pt [P,3,3].
    Returns (dataeps [9(i+3j)][P] complex64, datasigma, avesigma[3][3])."""
    databeta = np.zeros(9 * P, dtype=np.complex64)
    dataeps = np.zeros(9 * P, dtype=np.complex64)
    d = np.ascontiguousarray(data, dtype=np.complex64).copy()
    FFc = np.ascontiguousarray(FF, dtype=np.float64)
    FFv = np.zeros(NV * 9 * P)
    ev = np.ascontiguousarray(epsv, dtype=np.float64)
    ap = np.ascontiguousarray(avepst_pt, dtype=np.float64)
    lib().strain(databeta.ctypes.data, dataeps.ctypes.data, d.ctypes.data, FFc.ctypes.data, FFv.ctypes.data,
                 ev.ctypes.data, 1.0, 1.0, 1.0, 1.0, None, 5, 0, ap.ctypes.data)
    dataepsd = np.zeros(9 * P, dtype=np.complex64)
    datasigma = np.zeros(9 * P, dtype=np.complex64)
    sigmav = np.zeros(NV * P, dtype=np.complex64)
    x = np.ascontiguousarray(xi, dtype=np.complex128)
    e = np.ascontiguousarray(eps, dtype=np.float64)
    avesig = np.zeros((3, 3))
    theta = np.zeros((NMAT, 3, 3))
    theta[:] = np.eye(3)
    sd = np.array([1.0, 0.0, 0.0])
    xn = np.zeros((NS, 3))
    pen = np.zeros(N2)
    pen2 = np.zeros(N2)
    lib().stress(dataepsd.ctypes.data, datasigma.ctypes.data, dataeps.ctypes.data, sigmav.ctypes.data,
                 x.ctypes.data, e.ctypes.data, ev.ctypes.data, C11, C12, C44, None, 5, 0, avesig.ctypes.data,
                 theta.ctypes.data, sd.ctypes.data, xn.ctypes.data, pen.ctypes.data, pen2.ctypes.data,
                 7, 24, N1 // 2, 0, 24)
    return dataeps.reshape(9, P), datasigma.reshape(9, P), avesig


def setorient():
    th = np.zeros((NMAT, 3, 3))
    lib().setorient(th.ctypes.data)
    return th


def set3D1sys():
    xn = np.zeros((NS, 3))
    xb = np.zeros((NS, 3))
    lib().set3D1sys(xn.ctypes.data, xb.ctypes.data)
    return xn, xb


def seteps(xn, xb, dslip, theta):
    eps = np.zeros((NS, 3, 3))
    epsv = np.zeros((NV, 3, 3))
    AA = np.zeros((NMAT, 3, 3))
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (xn, xb, dslip, theta)]
    lib().seteps(eps.ctypes.data, epsv.ctypes.data, a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data,
                 AA.ctypes.data, a[3].ctypes.data)
    return eps, epsv


def resolSS(sigma_pt, eps):
    """tau[P][NS] = sigma(x):eps_a (G.c:1607-1622); sigma_pt [P,3,3]."""
    sg = np.ascontiguousarray(sigma_pt, dtype=np.float64)
    tau = np.zeros((P, NS))
    e = np.ascontiguousarray(eps, dtype=np.float64)
    xi_bc = np.zeros(2 * NSV * P)
    sigmal = np.zeros(9 * P)
    xi_o = np.zeros(P, dtype=np.int32)
    lib().resolSS(sg.ctypes.data, tau.ctypes.data, e.ctypes.data, xi_bc.ctypes.data, sigmal.ctypes.data,
                  xi_o.ctypes.data)
    return tau


def gl_loop(xi0, eps, epsv, C11, C12, C44, cd2, asf2, dslip, tau, nsteps):
    """The GL loop of main (G.c:4450-4452, virtualevolv G.c:1426-1447, then
    G.c:4574-4644) driven through the reference's own compiled functions.
    xi0: complex [NS,P] initial order parameter (virtual fields zero).
    Returns xi [NS,P] complex128 after nsteps and gamma per step."""
    BB = bmatrix(eps, epsv, C11, C12, C44)          # [NSV][NS][P]
    FF, DD = fmatrix(eps, epsv, C11, C12, C44)
    xi = np.zeros((NSV, P), dtype=np.complex128)
    xi[:NS] = xi0
    xi_bc = np.zeros((NSV, P), dtype=np.complex128)
    xi_bc[NS:] = 1.0
    tau_pt = np.tile(np.asarray(tau, dtype=np.float64), (P, 1))

    def forward(xi):
        data = xi.astype(np.complex64)
        for b in range(NSV):
            data[b] = fourn3d(data[b].reshape(N1, N2, N3), -1).reshape(P)
        return data

    def contract(data):
        # data2 zeroed, then DATA2 += DATA*BB for isa<NS, isb<NSV (G.c:4602-4629);
        # the sum is formed in double and rounded to float on every store.
        d2 = np.zeros((NSV, P), dtype=np.complex64)
        for a in range(NS):
            acc = d2[a]
            for b in range(NSV):
                re = (acc.real.astype(np.float64) + data[b].real.astype(np.float64) * BB[b, a]).astype(np.float32)
                im = (acc.imag.astype(np.float64) + data[b].imag.astype(np.float64) * BB[b, a]).astype(np.float32)
                acc = re + 1j * im
            d2[a] = acc.astype(np.complex64)
        return d2

    data = forward(xi)
    d2 = contract(data)
    data, d2 = virtualevolv(data, d2, DD.reshape(-1), xi, xi_bc)
    gammas = []
    for _ in range(nsteps):
        xi, g, _g1 = plasticevolv(xi, xi_bc, d2, cd2, asf2, tau_pt, dslip)
        gammas.append(g)
        data = forward(xi)
        d2 = contract(data)
        for a in range(NS):
            d2[a] = fourn3d(d2[a].reshape(N1, N2, N3), 1).reshape(P)
    return xi[:NS], np.array(gammas)


def avestrain(xi, eps, sigma, S11, S12, S44, mu, border, normal, ppoint):
    """The reference avestrain on xi [NSV][N1][N2][N3] (complex) with a uniform
    applied stress sigma[3][3]: (energy_stressTimePf, energy_in4 added)."""
    x = np.zeros((NSV, N1, N2, N3, 2))
    x[..., 0] = np.asarray(xi).real
    x[..., 1] = np.asarray(xi).imag
    avepsd = np.zeros((3, 3))
    avepst = np.zeros((P, 3, 3))
    e = np.ascontiguousarray(eps, dtype=np.float64).reshape(NS, 3, 3)
    ev = epsv_basis()
    sg = np.ascontiguousarray(np.broadcast_to(np.asarray(sigma, dtype=np.float64), (P, 3, 3)))
    e4 = ctypes.c_float(0.0)
    sa = ctypes.c_float(0.0)
    n = np.ascontiguousarray(normal, dtype=np.float64)
    pp = np.ascontiguousarray(ppoint, dtype=np.int32)
    r = lib().avestrain(avepsd.ctypes.data, avepst.ctypes.data, e.ctypes.data, ev.ctypes.data, x.ctypes.data, P,
                        sg.ctypes.data, S11, S12, S44, mu, ctypes.byref(e4), ctypes.byref(sa), border,
                        n.ctypes.data, pp.ctypes.data)
    return float(r), float(e4.value)


def residual_energy(xi, normal, ppoint, D00, D01, D10, D11):
    x = np.zeros((NSV, N1, N2, N3, 2))
    x[..., 0] = np.asarray(xi).real
    n = np.ascontiguousarray(normal, dtype=np.float64)
    return float(lib().ResidualEnergy(x.ctypes.data, n.ctypes.data, int(ppoint[0]), int(ppoint[1]),
                                      int(ppoint[2]), D00, D01, D10, D11))



def energy_calculation(data, eps, C11, C12, C44, normal, ppoint):
    """The reference Energy_calculation on the FFT array data [NSV][N1][N2][N3]
    (complex, stored as float like the reference's data)."""
    d = np.zeros((NSV, N1, N2, N3, 2), dtype=np.float32)
    d[..., 0] = np.asarray(data).real
    d[..., 1] = np.asarray(data).imag
    fx, fy, fz = frec()
    e = np.ascontiguousarray(eps, dtype=np.float64).reshape(NS, 3, 3)
    ev = epsv_basis()
    n = np.ascontiguousarray(normal, dtype=np.float64)
    return float(lib().Energy_calculation(fx.ctypes.data, fy.ctypes.data, fz.ctypes.data, e.ctypes.data,
                                          ev.ctypes.data, C11, C12, C44, d.ctypes.data, n.ctypes.data,
                                          int(ppoint[0]), int(ppoint[1]), int(ppoint[2])))
