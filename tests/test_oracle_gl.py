"""Pins the oracle's GL step -- rows a4-a7/a10 (contraction G.c:4609-4629,
virtualevolv side effect G.c:1426-1447, plasticevolv G.c:3470-3501, loop order
G.c:4574-4644)."""
import numpy as np
import pytest

import oraclepy as O
import physics as ph
import refpy as R


def test_plasticevolv_matches_reference_exactly():
    """Inputs chosen exact in FP32 so the reference's float term uq2*d2/nsize
    (G.c:3488) equals the double one; the update is then bit-identical."""
    if not R.available():
        pytest.skip("compiled reference not built")
    rng = np.random.default_rng(9)
    xi = np.zeros((R.NSV, R.P), dtype=np.complex128)
    xi[:R.NS] = rng.integers(-64, 64, (R.NS, R.P)) / 64.0 + 1j * rng.integers(-8, 8, (R.NS, R.P)) / 256.0
    bc = np.zeros((R.NSV, R.P), dtype=np.complex128)
    bc[R.NS:] = 1.0
    bc[0, ::7] = 1.0                                   # some frozen points
    d2 = np.zeros((R.NSV, R.P), dtype=np.complex64)
    d2[:R.NS] = (rng.integers(-512, 512, (R.NS, R.P)) + 1j * rng.integers(-512, 512, (R.NS, R.P))).astype(np.complex64)
    cd2, asf2, dslip = [0.1, 0.1], ph.asf2_reference(2), [1.0, 1.0]
    tau = np.array([0.01, -0.003])
    xr, gr, g1r = R.plasticevolv(xi, bc, d2, cd2, asf2, np.tile(tau, (R.P, 1)), dslip)
    xo, go, g1o = O.plasticevolv(xi[:R.NS], d2[:R.NS].astype(np.complex128), cd2, asf2, dslip, tau,
                                 xi_bc=bc[:R.NS])
    assert np.array_equal(xo, xr[:R.NS])
    assert abs(go - gr) <= 1e-13 * abs(gr)
    assert np.abs(g1o - g1r).max() <= 1e-13 * np.abs(g1r).max()
    assert np.array_equal(xo[0, ::7], xi[0, ::7])      # frozen points untouched


def test_uniform_xi_is_the_scalar_recurrence():
    """Uniform xi excites only k=0 where B(0)=0: each step is
    xi <- xi - CD2 (Asf2 pi sin(2 pi xi) - tau/dslip); from 0 the first step is CD2 tau."""
    n = (8, 6, 4)
    eps = ph.fcc_eps([0, 1, 2])
    p = O.make_params(n, eps, ph.C11, ph.C12, ph.C44, [0.1] * 3, ph.asf2_reference(3), [1.0] * 3)
    o = O.Oracle(p)
    x0 = np.array([0.0, 0.37, 0.81])
    o.set_xi(np.ones((3,) + n) * x0[:, None, None, None])
    tau = np.array([0.01, 0.02, -0.005])
    o.set_tau(tau)
    o.prologue(True)
    g, ga = o.gl_steps(20)
    x = x0.copy()
    asf = np.array(ph.asf2_reference(3))
    for s in range(20):
        x = x - 0.1 * (asf * ph.PI_REF * np.sin(2.0 * ph.PI_REF * x) - tau)
        if s == 0:
            assert ga[0, 0] == pytest.approx(0.1 * tau[0], abs=1e-15)
        assert np.abs(ga[s] - x).max() < 1e-13
    assert np.abs(o.get_xi().real - x[:, None, None, None]).max() < 1e-13
    assert np.abs(o.get_xi().imag).max() < 1e-14


def test_skipping_virtual_fields_is_exact():
    """With xi_v = 0 the virtual columns of B multiply FFT(0) = 0 exactly."""
    FX, FY, FZ = O.kgrid(6, 4, 8)
    ks = np.stack([FX.reshape(-1), FY.reshape(-1), FZ.reshape(-1)], 1)
    eps = ph.fcc_eps([0, 4])
    B = O.bmatrix_at(ks, eps, np.concatenate([eps, R.epsv_basis()]), ph.C11, ph.C12, ph.C44)
    assert np.isfinite(B).all()          # finite coefficients times exact zeros add exact zeros


@pytest.mark.skipif(not R.available(), reason="compiled reference oracle/_ref not built")
@pytest.mark.parametrize("quirk", [True, False])
def test_gl_loop_matches_reference_functions(quirk):
    """The GL loop of main driven through the reference's compiled fourn, Bmatrix,
    Fmatrix, virtualevolv and plasticevolv (32x2x32, NS=2, FP32 arrays) agrees
    with the FP64 oracle to FP32 rounding when the virtualevolv term is kept,
    and visibly disagrees when the oracle drops it."""
    th = ph.theta_nz()
    eps = ph.fcc_eps([0, 1], th)
    rng = np.random.default_rng(5)
    xi0 = np.zeros((2, R.P), dtype=np.complex128)
    X = np.zeros((32, 2, 32))
    X[8:20, :, 10:22] = 1.0
    xi0[0] = X.reshape(-1)
    xi0[1] = 0.3 * rng.random(R.P)
    cd2, asf2, dslip, tau = [0.1, 0.1], ph.asf2_reference(2), [1.0, 1.0], [0.01, -0.004]
    xr, gr = R.gl_loop(xi0, eps, R.epsv_basis(), ph.C11, ph.C12, ph.C44, cd2, asf2, dslip, tau, 10)
    p = O.make_params((32, 2, 32), eps, ph.C11, ph.C12, ph.C44, cd2, asf2, dslip)
    o = O.Oracle(p)
    o.set_xi(xi0.reshape(2, 32, 2, 32))
    o.set_tau(np.array(tau))
    o.prologue(quirk)
    g, _ = o.gl_steps(10)
    err = np.abs(o.get_xi().reshape(2, -1) - xr).max()
    if quirk:
        assert err < 1e-6
        assert np.abs(g - gr).max() < 1e-8
    else:
        assert err > 1e-2


def test_tau_ramp_uniform_closed_form():
    """Uniform xi with a linear tau ramp: step s uses tau0 + rate*s."""
    n = (8, 6, 4)
    eps = ph.fcc_eps([0, 1, 2])
    p = O.make_params(n, eps, ph.C11, ph.C12, ph.C44, [0.1] * 3, ph.asf2_reference(3), [1.0] * 3)
    o = O.Oracle(p)
    x0 = np.array([0.1, 0.5, 0.9])
    o.set_xi(np.ones((3,) + n) * x0[:, None, None, None])
    tau0, rate = np.array([0.0, 0.01, -0.02]), np.array([1e-3, -2e-4, 5e-4])
    o.set_tau_ramp(tau0, rate)
    o.prologue(True)
    _, ga = o.gl_steps(15)
    x = x0.copy()
    asf = np.array(ph.asf2_reference(3))
    for s in range(15):
        x = x - 0.1 * (asf * ph.PI_REF * np.sin(2.0 * ph.PI_REF * x) - (tau0 + rate * s))
        assert np.abs(ga[s] - x).max() < 1e-13
