"""The oracle's energy diagnostics on the interface line against the compiled
reference (avestrain G.c:641-749, ResidualEnergy G.c:3920-3947) at its
compile-time grid 32x2x32, NS=2, and the line rule against brute force; and
that the reference's Energy_calculation (G.c:3026-3467) reads only the
virtual-strain fields, so it is 0 on the GL path."""
import itertools

import numpy as np
import pytest

import oraclepy as O
import refpy as R

NORMAL = (np.sqrt(3) / 3.0, 0.0, np.sqrt(6) / 3.0)   # interface_n of main (G.c:4234-4236)


def s_consts(C11, C12, C44):
    S11 = 1. / 3. * (1 / (C11 + 2 * C12) + 2 / (C11 - C12))
    S12 = 1. / 3. * (1 / (C11 + 2 * C12) - 1 / (C11 - C12))
    return S11, S12, 1 / C44


@pytest.mark.parametrize("shape,pp", [((32, 2, 32), (16, 0, 24)), ((16, 12, 8), (3, 5, 6)), ((8, 8, 8), (0, 0, 0))])
def test_line_points_brute_force(shape, pp):
    want = [((k1 * shape[1] + k2) * shape[2] + k3)
            for k1, k2, k3 in itertools.product(range(shape[0]), range(shape[1]), range(shape[2]))
            if k1 == pp[0] and abs(NORMAL[0] * (k1 - pp[0]) + NORMAL[1] * (k2 - pp[1]) + NORMAL[2] * (k3 - pp[2])) <= 1.0]
    got = O.line_points(shape, NORMAL, pp)
    assert list(got) == want and len(want) > 0


def _state(rng):
    C11, C12, C44 = 2.0 + 2 * 0.49, 0.98, 1.0
    eps = rng.standard_normal((R.NS, 3, 3)) * 0.3
    p = O.make_params((R.N1, R.N2, R.N3), eps, C11, C12, C44, [0.1] * R.NS, [0.01] * R.NS, [1.0] * R.NS)
    o = O.Oracle(p)
    xi = rng.random((R.NS, R.N1, R.N2, R.N3)) + 1j * 1e-3 * rng.standard_normal((R.NS, R.N1, R.N2, R.N3))
    o.set_xi(xi)
    return o, xi, eps, (C11, C12, C44)


@pytest.mark.skipif(not R.available(), reason="compiled reference oracle/_ref not built")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_line_energy_vs_reference_avestrain(seed):
    rng = np.random.default_rng(seed)
    o, xi, eps, (C11, C12, C44) = _state(rng)
    s = rng.uniform(-0.02, 0.02, (3, 3))
    sigma = 0.5 * (s + s.T)
    S11, S12, S44 = s_consts(C11, C12, C44)
    mu = C44
    border = R.N3 * 3 // 4
    pp = (R.N1 // 2, 0, border)
    xfull = np.zeros((R.NSV, R.N1, R.N2, R.N3), dtype=np.complex128)
    xfull[:R.NS] = xi
    e3_ref, e4_ref = R.avestrain(xfull, eps, sigma, S11, S12, S44, mu, border, NORMAL, pp)
    e3, e4 = o.line_energy(sigma, S11, S12, S44, mu, O.line_points((R.N1, R.N2, R.N3), NORMAL, pp))
    assert abs(e3 - e3_ref) <= 1e-6 * max(1.0, abs(e3_ref)) and abs(e3) > 1e-4
    assert abs(e4 - e4_ref) <= 1e-6 * max(1.0, abs(e4_ref)) and abs(e4) > 1e-6


@pytest.mark.skipif(not R.available(), reason="compiled reference oracle/_ref not built")
def test_residual_energy_vs_reference():
    rng = np.random.default_rng(7)
    o, xi, eps, _ = _state(rng)
    pp = (R.N1 // 2, 0, R.N3 * 3 // 4)
    xfull = np.zeros((R.NSV, R.N1, R.N2, R.N3), dtype=np.complex128)
    xfull[:R.NS] = xi
    for D in ((0.3, -0.2, 0.1, 0.7), (0.0, 0.0, 0.0, 0.0), (1.0, 0.0, 0.0, -1.0)):
        ref = R.residual_energy(xfull, NORMAL, pp, *D)
        got = o.residual_energy(pp, 1, *D)
        assert abs(got - ref) <= 1e-6 * max(1.0, abs(ref))
    assert o.residual_energy(pp, 1, 0.0, 0.0, 0.0, 0.0) == 0.0


@pytest.mark.skipif(not R.available(), reason="compiled reference oracle/_ref not built")
def test_energy_calculation_is_zero_without_virtual_strain():
    rng = np.random.default_rng(11)
    C11, C12, C44 = 2.0 + 2 * 0.49, 0.98, 1.0
    eps = rng.standard_normal((R.NS, 3, 3)) * 0.3
    pp = (R.N1 // 2, 0, R.N3 * 3 // 4)
    data = np.zeros((R.NSV, R.N1, R.N2, R.N3), dtype=np.complex128)
    data[:R.NS] = rng.standard_normal((R.NS, R.N1, R.N2, R.N3)) + 1j * rng.standard_normal((R.NS, R.N1, R.N2, R.N3))
    assert R.energy_calculation(data, eps, C11, C12, C44, NORMAL, pp) == 0.0
    data[R.NS + 2] = rng.standard_normal((R.N1, R.N2, R.N3)) + 1j * rng.standard_normal((R.N1, R.N2, R.N3))
    assert R.energy_calculation(data, eps, C11, C12, C44, NORMAL, pp) != 0.0
