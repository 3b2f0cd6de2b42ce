"""Pins the oracle's strain/stress -- row a11 (strain G.c:752-835, stress
G.c:837-972, averages G.c:4058-4070)."""
import numpy as np
import pytest

import oraclepy as O
import physics as ph
import refpy as R


def s_tensor():
    """S_ijkl of avestrain (G.c:656-667) from C11, C12, C44 (G.c:4261-4263)."""
    C11, C12, C44 = ph.C11, ph.C12, ph.C44
    S11 = 1. / 3. * (1 / (C11 + 2 * C12) + 2 / (C11 - C12))
    S12 = 1. / 3. * (1 / (C11 + 2 * C12) - 1 / (C11 - C12))
    S44 = 1 / C44
    d = np.eye(3)
    S = np.zeros((3, 3, 3, 3))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                for l in range(3):
                    d4 = 1.0 if i == j == k == l else 0.0
                    S[i, j, k, l] = ((S11 + 2 * S12) / 3.0 * d[i, j] * d[k, l]
                                     + S44 / 4.0 * (d[i, k] * d[j, l] + d[i, l] * d[j, k])
                                     - S44 / 2.0 * d4 + (S11 - S12) * (d4 - d[i, j] * d[k, l] / 3.0))
    return S


def test_mean_stress_is_applied_stress():
    """Periodic cell: <sigma> = sigma_applied * mu whatever xi is."""
    n = (8, 8, 8)
    eps = ph.fcc_eps([0, 5, 9])
    p = O.make_params(n, eps, ph.C11, ph.C12, ph.C44, [0.1] * 3, ph.asf2_reference(3), [1.0] * 3)
    o = O.Oracle(p)
    rng = np.random.default_rng(2)
    o.set_xi(rng.random((3,) + n))
    sig = rng.uniform(-0.02, 0.02, (3, 3))
    sig = 0.5 * (sig + sig.T)
    avepst = np.einsum("ijkl,kl->ij", s_tensor(), sig) * ph.MU
    aveps, avesig = o.strain_stress(avepst)
    assert np.abs(avesig - sig * ph.MU).max() < 1e-13
    xim = o.get_xi().real.reshape(3, -1).mean(1)
    want_eps = avepst + 0.5 * np.einsum("aij,a->ij", eps + np.transpose(eps, (0, 2, 1)), xim)
    assert np.abs(aveps - want_eps).max() < 1e-13


@pytest.mark.skipif(not R.available(), reason="compiled reference oracle/_ref not built")
def test_strain_stress_match_reference():
    rng = np.random.default_rng(8)
    eps = ph.fcc_eps([0, 1], ph.theta_nz())
    epsv = R.epsv_basis()
    xi = np.zeros((R.NSV, R.P), dtype=np.complex128)
    X = np.zeros((32, 2, 32))
    X[4:18, :, 6:25] = 1.0
    xi[0] = X.reshape(-1)
    xi[1] = 0.5 * rng.random(R.P)
    sig = np.array([[0.0, 0.01, 0.0], [0.01, 0.0, -0.004], [0.0, -0.004, 0.005]])
    avepst_u = np.einsum("ijkl,kl->ij", s_tensor(), sig) * ph.MU
    avepsd = np.einsum("aij,a->ij", eps, xi[:2].real.mean(1))
    data = xi.astype(np.complex64)
    for b in range(R.NSV):
        data[b] = R.fourn3d(data[b].reshape(32, 2, 32), -1).reshape(-1)
    FF, _ = R.fmatrix(eps, epsv, ph.C11, ph.C12, ph.C44)
    ap = np.tile((avepst_u + avepsd).reshape(1, 3, 3), (R.P, 1, 1))
    deps, dsig, avesig = R.strain_stress(data, FF, xi, eps, epsv, ap, ph.C11, ph.C12, ph.C44)
    p = O.make_params((32, 2, 32), eps, ph.C11, ph.C12, ph.C44, [0.1] * 2, ph.asf2_reference(2), [1.0] * 2)
    o = O.Oracle(p)
    o.set_xi(xi[:2].reshape(2, 32, 2, 32))
    aveps, avesig_o, ef, sf = o.strain_stress(avepst_u, fields=True)
    # reference slot i + 3j holds component (i, j)
    ref_e = np.array([[deps[i + 3 * j].real for j in range(3)] for i in range(3)]).reshape(3, 3, 32, 2, 32)
    ref_s = np.array([[dsig[i + 3 * j].real for j in range(3)] for i in range(3)]).reshape(3, 3, 32, 2, 32)
    assert np.abs(ef - ref_e).max() < 2e-6
    assert np.abs(sf - ref_s).max() < 2e-5
    assert np.abs(avesig_o - avesig).max() < 1e-6
