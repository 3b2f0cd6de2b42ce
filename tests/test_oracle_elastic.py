"""Pins the oracle's interaction matrices -- rows a2/a3 (frec, Bmatrix
G.c:328-412) and the DD/FF matrices of Fmatrix (G.c:414-598)."""
import numpy as np
import pytest

import oraclepy as O
import refpy as R

NU = 0.33
LAM = 2 * NU / (1 - 2 * NU)
ISO = (LAM + 2.0, LAM, 1.0)          # C11, C12, C44 with mu = 1
CUBIC = (1.684, 1.214, 0.235 * 3)    # an anisotropic (cubic) set, mu' != 0


def rand_eps(rng, n):
    e = []
    for _ in range(n):
        b = rng.standard_normal(3)
        m = rng.standard_normal(3)
        e.append(np.outer(b, m))
    return np.array(e)


def cij(C11, C12, C44):
    C = np.zeros(81)
    O.lib().orc_cijkl(C.ctypes.data, C11, C12, C44)
    return C.reshape(3, 3, 3, 3)


def textbook_B(k, ea, eb, C11, C12, C44):
    """B_ab(k) = [e_a:C:e_b - (s_a n).K(n)^-1.(s_b n)]/mu with the acoustic tensor
    K_ik = C_ijkl n_j n_l (Khachaturyan's microelasticity kernel)."""
    C = cij(C11, C12, C44)
    k = np.asarray(k, dtype=float)
    if np.dot(k, k) == 0:
        return np.zeros((len(ea), len(eb)))
    n = k / np.linalg.norm(k)
    K = np.einsum("ijkl,j,l->ik", C, n, n)
    Om = np.linalg.inv(K)
    sa = np.einsum("ijkl,akl->aij", C, ea)
    sb = np.einsum("ijkl,akl->aij", C, eb)
    ua = sa @ n
    ub = sb @ n
    return (np.einsum("aij,bij->ab", sa, eb) - ua @ Om @ ub.T) / C44


@pytest.mark.parametrize("mat", [ISO, CUBIC])
def test_factored_equals_literal_loop_nest(mat):
    rng = np.random.default_rng(1)
    ks = rng.standard_normal((40, 3)) * 0.3
    ks[0] = 0.0
    ea, eb = rand_eps(rng, 4), rand_eps(rng, 5)
    lit = O.bmatrix_at(ks, ea, eb, *mat, literal=True)
    fac = O.bmatrix_at(ks, ea, eb, *mat)
    assert np.abs(lit - fac).max() <= 1e-13 * np.abs(lit).max()
    assert np.all(lit[0] == 0) and np.all(fac[0] == 0)


@pytest.mark.parametrize("mat", [ISO, CUBIC])
def test_B_matches_acoustic_tensor_closed_form(mat):
    rng = np.random.default_rng(2)
    ea = rand_eps(rng, 3)
    for _ in range(25):
        k = rng.standard_normal(3)
        got = O.bmatrix_at(k, ea, ea, *mat)[0]
        want = textbook_B(k, ea, ea, *mat)
        assert np.abs(got - want).max() <= 1e-12 * max(1, np.abs(want).max())


def test_B_invariants():
    rng = np.random.default_rng(3)
    ea = rand_eps(rng, 4)
    ks = rng.standard_normal((30, 3))
    B = O.bmatrix_at(ks, ea, ea, *ISO)
    Bm = O.bmatrix_at(-ks, ea, ea, *ISO)
    B5 = O.bmatrix_at(5.0 * ks, ea, ea, *ISO)
    assert np.abs(B - Bm).max() < 1e-13
    assert np.abs(B - B5).max() < 1e-12
    assert np.abs(B - np.transpose(B, (0, 2, 1))).max() < 1e-13
    # positive semi-definite on symmetric eigenstrains (elastic energy >= 0)
    es = np.array([0.5 * (e + e.T) for e in ea])
    Bs = O.bmatrix_at(ks, es, es, *ISO)
    assert np.linalg.eigvalsh(Bs).min() > -1e-12


def test_frec_matches_reference():
    if not R.available():
        pytest.skip("compiled reference not built")
    fx, fy, fz = R.frec()
    FX, FY, FZ = O.kgrid(R.N1, R.N2, R.N3)
    assert np.array_equal(fx, FX.reshape(-1))
    assert np.array_equal(fy, FY.reshape(-1))
    assert np.array_equal(fz, FZ.reshape(-1))


@pytest.mark.skipif(not R.available(), reason="compiled reference oracle/_ref not built")
def test_B_matches_reference_Bmatrix():
    rng = np.random.default_rng(4)
    eps = rand_eps(rng, R.NS)
    epsv = R.epsv_basis()
    BB = R.bmatrix(eps, epsv, *ISO)                      # [NSV][NS][P]
    FX, FY, FZ = O.kgrid(R.N1, R.N2, R.N3)
    ks = np.stack([FX.reshape(-1), FY.reshape(-1), FZ.reshape(-1)], 1)
    eb = np.concatenate([eps, epsv])
    mine = O.bmatrix_at(ks, eps, eb, *ISO)               # [P][NS][NSV]
    want = np.transpose(BB, (2, 1, 0))
    assert np.abs(mine - want).max() <= 1e-13 * np.abs(want).max()


@pytest.mark.skipif(not R.available(), reason="compiled reference oracle/_ref not built")
def test_DD_and_FF_match_reference_Fmatrix():
    rng = np.random.default_rng(5)
    eps = rand_eps(rng, R.NS)
    epsv = R.epsv_basis()
    FF, DD = R.fmatrix(eps, epsv, *ISO)                  # FF [j][i][a][P], DD [j][i][b][P]
    FX, FY, FZ = O.kgrid(R.N1, R.N2, R.N3)
    ks = np.stack([FX.reshape(-1), FY.reshape(-1), FZ.reshape(-1)], 1)
    eb = np.concatenate([eps, epsv])
    dds = O.ddsum_at(ks, eb, *ISO)                       # [P][NSV]
    want = DD.sum(axis=(0, 1)).T
    # Fmatrix accumulates D, A and G in float (G.c:423-430): FP32 tolerance
    assert np.abs(dds - want).max() <= 2e-6 * np.abs(want).max()
    ff = O.fmatrix_at(ks, eps, *ISO)                     # [P][a][i][j]
    want_ff = np.transpose(FF, (3, 2, 1, 0))
    assert np.abs(ff - want_ff).max() <= 2e-6 * np.abs(want_ff).max()
