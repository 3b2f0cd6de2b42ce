"""Pins the oracle FFT (orc_fft3d) -- row a1 of the hot path (G.c:209-273)."""
import os

import numpy as np
import pytest

import oraclepy as O
import refpy as R

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fft_golden.npz")


def brute_dft3(x, sign):
    out = x.astype(np.complex128)
    for ax in range(3):
        n = out.shape[ax]
        k = np.arange(n)
        W = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
        out = np.moveaxis(np.tensordot(W, np.moveaxis(out, ax, 0), axes=(1, 0)), 0, ax)
    return out


@pytest.mark.parametrize("shape", [(4, 4, 4), (6, 3, 2), (5, 4, 6), (8, 1, 3), (2, 2, 2), (1, 1, 1), (9, 4, 10)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_vs_brute_force(shape, sign):
    rng = np.random.default_rng(sum(shape) + sign)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    got = O.fft3d(x, sign)
    want = brute_dft3(x, sign)
    assert np.abs(got - want).max() <= 1e-13 * max(1.0, np.abs(want).max())


def test_fft_vs_numpy_golden():
    g = np.load(GOLD)
    for i, s in enumerate(g["shapes"]):
        x = g[f"x{i}"]
        for sign, key in ((-1, "fwd"), (1, "bwd")):
            want = g[f"{key}{i}"]
            got = O.fft3d(x, sign)
            assert np.abs(got - want).max() <= 2e-14 * np.abs(want).max(), (tuple(s), sign)


def test_fft_closed_forms():
    n = (16, 12, 10)
    P = np.prod(n)
    # delta -> constant
    d = np.zeros(n, dtype=complex)
    d[0, 0, 0] = 1.0
    assert np.abs(O.fft3d(d, -1) - 1.0).max() < 1e-15
    # plane wave exp(+2 pi i k0.x/N) -> P * delta(k0) under the forward sign
    k0 = (3, 5, 7)
    i, j, k = np.meshgrid(*[np.arange(m) for m in n], indexing="ij")
    pw = np.exp(2j * np.pi * (k0[0] * i / n[0] + k0[1] * j / n[1] + k0[2] * k / n[2]))
    F = O.fft3d(pw, -1)
    want = np.zeros(n, dtype=complex)
    want[k0] = P
    assert np.abs(F - want).max() < 1e-10
    # round trip and Parseval
    rng = np.random.default_rng(3)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    X = O.fft3d(x, -1)
    assert np.abs(O.fft3d(X, 1) / P - x).max() < 1e-14
    assert abs(np.sum(np.abs(X) ** 2) / P - np.sum(np.abs(x) ** 2)) < 1e-10 * np.sum(np.abs(x) ** 2)


def test_fft_batch_is_independent():
    rng = np.random.default_rng(11)
    x = rng.standard_normal((3, 6, 4, 8)) + 1j * rng.standard_normal((3, 6, 4, 8))
    got = O.fft3d(x, -1)
    for b in range(3):
        assert np.array_equal(got[b], O.fft3d(x[b], -1))


def test_fft_rejects_bad_size():
    with pytest.raises(ValueError):
        O.fft3d(np.zeros((7, 4, 4), dtype=complex), -1)


@pytest.mark.skipif(not R.available(), reason="compiled reference oracle/_ref not built")
@pytest.mark.parametrize("shape", [(32, 2, 32), (16, 8, 4), (64, 4, 2)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_vs_reference_fourn(shape, sign):
    """The reference's own FP32 fourn (G.c:2758) agrees to FP32 rounding."""
    rng = np.random.default_rng(7)
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    ref = R.fourn3d(x, sign)
    got = O.fft3d(x.astype(np.complex128), sign)
    assert np.abs(ref - got).max() <= 2e-6 * np.abs(got).max()
