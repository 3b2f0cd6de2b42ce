"""The allocation-free host seam call pfdd_fft_host_noalloc against the
allocating seam fft_forward / fft_backward (G.c:209-273), through the C ABI:
the same bits on two calls in a row (the second with a smaller batch), on a
second grid, and at the size limit of its fixed staging array."""
import sys

import numpy as np
import pytest

import pfdd_amd as pa
from pfdd_amd import hip

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-noalloc] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if hip.device_count() < 1:
        pytest.skip("no GPU visible")


def _field(rng, batch, shape):
    return (rng.standard_normal((batch,) + shape) + 1j * rng.standard_normal((batch,) + shape)).astype(np.complex64)


def _allocating(L, x, batch, direction):
    a = x.copy()
    (L.fft_forward if direction == pa.FORWARD else L.fft_backward)(a.ctypes.data, batch)
    assert L.pfdd_fft_host_status() == pa.OK
    return a


def test_noalloc_seam_equals_allocating_seam():
    L = pa.load()
    rng = np.random.default_rng(20261203)
    try:
        for shape, batches in (((32, 16, 64), (5, 2)), ((12, 8, 6), (3, 1))):
            assert L.pfdd_fft_host_init(*shape) == pa.OK
            for batch in batches:  # two calls in a row, the second with a smaller input
                for direction in (pa.FORWARD, pa.BACKWARD):
                    say(f"shape={shape} batch={batch} dir={direction}")
                    x = _field(rng, batch, shape)
                    want = _allocating(L, x, batch, direction)
                    got = x.copy()
                    assert L.pfdd_fft_host_noalloc(got.ctypes.data, batch, direction) == pa.OK
                    assert np.array_equal(got, want)
                    assert not np.array_equal(got, x)
    finally:
        L.pfdd_fft_host_release()


def test_noalloc_seam_size_limit():
    L = pa.load()
    rng = np.random.default_rng(5)
    shape = (128, 128, 128)  # one field fills the staging array exactly
    assert np.prod(shape) == pa.NOALLOC_MAX_POINTS
    try:
        assert L.pfdd_fft_host_init(*shape) == pa.OK
        x = _field(rng, 1, shape)
        say("128^3 batch 1 (the limit)")
        want = _allocating(L, x, 1, pa.FORWARD)
        got = x.copy()
        assert L.pfdd_fft_host_noalloc(got.ctypes.data, 1, pa.FORWARD) == pa.OK
        assert np.array_equal(got, want)
        big = np.zeros((2,) + shape, dtype=np.complex64)
        assert L.pfdd_fft_host_noalloc(big.ctypes.data, 2, pa.FORWARD) == pa.E_LIMIT
        assert L.pfdd_fft_host_noalloc(got.ctypes.data, 1, 0) == pa.E_SHAPE
    finally:
        L.pfdd_fft_host_release()
