"""PFDD_REF32: the reference's own float arithmetic on the GPU, checked
against the reference's compiled functions (oracle/_ref, built from
PFDD_Final_GPU.c with its FOURN backend): fourn bit for bit, and the GL loop
of main (G.c:4450-4452, virtualevolv 1426-1447, 4574-4644) at the
reference's compile-time grid 32x2x32, NS = 2, far closer than the FP64 path
can come to it; plus the c0 trajectory against the FP64 golden run."""
import os
import sys

import numpy as np
import pytest

import physics as ph
import pfdd_amd as pa
import refpy as R
from pfdd_amd import configs as C
from pfdd_amd import hip

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if hip.device_count() < 1:
        pytest.skip("no GPU visible")


def say(msg):
    print(f"[gpu-ref32] {msg}", flush=True)
    sys.stdout.flush()


def fft_workload(shape):
    return C.Workload("fft", shape, ph.fcc_eps([0]), np.zeros((1, 3)), np.zeros((1,) + shape, complex), np.zeros(1),
                      np.zeros((3, 3)), 0)


@pytest.mark.skipif(not R.available(), reason="compiled reference oracle/_ref not built")
@pytest.mark.parametrize("shape", [(32, 2, 32), (16, 8, 4), (64, 16, 32), (8, 4, 128), (2, 256, 2), (128, 64, 8)])
def test_ref32_fft_is_fourn_bit_for_bit(shape):
    """pfdd_fft with PFDD_REF32 returns the very floats of the compiled fourn."""
    say(f"fourn {shape}")
    rng = np.random.default_rng(sum(shape) + 11)
    x = (rng.standard_normal((2,) + shape) + 1j * rng.standard_normal((2,) + shape)).astype(np.complex64)
    with pa.Context(fft_workload(shape).params(pa.REF32)) as c:
        for sign in (pa.FORWARD, pa.BACKWARD):
            buf = hip.DeviceBuffer.from_array(x)
            c.fft(buf.ptr, 2, sign, pa.REF32)
            c.sync()
            got = buf.to_array(np.complex64, x.shape)
            for b in range(2):
                ref = R.fourn3d(x[b], sign)
                assert np.array_equal(got[b].view(np.uint32), ref.view(np.uint32)), (shape, sign, b)


def reference_grid_case():
    eps = ph.fcc_eps([0, 1], ph.theta_nz())
    rng = np.random.default_rng(5)
    xi0 = np.zeros((2, R.P), dtype=np.complex128)
    X = np.zeros((32, 2, 32))
    X[8:20, :, 10:22] = 1.0
    xi0[0] = X.reshape(-1)
    xi0[1] = 0.3 * rng.random(R.P)
    tau = np.array([0.01, -0.004])
    w = C.Workload("ref", (32, 2, 32), eps, np.zeros((2, 3)), xi0.reshape(2, 32, 2, 32), tau, np.zeros((3, 3)), 10)
    return eps, xi0, tau, w


def run_gpu(w, steps, precision, quirk=True, chunks=None):
    with pa.Context(w.params(precision)) as c:
        c.set_tau(w.tau)
        c.upload_xi(w.xi)
        c.begin_outer(quirk)
        if chunks is None:
            g, _ = c.gl_steps(steps)
        else:
            g = np.concatenate([c.gl_steps(k)[0] for k in chunks])
        return c.download_xi(), g


@pytest.mark.skipif(not R.available(), reason="compiled reference oracle/_ref not built")
def test_ref32_gl_loop_vs_compiled_reference():
    """10 GL steps after the prologue (with virtualevolv's term), against the
    loop run through the reference's compiled fourn, Bmatrix, Fmatrix,
    virtualevolv and plasticevolv.  The FP64 path differs from it by the
    reference's float rounding; PFDD_REF32 reproduces that rounding, all but
    the last bits of the double sine."""
    eps, xi0, tau, w = reference_grid_case()
    xr, gr = R.gl_loop(xi0, eps, R.epsv_basis(), ph.C11, ph.C12, ph.C44, [0.1, 0.1], ph.asf2_reference(2),
                       [1.0, 1.0], tau, 10)
    x32, g32 = run_gpu(w, 10, pa.REF32)
    x64, g64 = run_gpu(w, 10, pa.FP64)
    e32 = np.abs(x32.reshape(2, -1) - xr).max()
    e64 = np.abs(x64.reshape(2, -1) - xr).max()
    say(f"max|dxi| vs compiled reference: REF32 {e32:.3e}, FP64 {e64:.3e}; gamma REF32 {np.abs(g32 - gr).max():.3e}")
    assert e32 <= 1e-14          # measured 2.8e-17: the GPU's double sine against glibc's, last bits
    assert e32 * 1e6 <= e64      # FP64: 2.4e-8
    assert np.abs(g32 - gr).max() <= 1e-12


def test_ref32_graph_chunks_equal_single_steps():
    """Graph-replayed chunks and single eager steps give the same bits."""
    _, _, _, w = reference_grid_case()
    xa, ga = run_gpu(w, 20, pa.REF32)
    xb, gb = run_gpu(w, 20, pa.REF32, chunks=[1] * 5 + [8, 3, 4])
    assert np.array_equal(xa, xb)
    assert np.array_equal(ga, gb)


def test_ref32_c0_against_fp64_golden():
    """c0 (32^3 x 3, 500 steps) in the reference's float arithmetic against
    the FP64 oracle's golden trajectory: within the FP32 bounds of the mixed
    modes (xi 1e-4, slipped area 1e-3 relative)."""
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gl_c0.npz"))
    w = C.c0()
    x32, g32 = run_gpu(w, int(g["steps"]), pa.REF32)
    xo = g["xi"]
    err = np.abs(x32 - xo).max()
    area_o = float(np.sum(xo.real))
    area = float(np.sum(x32.real))
    say(f"c0 REF32 vs FP64 golden: max|dxi| {err:.3e}, slipped area rel {abs(area - area_o) / abs(area_o):.3e}")
    assert err <= 1e-4
    assert abs(area - area_o) <= 1e-3 * abs(area_o)


def test_ref32_rejects_what_it_does_not_run():
    w = C.small_fcc((32, 32, 32), ns=3)
    with pytest.raises(pa.PfddError) as ei:
        with pa.Context(w.params(pa.REF32, pa.B_RECOMPUTE, pa.PLAN_INTERLEAVED)):
            pass
    assert ei.value.rc == pa.E_SHAPE
    with pa.Context(w.params(pa.REF32)) as c:
        c.set_tau(w.tau)
        c.upload_xi(w.xi)
        with pytest.raises(pa.PfddError) as ei:
            c.autotune(2)
        assert ei.value.rc == pa.E_UNSUPPORTED
    with pytest.raises(pa.PfddError) as ei:
        with pa.Ensemble([w.params(pa.REF32)]):
            pass
    assert ei.value.rc == pa.E_UNSUPPORTED
