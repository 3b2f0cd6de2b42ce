"""GPU tests of PFDD_MIXED_LEAN: mixed precision with the master xi kept as a
plain FP32 complex (8 bytes a point), through the C ABI.  The update still
runs in FP64; xi is rounded to FP32 after every update, so these tests bound
its error against the FP64 path and against the compensated mixed mode, and
pin the storage form itself."""
import sys

import numpy as np
import pytest

import pfdd_amd as pa
from pfdd_amd import configs as C
from pfdd_amd import hip

pytestmark = pytest.mark.gpu
LEAN = pa.MIXED_LEAN


def say(msg):
    print(f"[gpu-lean] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if hip.device_count() < 1:
        pytest.skip("no GPU visible")


def run_gpu(w, steps, precision, bmode=pa.B_RECOMPUTE, plan=0, xi_bc=None):
    with pa.Context(w.params(precision, bmode, plan)) as c:
        c.set_tau(w.tau)
        c.upload_xi(w.xi, xi_bc)
        c.begin_outer(True)
        g, ga = c.gl_steps(steps)
        return c.download_xi(), g, ga


def f32(z):
    return z.real.astype(np.float32).astype(np.float64) + 1j * z.imag.astype(np.float32).astype(np.float64)


def test_lean_master_is_plain_fp32():
    """Upload rounds both parts of xi to FP32 and download returns them; after
    400 steps on a uniform field (exact FP32 transforms, B(0) = 0) every xi is
    an FP32 value that follows the FP64 scalar recurrence to within the
    accumulated FP32 rounding (400 x 2^-24 relative)."""
    say("lean master round trip + uniform 32^3 x 12")
    w = C.small_fcc((32, 32, 32), ns=12)
    rng = np.random.default_rng(11)
    z = rng.standard_normal((w.ns,) + w.shape) + 1j * 1e-3 * rng.standard_normal((w.ns,) + w.shape)
    with pa.Context(w.params(LEAN)) as c:
        c.upload_xi(z)
        back = c.download_xi()
    assert np.array_equal(back, f32(z))

    x0 = np.linspace(0.05, 0.95, w.ns) + 1e-9 * np.arange(w.ns)
    w.xi = np.ascontiguousarray(np.ones((w.ns,) + w.shape, dtype=np.complex128) * x0[:, None, None, None])
    xg, _, gag = run_gpu(w, 400, LEAN)
    x = x0.copy()
    asf = np.array(C.asf2_reference(w.ns))
    for _ in range(400):
        x = x - 0.1 * (asf * C.PI_REF * np.sin(2.0 * C.PI_REF * x) - w.tau)
    assert np.array_equal(xg, f32(xg))
    assert np.abs(xg.real - x[:, None, None, None]).max() <= 400 * 2.0 ** -24
    assert np.abs(gag[-1] - x).max() <= 400 * 2.0 ** -24
    assert np.abs(xg.imag).max() == 0.0


def test_lean_error_c1_1000_steps():
    """c1 (128^3 x 12) for 1000 steps: the lean mode against FP64 within the
    mixed-precision bounds (xi <= 1e-4, slipped area <= 1e-3 relative)."""
    say("c1 lean 1000 steps vs fp64")
    w = C.c1()
    x64, _, ga64 = run_gpu(w, 1000, pa.FP64)
    xl, _, gal = run_gpu(w, 1000, LEAN)
    err = np.abs(xl - x64).max()
    a64, al = ga64[-1].sum(), gal[-1].sum()
    say(f"lean: max|dxi| = {err:.3e}, slipped area rel = {abs(al - a64) / abs(a64):.3e}")
    assert err <= 1e-4
    assert abs(al - a64) <= 1e-3 * abs(a64)


@pytest.mark.parametrize("bmode", [pa.B_RECOMPUTE, pa.B_STORED])
def test_lean_close_to_compensated(bmode):
    """Over 200 steps at 32^3 x 12 the two mixed masters differ by the FP32
    rounding of xi only."""
    say(f"lean vs compensated bmode={bmode}")
    w = C.small_fcc((32, 32, 32), ns=12)
    xm, _, gam = run_gpu(w, 200, pa.MIXED, bmode)
    xl, _, gal = run_gpu(w, 200, LEAN, bmode)
    assert np.abs(xl - xm).max() <= 1e-5
    assert np.abs(gal - gam).max() <= 1e-6


@pytest.mark.parametrize("plan", [pa.PLAN_INTERLEAVED, pa.PLAN_INTERLEAVED | pa.PLAN_UNFUSED_COLS,
                                  pa.PLAN_PLANE | pa.PLAN_FUSED_COLS, pa.PLAN_NO_PLANE | pa.PLAN_UNFUSED_ROWS])
def test_lean_plans_agree(plan):
    """The lean master through the interleaved, plane and unfused kernels."""
    say(f"lean plan {pa.describe_plan(plan)}")
    w = C.small_fcc((32, 32, 32), ns=12)
    x0, _, _ = run_gpu(w, 20, LEAN, plan=pa.PLAN_FIELD_MAJOR | pa.PLAN_NO_PLANE)
    x1, _, _ = run_gpu(w, 20, LEAN, plan=plan)
    assert np.abs(x1 - x0).max() <= 2e-6


def test_lean_frozen_points():
    """Frozen points (xi_bc != 0) keep their uploaded FP32 values."""
    say("lean xi_bc mask")
    w = C.small_fcc((16, 16, 32), ns=4)
    bc = np.zeros((w.ns,) + w.shape, dtype=np.complex128)
    bc[1, 3:7] = 1.0
    bc[2, :, :, ::5] = 1.0
    xg, _, _ = run_gpu(w, 15, LEAN, xi_bc=bc)
    assert np.array_equal(xg[1, 3:7], f32(w.xi[1, 3:7]))
    assert np.array_equal(xg[2, :, :, ::5], f32(w.xi[2, :, :, ::5]))
    xm, _, _ = run_gpu(w, 15, pa.MIXED, xi_bc=bc)
    assert np.abs(xg - xm).max() <= 1e-5


def test_lean_raw_fft_precision_rejected():
    """The raw FFT seam takes PFDD_FP64 or PFDD_MIXED buffers only."""
    w = C.small_fcc((16, 16, 16), ns=1)
    with pa.Context(w.params(LEAN)) as c:
        buf = hip.DeviceBuffer(16 * 16 * 16 * 8)
        with pytest.raises(pa.PfddError) as ei:
            c.fft(buf.ptr, 1, pa.FORWARD, LEAN)
        assert ei.value.rc == pa.E_SHAPE
        buf.free()
