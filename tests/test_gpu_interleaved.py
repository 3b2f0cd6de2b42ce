"""GPU parity of the interleaved work-array layout (PFDD_PLAN_INTERLEAVED):
the interleaved row kernel, the interleaved fused column kernel and the
layout-aware passes, against the FP64 oracle and against the field-major
kernels, through the C ABI."""
import sys

import numpy as np
import pytest

import oraclepy as O
import pfdd_amd as pa
from pfdd_amd import configs as C
from pfdd_amd import hip

pytestmark = pytest.mark.gpu
IL = pa.PLAN_INTERLEAVED
# whether AUTO takes the compile-time interleaved kernels, per (N1 = N3, S) and precision
AUTO_IL = {(128, 12): {pa.FP64: True, pa.MIXED: True, pa.MIXED_LEAN: True},
           (256, 12): {pa.FP64: False, pa.MIXED: True, pa.MIXED_LEAN: True},
           (512, 3): {pa.FP64: False, pa.MIXED: True, pa.MIXED_LEAN: True}}


def say(msg):
    print(f"[gpu-interleaved] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if hip.device_count() < 1:
        pytest.skip("no GPU visible")


def run_gpu(w, steps, plan, precision=pa.FP64, bmode=pa.B_RECOMPUTE, xi_bc=None, tau=None, tau_pp=False,
            strain=False):
    with pa.Context(w.params(precision, bmode, plan)) as c:
        c.set_tau(w.tau if tau is None else tau, per_point=tau_pp)
        c.upload_xi(w.xi, xi_bc)
        c.begin_outer(True)
        g, ga = c.gl_steps(steps)
        sp = c.step_plan()
        ss = c.strain_stress(w.s_tensor_avepst()) if strain else None
        return c.download_xi(), g, ga, sp, ss


def run_oracle(w, steps, xi_bc=None, tau=None, tau_pp=False, strain=False):
    p = O.make_params(w.shape, w.eps, C.C11, C.C12, C.C44, [0.1] * w.ns, C.asf2_reference(w.ns), [1.0] * w.ns)
    o = O.Oracle(p)
    o.set_xi(w.xi, xi_bc)
    o.set_tau(w.tau if tau is None else tau, per_point=tau_pp)
    o.prologue(True)
    g, ga = o.gl_steps(steps)
    ss = o.strain_stress(w.s_tensor_avepst()) if strain else None
    return o.get_xi(), g, ga, ss


CASES = [
    ((32, 32, 32), 12, IL),
    ((32, 32, 32), 12, IL | pa.PLAN_UNFUSED_COLS),
    ((32, 32, 32), 12, IL | pa.plan_row_tile(2) | pa.plan_col_tile(2)),
    ((32, 32, 32), 12, IL | pa.plan_row_tile(4) | pa.plan_col_tile(4)),
    ((16, 24, 48), 3, IL),
    ((48, 16, 32), 5, IL | pa.plan_col_tile(8)),
    ((64, 8, 96), 2, IL | pa.plan_row_tile(2)),
    ((12, 16, 64), 12, IL),          # N1 = 12: no fused column kernel, plain N1 passes
    ((8, 96, 128), 4, IL | pa.PLAN_RADIX4_AXIS2),
]


@pytest.mark.parametrize("shape,ns,plan", CASES)
def test_interleaved_fp64_vs_oracle(shape, ns, plan):
    say(f"IL fp64 {shape} ns={ns} plan={pa.describe_plan(plan)}")
    w = C.small_fcc(shape, ns=ns)
    strain = shape == (16, 24, 48)
    xo, go, gao, sso = run_oracle(w, 12, strain=strain)
    xg, gg, gag, sp, ssg = run_gpu(w, 12, plan, strain=strain)
    assert np.abs(xg - xo).max() <= 1e-12
    assert np.abs(gag - gao).max() <= 1e-12 * max(1.0, np.abs(gao).max())
    assert sp[pa.STAGE_FFT_AXIS3] == 0 and sp[pa.STAGE_FFT_AXIS2] == 2
    fused_cols = shape[0] in (32, 48, 64, 96, 128, 256, 512) and not plan & pa.PLAN_UNFUSED_COLS
    assert sp[pa.STAGE_FFT_AXIS1] == (0 if fused_cols else 2)
    if strain:
        for a, b in zip(ssg, sso):
            assert np.abs(np.asarray(a) - np.asarray(b)).max() <= 1e-12


@pytest.mark.parametrize("shape,ns,plan", [((32, 32, 32), 12, IL), ((32, 32, 32), 12, IL | pa.PLAN_UNFUSED_COLS),
                                           ((16, 24, 48), 3, IL), ((48, 16, 32), 5, IL | pa.plan_row_tile(2))])
def test_interleaved_mixed_matches_field_major(shape, ns, plan):
    say(f"IL mixed {shape} ns={ns} plan={pa.describe_plan(plan)}")
    w = C.small_fcc(shape, ns=ns)
    x0, g0, _, _, _ = run_gpu(w, 20, pa.PLAN_FIELD_MAJOR | pa.PLAN_NO_PLANE, precision=pa.MIXED)
    x1, g1, _, _, _ = run_gpu(w, 20, plan, precision=pa.MIXED)
    assert np.abs(x1 - x0).max() <= 2e-6
    xo, _, _, _ = run_oracle(w, 20)
    assert np.abs(x1 - xo).max() <= 1e-4


@pytest.mark.parametrize("prec", [pa.FP64, pa.MIXED])
def test_interleaved_stored_B(prec):
    say(f"IL stored B prec={prec}")
    w = C.small_fcc((32, 32, 32), ns=12)
    x1, _, _, sp, _ = run_gpu(w, 12, IL, precision=prec, bmode=pa.B_STORED)
    assert sp[pa.STAGE_FFT_AXIS1] == 2       # stored B takes the plain N1 passes
    if prec == pa.FP64:
        xo, _, _, _ = run_oracle(w, 12)
        assert np.abs(x1 - xo).max() <= 1e-12
    else:
        x0, _, _, _, _ = run_gpu(w, 12, pa.PLAN_FIELD_MAJOR | pa.PLAN_NO_PLANE, precision=prec, bmode=pa.B_STORED)
        assert np.abs(x1 - x0).max() <= 2e-6


def test_interleaved_frozen_points_and_per_point_tau():
    say("IL xi_bc mask + per-point tau")
    w = C.small_fcc((16, 16, 32), ns=4)
    rng = np.random.default_rng(3)
    bc = np.zeros((w.ns,) + w.shape, dtype=np.complex128)
    bc[1, 3:7] = 1.0
    bc[2, :, :, ::5] = 1.0
    tau = rng.uniform(-0.02, 0.02, (w.P, w.ns))
    xo, _, _, _ = run_oracle(w, 15, xi_bc=bc, tau=tau, tau_pp=True)
    xg, _, _, _, _ = run_gpu(w, 15, IL, xi_bc=bc, tau=tau, tau_pp=True)
    assert np.abs(xg - xo).max() <= 1e-12
    assert np.array_equal(xg[1, 3:7], w.xi[1, 3:7])


@pytest.mark.parametrize("prec", [pa.FP64, pa.MIXED, pa.MIXED_LEAN])
def test_interleaved_ensemble_equals_single(prec):
    """A batched class in the interleaved layout gives each instance the bits
    of its single-context run with the same plan."""
    say(f"IL ensemble prec={prec}")
    members = [C.small_fcc((32, 32, 32), ns=3, seed=C.BASE_SEED + 300 + i) for i in range(3)]
    params = [m.params(prec, pa.B_RECOMPUTE, IL) for m in members]
    with pa.Ensemble(params) as e:
        for i, m in enumerate(members):
            e.set_tau(i, m.tau)
            e.upload_xi(i, m.xi)
        e.begin_outer(True)
        g = e.gl_steps(20)
        xs = [e.download_xi(i) for i in range(3)]
    for i, m in enumerate(members):
        x1, g1, _, _, _ = run_gpu(m, 20, IL, precision=prec)
        assert np.array_equal(xs[i], x1)
        assert np.array_equal(g[:, i], g1)


@pytest.mark.parametrize("n,ns", [(128, 12), (256, 12), (512, 3)])
@pytest.mark.parametrize("prec", [pa.FP64, pa.MIXED, pa.MIXED_LEAN])
def test_fixed_kernels_equal_field_major(n, ns, prec):
    """With recomputed B the interleaved plan runs the compile-time row and
    column kernels (k_row_il_fixed, k_col_il_fixed) at S = 12 with N1 = N3
    in {128, 256} and at S = 3 with N1 = N3 = 512.  They do the field-major
    kernels' arithmetic per pencil and per k-point, so xi comes out bit for
    bit the same; gamma only regroups its block sums.  AUTO takes them except
    in FP64 at 256."""
    say(f"fixed interleaved kernels n={n} ns={ns} prec={prec}")
    w = C.small_fcc((n, 8, n), ns=ns, nloops=8, rmin=6.0, rmax=20.0)
    x0, g0, _, _, _ = run_gpu(w, 16, pa.PLAN_FIELD_MAJOR | pa.PLAN_NO_PLANE, precision=prec)
    x1, g1, _, sp, _ = run_gpu(w, 16, IL, precision=prec)
    assert sp[pa.STAGE_FFT_AXIS3] == 0 and sp[pa.STAGE_FFT_AXIS1] == 0   # fused rows and fused columns
    assert np.array_equal(x1, x0)
    assert np.abs(g1 - g0).max() <= 1e-14 * max(1.0, np.abs(g0).max())
    x2, _, _, sp2, _ = run_gpu(w, 16, 0, precision=prec)
    auto_il = AUTO_IL[(n, ns)][prec]
    assert sp2[pa.STAGE_FFT_AXIS1] == (0 if auto_il else 2)
    assert np.array_equal(x2, x0)
    if prec == pa.FP64:
        xo, go, _, _ = run_oracle(w, 16)
        assert np.abs(x1 - xo).max() <= 1e-12


def test_interleaved_plan_errors():
    w = C.small_fcc((32, 32, 32), ns=12)
    for plan in (IL | pa.PLAN_FIELD_MAJOR, IL | pa.PLAN_UNFUSED_ROWS, IL | pa.PLAN_PLANE,
                 IL | pa.plan_row_tile(128), IL | pa.plan_col_tile(1024)):
        with pytest.raises(pa.PfddError) as ei:
            with pa.Context(w.params(pa.FP64, pa.B_RECOMPUTE, plan)):
                pass
        assert ei.value.rc == pa.E_SHAPE
    w = C.small_fcc((32, 32, 12), ns=3)                  # N3 = 12: no fused row kernel, so no interleaved layout
    with pytest.raises(pa.PfddError):
        with pa.Context(w.params(pa.FP64, pa.B_RECOMPUTE, IL)):
            pass
