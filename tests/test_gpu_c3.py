"""GPU: the c3 batched-ensemble workload at full size, through the C ABI:
512 independent NS=3 instances on grids drawn from {32,48,64,96}^3 (1.74e8
grid points), 1-4 seeded loops each, resolved shear ~ U[0, 0.05] mu.

Twenty GL steps (two captured graph chunks of 8 and four eager steps) of the
FP64 ensemble are the reference of the module: one sampled instance per grid
class must equal its own single-context run bit for bit, one FP64 instance
per grid edge must match the oracle to 1e-10, and the mixed-precision
ensembles must stay within 1e-4 of it in xi and 1e-3 in slipped area.  The
persistent step runs the whole ensemble too (96^3 included)."""
import sys

import numpy as np
import pytest

import oraclepy as O
import pfdd_amd as pa
from pfdd_amd import configs as C
from pfdd_amd import hip

pytestmark = pytest.mark.gpu
STEPS = 20


def say(msg):
    print(f"[gpu-c3] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if hip.device_count() < 1:
        pytest.skip("no GPU visible")


@pytest.fixture(scope="module")
def w():
    return C.c3()


def run_ensemble(w, prec, mode=pa.ENS_PHASED):
    """The c3 ensemble after STEPS GL steps (left open for downloads) and its gamma [STEPS][count]."""
    e = pa.Ensemble(w.params_list(prec))
    for i in range(w.count):
        m = w.member(i)
        e.set_tau(i, m.tau)
        e.upload_xi(i, m.xi)
    if mode != pa.ENS_PHASED:
        e.set_mode(mode)
    e.begin_outer(True)
    g = e.gl_steps(STEPS)
    return e, g


@pytest.fixture(scope="module")
def fp64_run(w):
    say(f"c3 fp64 ensemble: {w.count} instances, {w.total_points()} points, {STEPS} steps")
    e, g = run_ensemble(w, pa.FP64)
    yield e, g
    e.close()


def samples(w):
    """The first instance of each grid edge."""
    first = {}
    for i in range(w.count):
        first.setdefault(w.shape(i)[0], i)
    return first


def test_c3_workload_shape(w):
    edges = sorted({w.shape(i)[0] for i in range(w.count)})
    assert w.count == 512 and edges == [32, 48, 64, 96]
    assert w.total_points() > 1.6e8


def test_c3_fp64_instances_equal_single_contexts_and_oracle(w, fp64_run):
    e, g = fp64_run
    ncls, cls, _ = e.info()
    assert ncls == 4
    for n, i in sorted(samples(w).items()):
        m = w.member(i)
        say(f"instance {i} ({n}^3): single context and oracle")
        with pa.Context(w.params(i, pa.FP64)) as c:
            c.set_tau(m.tau)
            c.upload_xi(m.xi)
            c.begin_outer(True)
            gs, _ = c.gl_steps(STEPS)
            xs = c.download_xi()
        xe = e.download_xi(i)
        assert np.array_equal(xe, xs), i
        assert np.array_equal(g[:, i], gs), i
        p = O.make_params(m.shape, w.eps, C.C11, C.C12, C.C44, [0.1] * 3, C.asf2_reference(3), [1.0] * 3)
        o = O.Oracle(p)
        o.set_xi(m.xi)
        o.set_tau(m.tau)
        o.prologue(True)
        go, _ = o.gl_steps(STEPS)
        assert np.abs(xe - o.get_xi()).max() <= 1e-10, i
        assert np.abs(g[:, i] - go).max() <= 1e-10, i
        o.close()


@pytest.mark.parametrize("prec", [pa.MIXED, pa.MIXED_LEAN])
def test_c3_mixed_vs_fp64(w, fp64_run, prec):
    e64, g64 = fp64_run
    say(f"c3 precision={prec} ensemble")
    e, g = run_ensemble(w, prec)
    try:
        worst_xi = worst_area = 0.0
        for i in range(w.count):
            x64, x32 = e64.download_xi(i), e.download_xi(i)
            worst_xi = max(worst_xi, float(np.abs(x32 - x64).max()))
            a64 = x64.real.sum()
            worst_area = max(worst_area, abs(x32.real.sum() - a64) / abs(a64))
        say(f"c3 precision={prec}: max|dxi| = {worst_xi:.3e}, slipped area rel {worst_area:.3e}")
        assert worst_xi <= 1e-4
        assert worst_area <= 1e-3
        assert np.abs(g - g64).max() <= 1e-3 * np.abs(g64).max()
        # the sampled instances against their single contexts, bit for bit
        for n, i in sorted(samples(w).items()):
            m = w.member(i)
            with pa.Context(w.params(i, prec)) as c:
                c.set_tau(m.tau)
                c.upload_xi(m.xi)
                c.begin_outer(True)
                gs, _ = c.gl_steps(STEPS)
                assert np.array_equal(e.download_xi(i), c.download_xi()), i
            assert np.array_equal(g[:, i], gs), i
    finally:
        e.close()


def test_c3_persistent_step_full_ensemble(w, fp64_run):
    """The persistent step over the whole FP64 ensemble (all four grid
    classes, 96^3 included) against the phased AUTO-plan run: the plans
    differ (plane + fused columns everywhere), so the results agree to
    rounding (1e-12)."""
    e64, g64 = fp64_run
    say("c3 fp64 ensemble, persistent step")
    e, g = run_ensemble(w, pa.FP64, pa.ENS_PERSISTENT)
    try:
        assert e.mode()[0] == pa.ENS_PERSISTENT
        e.sync()
        worst = 0.0
        for i in range(w.count):
            worst = max(worst, float(np.abs(e.download_xi(i) - e64.download_xi(i)).max()))
        say(f"persistent vs phased: max|dxi| = {worst:.3e}")
        assert worst <= 1e-12
        assert np.abs(g - g64).max() <= 1e-12
    finally:
        e.close()


@pytest.mark.parametrize("prec", [pa.FP64, pa.MIXED])
def test_persistent_96_cubed_equals_phased(prec):
    """A small ensemble with two 96^3 instances and one 32^3: the persistent
    step against the phased step with the same plans, bit for bit (xi and
    gamma), and in FP64 against the oracle."""
    w = C.EnsembleWorkload("p96", 3, 0, sizes=(96, 96, 32), first=301)
    shapes = [w.shape(i)[0] for i in range(w.count)]
    assert shapes.count(96) == 2
    plan = pa.PLAN_PLANE | pa.PLAN_FUSED_COLS
    say(f"persistent vs phased at 96^3, prec={prec}, shapes={shapes}")
    out = []
    for mode in (pa.ENS_PHASED, pa.ENS_PERSISTENT):
        with pa.Ensemble(w.params_list(prec, plan=plan)) as e:
            for i in range(w.count):
                m = w.member(i)
                e.set_tau(i, m.tau)
                e.upload_xi(i, m.xi)
            e.set_mode(mode)
            e.begin_outer(True)
            g = e.gl_steps(STEPS)
            e.sync()
            out.append((g, [e.download_xi(i) for i in range(w.count)]))
    (g0, x0), (g1, x1) = out
    assert np.array_equal(g0, g1)
    for i in range(w.count):
        assert np.array_equal(x0[i], x1[i]), i
    if prec == pa.FP64:
        i = shapes.index(96)
        m = w.member(i)
        p = O.make_params(m.shape, w.eps, C.C11, C.C12, C.C44, [0.1] * 3, C.asf2_reference(3), [1.0] * 3)
        o = O.Oracle(p)
        o.set_xi(m.xi)
        o.set_tau(m.tau)
        o.prologue(True)
        o.gl_steps(STEPS)
        assert np.abs(x1[i] - o.get_xi()).max() <= 1e-10
        o.close()
