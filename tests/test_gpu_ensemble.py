"""GPU: batched ensembles and the autotuner, through the C ABI.

An instance of an ensemble runs the same per-field arithmetic as a single
context of the same plan, so the two are compared bit for bit; instances
are also checked against the FP64 oracle."""
import sys

import numpy as np
import pytest

import oraclepy as O
import pfdd_amd as pa
from pfdd_amd import configs as C
from pfdd_amd import hip

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-ensemble] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if hip.device_count() < 1:
        pytest.skip("no GPU visible")


def small_ensemble(count=6, sizes=(16, 24, 32), first=0):
    return C.EnsembleWorkload("t", count, 0, sizes=sizes, rmin=2.0, rmax=7.0, first=first)


def single_run(w, m, steps, precision, plan=0, tau=None, tau_pp=False, xi_bc=None, quirk=True):
    with pa.Context(w.params(m.index, precision, plan=plan)) as c:
        c.set_tau(m.tau if tau is None else tau, per_point=tau_pp)
        c.upload_xi(m.xi, xi_bc)
        c.begin_outer(quirk)
        g, _ = c.gl_steps(steps)
        return c.download_xi(), g


def oracle_run(w, m, steps, quirk=True):
    p = O.make_params(m.shape, w.eps, C.C11, C.C12, C.C44, [0.1] * 3, C.asf2_reference(3), [1.0] * 3)
    o = O.Oracle(p)
    o.set_xi(m.xi)
    o.set_tau(m.tau)
    o.prologue(quirk)
    g, _ = o.gl_steps(steps)
    return o.get_xi(), g


@pytest.mark.parametrize("prec", [pa.FP64, pa.MIXED, pa.MIXED_LEAN])
def test_ensemble_equals_single_contexts(prec):
    """43 steps (five graph replays + three eager steps) of a 6-instance,
    3-class ensemble against each instance run alone: bit-identical."""
    w = small_ensemble()
    members = [w.member(i) for i in range(w.count)]
    say(f"ensemble vs single prec={prec} shapes={[m.shape[0] for m in members]}")
    with pa.Ensemble(w.params_list(prec)) as e:
        ncls, cls, plans = e.info()
        assert ncls == len({m.shape for m in members})
        for m in members:
            e.set_tau(m.index, m.tau)
            e.upload_xi(m.index, m.xi)
        e.begin_outer(True)
        g = e.gl_steps(43)
        got = [e.download_xi(i) for i in range(w.count)]
    for m in members:
        xs, gs = single_run(w, m, 43, prec)
        assert np.array_equal(got[m.index], xs), m.index
        assert np.array_equal(g[:, m.index], gs), m.index


def test_ensemble_fp64_vs_oracle():
    """Four instances of the c3 generator's kind (one per grid class) against
    the oracle: 100 steps, max |xi - oracle| <= 1e-10."""
    w = small_ensemble(count=4, sizes=(16, 24, 32, 48), first=7)
    members = [w.member(i) for i in range(w.count)]
    say(f"ensemble fp64 vs oracle shapes={[m.shape[0] for m in members]}")
    with pa.Ensemble(w.params_list(pa.FP64)) as e:
        for m in members:
            e.set_tau(m.index, m.tau)
            e.upload_xi(m.index, m.xi)
        e.begin_outer(True)
        g = e.gl_steps(100)
        got = [e.download_xi(i) for i in range(w.count)]
    for m in members:
        xo, go = oracle_run(w, m, 100)
        err = np.abs(got[m.index] - xo).max()
        assert err <= 1e-10, (m.index, m.shape, err)
        assert np.abs(g[:, m.index] - go).max() <= 1e-10


def test_ensemble_mixed_frozen_and_per_point_tau():
    """One instance with frozen points, one with a per-point shear, the rest
    uniform, in one class: each equals its single-context run bit for bit,
    also after a later switch back to a uniform shear."""
    w = small_ensemble(count=4, sizes=(24,))
    members = [w.member(i) for i in range(w.count)]
    rng = np.random.default_rng(5)
    P = members[0].P
    bc = np.zeros((3,) + members[0].shape, dtype=np.complex128)
    bc.reshape(3, -1)[:, rng.random(P) < 0.05] = 1.0
    tpp = 0.02 * rng.random((P, 3))
    say("ensemble frozen + per-point tau")
    with pa.Ensemble(w.params_list(pa.FP64)) as e:
        assert e.info()[0] == 1
        for m in members:
            e.set_tau(m.index, m.tau)
            e.upload_xi(m.index, m.xi, bc if m.index == 1 else None)
        e.set_tau(2, tpp, per_point=True)
        e.begin_outer(True)
        e.gl_steps(20)
        got = [e.download_xi(i) for i in range(w.count)]
        e.set_tau(2, members[2].tau)  # back to uniform: the per-point array goes away
        e.gl_steps(9)
        got2 = e.download_xi(2)
    for m in members:
        xs, _ = single_run(w, m, 20, pa.FP64, tau=tpp if m.index == 2 else None, tau_pp=m.index == 2,
                           xi_bc=bc if m.index == 1 else None)
        assert np.array_equal(got[m.index], xs), m.index
    with pa.Context(w.params(2, pa.FP64)) as c:
        c.set_tau(tpp, per_point=True)
        c.upload_xi(members[2].xi)
        c.begin_outer(True)
        c.gl_steps(20)
        c.set_tau(members[2].tau)
        c.gl_steps(9)
        assert np.array_equal(got2, c.download_xi())


@pytest.mark.parametrize("prec", [pa.FP64, pa.MIXED])
def test_autotune_restores_state_and_agrees(prec):
    """The autotuner leaves xi as uploaded, applies a valid plan, and the run
    under that plan agrees with the AUTO plan to rounding."""
    w = small_ensemble(count=1, sizes=(32,))
    m = w.member(0)
    say(f"autotune prec={prec}")
    with pa.Context(w.params(0, prec)) as c:
        c.set_tau(m.tau)
        c.upload_xi(m.xi)
        plan, ms = c.autotune(8)
        assert ms > 0 and c.plan == plan
        assert np.array_equal(c.download_xi(), m.xi)
        c.begin_outer(True)
        c.gl_steps(30)
        x_tuned = c.download_xi()
        say(f"  plan {pa.describe_plan(plan)} {ms:.4f} ms/step")
    x_auto, _ = single_run(w, m, 30, prec)
    tol = 1e-12 if prec == pa.FP64 else 2e-5
    assert np.abs(x_tuned - x_auto).max() <= tol


def test_plan_tiles_and_errors():
    """Explicit tiles of the fused kernels give the AUTO results to rounding;
    tiles that do not fit are rejected."""
    w = small_ensemble(count=1, sizes=(32,))
    m = w.member(0)
    base, _ = single_run(w, m, 12, pa.FP64)
    for plan in (pa.PLAN_NO_PLANE | pa.plan_row_tile(2), pa.PLAN_NO_PLANE | pa.plan_row_tile(8),
                 pa.PLAN_NO_PLANE | pa.plan_col_tile(4), pa.plan_col_tile(1) | pa.PLAN_FUSED_COLS):
        say(f"plan {pa.describe_plan(plan)}")
        x, _ = single_run(w, m, 12, pa.FP64, plan=plan)
        assert np.abs(x - base).max() <= 1e-12, plan
    with pa.Context(w.params(0, pa.FP64)) as c:
        for bad in (pa.plan_row_tile(4096), pa.PLAN_UNFUSED_ROWS | pa.plan_row_tile(2), 1 << 24):
            with pytest.raises(pa.PfddError) as ei:
                c.set_plan(bad)
            assert ei.value.rc == pa.E_SHAPE


def test_ensemble_errors():
    w = small_ensemble(count=2, sizes=(16,))
    m = w.member(0)
    with pa.Ensemble(w.params_list(pa.FP64)) as e:
        with pytest.raises(pa.PfddError) as ei:
            e.begin_outer(True)  # nothing uploaded
        assert ei.value.rc == pa.E_STATE
        with pytest.raises(pa.PfddError) as ei:
            e.set_tau(2, m.tau)
        assert ei.value.rc == pa.E_SHAPE
        bad = m.xi.copy()
        bad[0, 0, 0, 0] = np.nan
        with pytest.raises(pa.PfddError) as ei:
            e.upload_xi(0, bad)
        assert ei.value.rc == pa.E_NAN
        with pytest.raises(pa.PfddError) as ei:
            e.gl_steps(1)
        assert ei.value.rc == pa.E_STATE


def test_ensemble_autotune_then_oracle():
    """Per-class autotuning of a 3-class FP64 ensemble, then 40 steps against
    the oracle (<= 1e-10)."""
    w = small_ensemble(count=3, sizes=(16, 24, 32), first=20)
    members = [w.member(i) for i in range(w.count)]
    say(f"ensemble autotune shapes={[m.shape[0] for m in members]}")
    with pa.Ensemble(w.params_list(pa.FP64)) as e:
        for m in members:
            e.set_tau(m.index, m.tau)
            e.upload_xi(m.index, m.xi)
        plans, ms = e.autotune(8)
        assert len(plans) == e.info()[0] and np.all(ms > 0)
        e.begin_outer(True)
        e.gl_steps(40)
        got = [e.download_xi(i) for i in range(w.count)]
    for m in members:
        xo, _ = oracle_run(w, m, 40)
        assert np.abs(got[m.index] - xo).max() <= 1e-10, m.index


def test_time_step_kernels_follows_the_plan_and_advances_state():
    """The in-sequence kernel timing runs real GL steps: its launch counts are
    the step plan's, and the state it leaves equals that of pfdd_gl_steps."""
    w = small_ensemble(count=1, sizes=(32,))
    m = w.member(0)
    for plan in (pa.PLAN_AUTO, pa.PLAN_NO_PLANE, pa.PLAN_NO_PLANE | pa.PLAN_UNFUSED_ROWS | pa.PLAN_UNFUSED_COLS):
        with pa.Context(w.params(0, pa.FP64, plan=plan)) as c:
            c.set_tau(m.tau)
            c.upload_xi(m.xi)
            c.begin_outer(True)
            t = c.time_step_kernels(5)
            steps_plan = c.step_plan()
            for st, (ms, n) in t.items():
                assert ms > 0 and n == steps_plan[st], (plan, st)
            x = c.download_xi()
        xs, _ = single_run(w, m, 5, pa.FP64, plan=plan)
        assert np.array_equal(x, xs), plan


@pytest.mark.parametrize("prec", [pa.FP64, pa.MIXED, pa.MIXED_LEAN])
def test_persistent_step_equals_phased(prec):
    """The persistent step (one launch per step) runs the same
    item code as the per-class kernels: an ensemble stepped both ways, with
    the plans the persistent step uses, agrees bit for bit (xi and gamma)."""
    w = small_ensemble(count=5, sizes=(32, 48, 64), first=40)
    members = [w.member(i) for i in range(w.count)]
    plan = pa.PLAN_PLANE | pa.PLAN_FUSED_COLS
    say(f"persistent vs phased prec={prec} shapes={[m.shape[0] for m in members]}")
    out = []
    for mode in (pa.ENS_PHASED, pa.ENS_PERSISTENT):
        with pa.Ensemble(w.params_list(prec, plan=plan)) as e:
            for m in members:
                e.set_tau(m.index, m.tau)
                e.upload_xi(m.index, m.xi)
            e.set_mode(mode)
            assert e.mode()[0] == mode
            e.begin_outer(True)
            g = e.gl_steps(21)
            out.append((g, [e.download_xi(i) for i in range(w.count)]))
    (g0, x0), (g1, x1) = out
    assert np.array_equal(g0, g1)
    for i in range(w.count):
        assert np.array_equal(x0[i], x1[i]), i


def test_persistent_step_vs_oracle_and_limits():
    """Persistent mode against the oracle (FP64, 30 steps, <= 1e-10); grids
    and inputs it does not cover are refused."""
    w = small_ensemble(count=3, sizes=(32, 48), first=60)
    members = [w.member(i) for i in range(w.count)]
    say("persistent vs oracle")
    with pa.Ensemble(w.params_list(pa.FP64)) as e:
        for m in members:
            e.set_tau(m.index, m.tau)
            e.upload_xi(m.index, m.xi)
        e.set_mode(pa.ENS_PERSISTENT)
        e.begin_outer(True)
        e.gl_steps(30)
        got = [e.download_xi(i) for i in range(w.count)]
        # a per-point shear is outside the persistent step
        e.set_tau(0, np.full((members[0].P, 3), 0.01), per_point=True)
        with pytest.raises(pa.PfddError) as ei:
            e.gl_steps(1)
        assert ei.value.rc == pa.E_UNSUPPORTED
    for m in members:
        xo, _ = oracle_run(w, m, 30)
        assert np.abs(got[m.index] - xo).max() <= 1e-10, m.index
    w16 = small_ensemble(count=2, sizes=(16,))
    with pa.Ensemble(w16.params_list(pa.FP64)) as e:
        with pytest.raises(pa.PfddError) as ei:
            e.set_mode(pa.ENS_PERSISTENT)
        assert ei.value.rc == pa.E_UNSUPPORTED
        assert e.mode()[0] == pa.ENS_PHASED


def test_ensemble_autotune_chooses_a_mode():
    """The ensemble autotuner times both modes, keeps the faster, and the run
    after it matches single contexts to rounding."""
    w = small_ensemble(count=3, sizes=(32, 48), first=80)
    members = [w.member(i) for i in range(w.count)]
    with pa.Ensemble(w.params_list(pa.MIXED)) as e:
        for m in members:
            e.set_tau(m.index, m.tau)
            e.upload_xi(m.index, m.xi)
        e.autotune(8)
        mode, t_ph, t_pe = e.mode()
        say(f"ensemble mode {mode}: phased {t_ph:.4f} ms, persistent {t_pe:.4f} ms")
        assert t_ph > 0 and t_pe > 0 and mode == (pa.ENS_PERSISTENT if t_pe < t_ph else pa.ENS_PHASED)
        for m in members:
            assert np.array_equal(e.download_xi(m.index), m.xi)
        e.begin_outer(True)
        e.gl_steps(16)
        got = [e.download_xi(i) for i in range(w.count)]
    for m in members:
        xs, _ = single_run(w, m, 16, pa.MIXED)
        assert np.abs(got[m.index] - xs).max() <= 2e-5, m.index
