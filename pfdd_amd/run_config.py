"""Run a benchmark configuration on one MI355X and report throughput and the
mixed-precision error against the FP64 run of the same instance.

    python -m pfdd_amd.run_config c2 [--steps N] [--json out.json]
    python -m pfdd_amd.run_config c4 [--steps N] [--every 50] [--json out.json]
    python -m pfdd_amd.run_config c3 [--steps N] [--count 512] [--json out.json]

c2: 256^3 x 12, 64 seeded loops: FP64, mixed (stored and recomputed B) and
    lean mixed (plain FP32 master, recomputed B) GL
    runs of N steps (default 2000); reports GL steps/s, grid-points x systems
    / s, and the errors of the mixed runs in xi (max abs), slipped area
    (sum_a gamma_a P) and mean stress.
c4: 512^3 x 3 with the applied-shear ramp 0 -> 0.03 mu: mixed, lean mixed
    (recompute B) and FP64 runs with <sigma>, <eps> sampled every `every` steps (the
    stress-strain curve), plus mixed with stored B over a shorter window for
    the stored-vs-recompute comparison.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import (B_RECOMPUTE, B_STORED, ENS_PERSISTENT, ENS_PHASED, FP64, MIXED, MIXED_LEAN, Context, Ensemble, PfddError,
               describe_plan)
from . import configs as C
from . import hip


def _timed(ctx, n, record=True):
    e0, e1 = hip.Event(), hip.Event()
    e0.record(ctx.stream)
    g, ga = ctx.gl_steps(n, record=record)
    e1.record(ctx.stream)
    e1.synchronize()
    return e0.elapsed_ms(e1), g, ga


def run_plain(w, steps, precision, bmode, chunk=500):
    ctx = Context(w.params(precision, bmode))
    ctx.set_tau(w.tau)
    ctx.upload_xi(w.xi)
    ctx.begin_outer(True)
    ms = 0.0
    ga_all = []
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        t, _, ga = _timed(ctx, n)
        ms += t
        ga_all.append(ga)
        done += n
    ap = w.s_tensor_avepst()
    aveps, avesig = ctx.strain_stress(ap)
    xi = ctx.download_xi()
    out = {"steps": steps, "ms": ms, "steps_per_s": steps / (ms / 1e3), "device_bytes": ctx.device_bytes}
    ctx.close()
    return out, xi, np.concatenate(ga_all), aveps, avesig


def run_ramp(w, steps, precision, bmode, every):
    ctx = Context(w.params(precision, bmode))
    ctx.set_tau_ramp(w.tau, w.tau_rate)
    ctx.upload_xi(w.xi)
    ctx.begin_outer(True)
    ms = 0.0
    curve = []
    done = 0
    while done < steps:
        n = min(every, steps - done)
        t, _, ga = _timed(ctx, n)
        ms += t
        done += n
        sig = w.sigma_unit * 0.03 * (done - 1) / (steps - 1)   # stress of the last step taken
        aveps, avesig = ctx.strain_stress(w.s_tensor_avepst(sig))
        ctx.begin_outer(False)                                 # rebuild data2 from the current xi
        curve.append({"step": done, "applied_tau0": float(np.sum(sig * w.eps[0])),
                      "gamma_a": ga[-1].tolist(), "aveps": aveps.tolist(), "avesigma": avesig.tolist()})
    xi = ctx.download_xi()
    out = {"steps": steps, "ms_gl_only": ms, "steps_per_s": steps / (ms / 1e3), "device_bytes": ctx.device_bytes}
    ctx.close()
    return out, xi, curve


def _timed_ens(e, n):
    e0, e1 = hip.Event(), hip.Event()
    e0.record(e.stream)
    e.gl_steps(n, record=False)
    e1.record(e.stream)
    e1.synchronize()
    return e0.elapsed_ms(e1)


def build_ensemble(w, precision):
    e = Ensemble(w.params_list(precision, B_RECOMPUTE))
    for i in range(w.count):
        m = w.member(i)
        e.set_tau(i, m.tau)
        e.upload_xi(i, m.xi)
    return e


def run_ensemble(w, steps, precision, tune_steps=8):
    """AUTO-plan timing, then per-class autotuning and the timed run."""
    t0 = time.time()
    e = build_ensemble(w, precision)
    setup_s = time.time() - t0
    ncls, cls, _ = e.info()
    e.begin_outer(True)
    e.gl_steps(8, record=False)
    auto_ms = _timed_ens(e, 64) / 64
    # back to the initial state for the measured run
    for i in range(w.count):
        e.upload_xi(i, w.member(i).xi)
    t1 = time.time()
    plans, tuned_ms = e.autotune(tune_steps)
    tune_s = time.time() - t1
    mode, t_ph, t_pe = e.mode()
    e.begin_outer(True)
    ms = _timed_ens(e, steps)
    # the mode the autotuner did not keep, over 64 steps from the initial state
    other = {}
    try:
        xs = [e.download_xi(i) for i in range(w.count)]
        for i in range(w.count):
            e.upload_xi(i, w.member(i).xi)
        if mode == ENS_PHASED:
            e.set_mode(ENS_PERSISTENT)
        else:
            e.set_mode(ENS_PHASED)
            for k in range(ncls):
                e.set_plan(k, int(plans[k]))
        e.begin_outer(True)
        e.gl_steps(8, record=False)
        other = {"mode": "persistent" if mode == ENS_PHASED else "phased", "ms_per_step": _timed_ens(e, 64) / 64}
        for i in range(w.count):  # leave the final state of the measured run in place
            e.upload_xi(i, xs[i])
        del xs
    except PfddError as err:
        other = {"error": str(err)}
    classes = []
    for k in range(ncls):
        members = np.nonzero(cls == k)[0]
        classes.append({"edge": w.shape(int(members[0]))[0], "instances": int(len(members)),
                        "plan": int(plans[k]), "plan_name": describe_plan(int(plans[k])),
                        "ms_per_step": float(tuned_ms[k])})
    pts = w.total_points() * w.ns
    out = {"steps": steps, "ms": ms, "steps_per_s": steps / (ms / 1e3),
           "gridpoints_x_systems_per_s": pts * steps / (ms / 1e3),
           "auto_plan_ms_per_step": auto_ms, "auto_plan_gridpoints_x_systems_per_s": pts / (auto_ms / 1e3),
           "classes": classes, "setup_s": setup_s, "autotune_s": tune_s, "device_bytes": e.device_bytes,
           "mode": "persistent" if mode == ENS_PERSISTENT else "phased",
           "autotune_ms_per_step": {"phased": t_ph, "persistent": t_pe}, "other_mode": other}
    return out, e


def step_bytes(S, P, fp64, stored, lean=False):
    """Stage-minimum bytes of a GL step (the mixed master xi takes 12 bytes a
    point, 8 in the lean form)."""
    if fp64:
        return (144 * S + (4 * S * (S + 1) if stored else 0)) * P
    return ((80 if lean else 88) * S + (2 * S * (S + 1) if stored else 0)) * P


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("config", choices=["c2", "c3", "c4"])
    ap.add_argument("--count", type=int, default=512)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--every", type=int, default=50)
    ap.add_argument("--json", default=None)
    a = ap.parse_args()
    t0 = time.time()
    res = {"config": a.config}
    if a.config == "c2":
        w = C.c2()
        steps = a.steps or w.steps
        print(f"[run_config] c2 {w.shape} x {w.ns}, {steps} steps, setup {time.time() - t0:.1f}s", flush=True)
        r64, x64, ga64, e64, s64 = run_plain(w, steps, FP64, B_RECOMPUTE)
        print(f"[run_config] fp64 {r64['steps_per_s']:.2f} steps/s", flush=True)
        r32, x32, ga32, e32, s32 = run_plain(w, steps, MIXED, B_STORED)
        print(f"[run_config] mixed stored B {r32['steps_per_s']:.2f} steps/s", flush=True)
        rr, xr, gar, er, sr = run_plain(w, steps, MIXED, B_RECOMPUTE)
        print(f"[run_config] mixed recomputed B {rr['steps_per_s']:.2f} steps/s", flush=True)
        res["mixed_recompute_B"] = dict(rr, speedup=rr["steps_per_s"] / r64["steps_per_s"],
                                        xi_max_abs_err=float(np.abs(xr - x64).max()),
                                        slipped_area_rel_err=float(abs(gar[-1].sum() * w.P - ga64[-1].sum() * w.P)
                                                                   / abs(ga64[-1].sum() * w.P)),
                                        mean_stress_max_abs_err=float(np.abs(sr - s64).max()),
                                        step_roofline_frac=step_bytes(w.ns, w.P, False, False) * rr["steps_per_s"] / 8e12)
        del xr
        rl, xl, gal, el, sl = run_plain(w, steps, MIXED_LEAN, B_RECOMPUTE)
        print(f"[run_config] mixed lean (FP32 master) recomputed B {rl['steps_per_s']:.2f} steps/s", flush=True)
        res["mixed_lean_recompute_B"] = dict(rl, speedup=rl["steps_per_s"] / r64["steps_per_s"],
                                             xi_max_abs_err=float(np.abs(xl - x64).max()),
                                             slipped_area_rel_err=float(abs(gal[-1].sum() - ga64[-1].sum())
                                                                        / abs(ga64[-1].sum())),
                                             mean_stress_max_abs_err=float(np.abs(sl - s64).max()),
                                             step_roofline_frac=step_bytes(w.ns, w.P, False, False, True)
                                             * rl["steps_per_s"] / 8e12)
        del xl
        area64, area32 = ga64[-1].sum() * w.P, ga32[-1].sum() * w.P
        res.update({"shape": w.shape, "ns": w.ns, "steps": steps, "fp64": r64, "mixed_stored_B": r32,
                    "speedup": r32["steps_per_s"] / r64["steps_per_s"],
                    "gridpoints_x_systems_per_s_mixed": w.P * w.ns * r32["steps_per_s"],
                    "xi_max_abs_err": float(np.abs(x32 - x64).max()),
                    "slipped_area_fp64": float(area64), "slipped_area_rel_err": float(abs(area32 - area64) / abs(area64)),
                    "mean_stress_max_abs_err": float(np.abs(s32 - s64).max()),
                    "mean_stress_fp64": s64.tolist(),
                    "step_roofline_frac_mixed": step_bytes(w.ns, w.P, False, True) * r32["steps_per_s"] / 8e12,
                    "step_roofline_frac_fp64": step_bytes(w.ns, w.P, True, False) * r64["steps_per_s"] / 8e12})
    elif a.config == "c3":
        w = C.c3(count=a.count)
        steps = a.steps or w.steps
        print(f"[run_config] c3 {w.count} instances, {w.total_points()} points x 3, {steps} steps", flush=True)
        r32, e32 = run_ensemble(w, steps, MIXED)
        print(f"[run_config] mixed {r32['steps_per_s']:.2f} ensemble steps/s "
              f"({r32['gridpoints_x_systems_per_s']:.3e} pts*sys/s)", flush=True)
        r64, e64 = run_ensemble(w, steps, FP64)
        print(f"[run_config] fp64 {r64['steps_per_s']:.2f} ensemble steps/s", flush=True)
        xerr, aerr = 0.0, 0.0
        for i in range(w.count):
            x32, x64 = e32.download_xi(i), e64.download_xi(i)
            xerr = max(xerr, float(np.abs(x32 - x64).max()))
            a32, a64 = float(np.sum(x32.real)), float(np.sum(x64.real))   # slipped area, sum_a gamma_a P
            aerr = max(aerr, abs(a32 - a64) / max(abs(a64), 1e-300))
        e32.close()
        rl, el = run_ensemble(w, steps, MIXED_LEAN)
        print(f"[run_config] lean mixed {rl['steps_per_s']:.2f} ensemble steps/s "
              f"({rl['gridpoints_x_systems_per_s']:.3e} pts*sys/s)", flush=True)
        lxerr, laerr = 0.0, 0.0
        for i in range(w.count):
            xl, x64 = el.download_xi(i), e64.download_xi(i)
            lxerr = max(lxerr, float(np.abs(xl - x64).max()))
            al, a64 = float(np.sum(xl.real)), float(np.sum(x64.real))
            laerr = max(laerr, abs(al - a64) / max(abs(a64), 1e-300))
        el.close()
        e64.close()
        # algorithmic bytes of one ensemble step (recompute B)
        pts = w.total_points()
        res["mixed_lean"] = dict(rl, speedup=rl["steps_per_s"] / r64["steps_per_s"], xi_max_abs_err=lxerr,
                                 slipped_area_max_rel_err=laerr,
                                 step_roofline_frac=step_bytes(3, pts, False, False, True) * rl["steps_per_s"] / 8e12)
        res.update({"count": w.count, "ns": w.ns, "total_points": pts, "steps": steps,
                    "mixed": r32, "fp64": r64, "speedup": r32["steps_per_s"] / r64["steps_per_s"],
                    "xi_max_abs_err": xerr, "slipped_area_max_rel_err": aerr,
                    "step_roofline_frac_mixed": step_bytes(3, pts, False, False) * r32["steps_per_s"] / 8e12,
                    "step_roofline_frac_fp64": step_bytes(3, pts, True, False) * r64["steps_per_s"] / 8e12})
    else:
        w = C.c4(steps=a.steps or 5000)
        steps = w.steps
        print(f"[run_config] c4 {w.shape} x {w.ns}, {steps} steps, setup {time.time() - t0:.1f}s", flush=True)
        r32, x32, c32 = run_ramp(w, steps, MIXED, B_RECOMPUTE, a.every)
        print(f"[run_config] mixed recompute {r32['steps_per_s']:.2f} steps/s", flush=True)
        r64, x64, c64 = run_ramp(w, steps, FP64, B_RECOMPUTE, a.every)
        print(f"[run_config] fp64 recompute {r64['steps_per_s']:.2f} steps/s", flush=True)
        rl, xl, cl = run_ramp(w, steps, MIXED_LEAN, B_RECOMPUTE, a.every)
        print(f"[run_config] lean mixed recompute {rl['steps_per_s']:.2f} steps/s", flush=True)
        res["mixed_lean_recompute"] = dict(
            rl, speedup=rl["steps_per_s"] / r64["steps_per_s"], xi_max_abs_err=float(np.abs(xl - x64).max()),
            stress_strain_max_abs_err={
                "avesigma": float(max(abs(np.array(p["avesigma"]) - np.array(q["avesigma"])).max()
                                      for p, q in zip(cl, c64))),
                "aveps": float(max(abs(np.array(p["aveps"]) - np.array(q["aveps"])).max() for p, q in zip(cl, c64)))})
        del xl
        win = min(steps, 200)
        rs, _, _ = run_ramp(C.c4(B_STORED, steps=steps), win, MIXED, B_STORED, win)
        rr, _, _ = run_ramp(w, win, MIXED, B_RECOMPUTE, win)
        print(f"[run_config] stored {rs['steps_per_s']:.2f} vs recompute {rr['steps_per_s']:.2f} steps/s", flush=True)
        curve_err = max(abs(np.array(p["avesigma"]) - np.array(q["avesigma"])).max() for p, q in zip(c32, c64))
        ceps_err = max(abs(np.array(p["aveps"]) - np.array(q["aveps"])).max() for p, q in zip(c32, c64))
        res.update({"shape": w.shape, "ns": w.ns, "steps": steps, "every": a.every,
                    "mixed_recompute": r32, "fp64_recompute": r64,
                    "stored_vs_recompute_window": {"steps": win, "mixed_stored_steps_per_s": rs["steps_per_s"],
                                                   "mixed_recompute_steps_per_s": rr["steps_per_s"],
                                                   "stored_GB_s": step_bytes(3, w.P, False, True) * rs["steps_per_s"] / 1e9,
                                                   "recompute_GB_s": step_bytes(3, w.P, False, False) * rr["steps_per_s"] / 1e9},
                    "xi_max_abs_err": float(np.abs(x32 - x64).max()),
                    "stress_strain_max_abs_err": {"avesigma": float(curve_err), "aveps": float(ceps_err)},
                    "curve_fp64": c64, "curve_mixed": c32})
    res["wall_s"] = time.time() - t0
    txt = json.dumps(res)
    if a.json:
        with open(a.json, "w") as fh:
            fh.write(txt)
    summary = {k: v for k, v in res.items() if not k.startswith("curve")}
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    sys.exit(main())
