"""Benchmark: PFDD GL steps/s on the 128^3 x 12-slip-system workload (c1) on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-mixed]

A step is one GL iteration (update, forward 3D FFT, B(k) contraction, inverse
3D FFT) over the whole 128^3 x 12 instance, inputs resident in HBM.  Prints one
JSON line.  value = FP64 GL steps/s; the mixed-precision paths (FP32 FFT and
contraction, FP64 update of a compensated FP32 master xi, or of a plain FP32
one) are timed beside it with their error against FP64.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import pfdd_amd as pa  # noqa: E402
from pfdd_amd import configs as C  # noqa: E402
from pfdd_amd import hip  # noqa: E402

METRIC = "GL steps/s at 128^3×12 slip, FP64 vs mixed FP32 at stated error; % HBM peak"
HBM_PEAK_GBS = 8000.0
STAGES = [("k_update", pa.STAGE_UPDATE, 1), ("k_fft_rows(N3)", pa.STAGE_FFT_AXIS3, 2),
          ("k_fft_cols(N2)", pa.STAGE_FFT_AXIS2, 2), ("k_fft_cols(N1)", pa.STAGE_FFT_AXIS1, 2),
          ("k_contract", pa.STAGE_CONTRACT, 1)]


def stage_bytes(stage, S, P, ntri, fp64, stored, lean=False):
    """Algorithmic bytes of one launch: each input read once, each output written once."""
    c = 16 if fp64 else 8
    m = 16 if fp64 else 8 if lean else 12         # master xi: double2, plain FP32 or the compensated FP32 form
    if stage == pa.STAGE_FINISH:
        return 0
    r = 8 if fp64 else 4
    if stage == pa.STAGE_UPDATE:
        return S * P * (c + m + m + c)            # read W, read xi, write xi, write W
    if stage in (pa.STAGE_FFT_AXIS3, pa.STAGE_FFT_AXIS2, pa.STAGE_FFT_AXIS1):
        return S * P * 2 * c
    if stage == pa.STAGE_CONTRACT:
        return S * P * 2 * c + (ntri * P * r if stored else 0)
    if stage == pa.STAGE_COPY:
        return S * P * (m + c)
    raise ValueError(stage)


def fixed_interleaved(shape, S, stored):
    """AUTO runs the interleaved layout with the compile-time row and column
    kernels (k_row_il_fixed, k_col_il_fixed) for S = 12 at N1 = N3 = 128 with
    recomputed B (apply_plan in pfdd_api.hip)."""
    return S == 12 and shape[0] == 128 and shape[2] == 128 and not stored


def rocprof_name(stage, plan, fp64, shape, S, stored, lean=False):
    """The kernel symbol rocprofv3 reports for a stage of this context (AUTO plan)."""
    T = "double" if fp64 else "float"
    M = "pfdd::Master<double> " if fp64 else "pfdd::MasterF32L" if lean else "pfdd::MasterF32C"
    n1, n2, n3 = shape
    sm = 4 if S <= 4 else 8 if S <= 8 else 12 if S <= 12 else 16
    if fixed_interleaved(shape, S, stored):
        if stage == pa.STAGE_UPDATE:
            return f"void pfdd::k_row_il_fixed<{T}, {n3}, 12, 1, {M}>"
        if stage == pa.STAGE_CONTRACT:
            return f"void pfdd::k_col_il_fixed<{T}, {n1}, 12, {1 if fp64 else 2}>"
        if stage == pa.STAGE_FFT_AXIS2:
            return f"void pfdd::k_fft_cols<{T}, {n2}, 8, false>"
        return None
    if stage == pa.STAGE_UPDATE:
        return (f"void pfdd::k_row_step<{T}, {n3}, false, false, {M}>" if plan[pa.STAGE_FFT_AXIS3] == 0
                else f"void pfdd::k_update<{T}, false, false>")
    # (the AUTO plan: radix-8 sequences, field-major layout, narrow column tiles)
    if stage == pa.STAGE_FFT_AXIS3:
        return f"void pfdd::k_fft_rows<{T}, {n3}, 8>"
    if stage == pa.STAGE_FFT_AXIS2:
        return f"void pfdd::k_fft_cols<{T}, {n2}, 8, false>"
    if stage == pa.STAGE_FFT_AXIS1:
        return f"void pfdd::k_fft_cols<{T}, {n1}, 8, false>"
    if stage == pa.STAGE_CONTRACT:
        return (f"void pfdd::k_col_contract<{T}, {n1}, {sm}>" if plan[pa.STAGE_FFT_AXIS1] == 0
                else f"void pfdd::k_contract<{T}, {sm}, {'true' if stored else 'false'}, false>")
    return None


def measured_traffic(name):
    """HBM bytes per dispatch from the committed rocprofv3 PMC summary (FETCH_SIZE and
    WRITE_SIZE runs, calibrated on the plain copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", "c1_128cubed_ns12_pmc_traffic.json")
    if not name or not os.path.exists(path):
        return None
    with open(path) as fh:
        kernels = json.load(fh)["kernels"]
    # rocprofv3 names carry the argument list after the template arguments
    k = kernels.get(name) or next((v for key, v in kernels.items() if key.startswith(name + "(")), None)
    return None if k is None else int(round(k["traffic_bytes"]))


def step_alg_bytes(S, P, fp64, stored, lean=False):
    """Stage-minimum bytes of one GL step (C2C, FP64: (144 S [+4 S(S+1)]) P;
    mixed with the 12-byte compensated master: (88 S [+2 S(S+1)]) P; with the
    8-byte plain FP32 master: (80 S [+2 S(S+1)]) P)."""
    if fp64:
        return (144 * S + (4 * S * (S + 1) if stored else 0)) * P
    return ((80 if lean else 88) * S + (2 * S * (S + 1) if stored else 0)) * P


# GL steps per captured hipGraph chunk (GRAPH_STEPS in pfdd_api.hip): the
# first gl_steps call of at least this many steps captures and instantiates
# the graph, so the warm-up runs at least that many and the timed region only
# replays it.
GRAPH_STEPS = 8


def progress(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def timed_steps(ctx, steps, warmup):
    """(ms for `steps` GL steps, host wall s, warm-up steps done): the warm-up
    is at least 4 GRAPH_STEPS steps, so the graph is captured before e0 and
    the clocks have settled (a 5-step warm-up left 20-step runs 5% slower
    than long ones)."""
    wu = max(warmup, 4 * GRAPH_STEPS)
    ctx.gl_steps(wu, record=False)
    ctx.sync()
    e0, e1 = hip.Event(), hip.Event()
    t0 = time.perf_counter()
    e0.record(ctx.stream)
    ctx.gl_steps(steps, record=False)
    e1.record(ctx.stream)
    e1.synchronize()
    ctx.sync()
    wall = time.perf_counter() - t0
    return e0.elapsed_ms(e1), wall, wu


def time_kernels(ctx, nsteps=20):
    """Per kernel of the step: (stage, launches per step, average ms per launch),
    from HIP events recorded between the launches of nsteps eager GL steps on
    the context's stream (pfdd_time_step_kernels), so every launch is timed in
    the sequence it runs in during a step."""
    plan = ctx.step_plan()
    names = {st: n for n, st, _ in STAGES}
    names[pa.STAGE_FINISH] = "k_finish_sums"
    fixed = fixed_interleaved(ctx.shape, ctx.ns, ctx.params.bmode == pa.B_STORED)
    if plan[pa.STAGE_FFT_AXIS3] == 0:
        names[pa.STAGE_UPDATE] = ("k_row_il_fixed" if fixed else "k_row_step") + "(N3 inv + update + N3 fwd)"
    if plan[pa.STAGE_FFT_AXIS1] == 0:
        names[pa.STAGE_CONTRACT] = ("k_col_il_fixed" if fixed else "k_col_contract") + "(N1 fwd + B(k) + N1 inv)"
    out = {}
    for st, (ms, n) in ctx.time_step_kernels(nsteps).items():
        out[names[st]] = (st, n, ms / n)
    return out


def copy_reference(ctx, reps=10):
    """The plain copy kernel xi -> W (bandwidth reference, and the calibration
    kernel of the PMC traffic runs).  Leaves the work array overwritten."""
    ctx.enqueue_stage(pa.STAGE_COPY, 2)
    e0, e1 = hip.Event(), hip.Event()
    e0.record(ctx.stream)
    ctx.enqueue_stage(pa.STAGE_COPY, reps)
    e1.record(ctx.stream)
    e1.synchronize()
    return e0.elapsed_ms(e1) / reps


def make_ctx(w, precision, bmode):
    ctx = pa.Context(w.params(precision, bmode))
    ctx.set_tau(w.tau)
    ctx.upload_xi(w.xi)
    ctx.begin_outer(True)
    return ctx


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, budget_s=20.0):
    """The FP64 oracle on the same instance, on the cores this process may use
    (at most 16, the host share of one GPU) and on one core."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oraclepy as O
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    cores = max(1, min(16, avail))

    def rate(nthreads):
        p = O.make_params(w.shape, w.eps, C.C11, C.C12, C.C44, [0.1] * w.ns, C.asf2_reference(w.ns),
                          [1.0] * w.ns, nthreads=nthreads)
        o = O.Oracle(p)
        o.set_xi(w.xi)
        o.set_tau(w.tau)
        o.prologue(True)
        t = time.perf_counter()
        o.gl_steps(1)
        t1 = time.perf_counter() - t
        n = int(max(1, min(100, budget_s / max(t1, 1e-6))))
        t = time.perf_counter()
        o.gl_steps(n)
        dt = time.perf_counter() - t
        o.close()
        return n, dt

    n, dt = rate(cores)
    n1, dt1 = rate(1)
    return {"value": n / dt, "unit": "GL steps/s", "cores": cores, "kind": "port",
            "sample": f"FP64 oracle (oracle/pfdd_oracle.c, OpenMP, own Stockham FFT) on the same 128^3x12 c1 "
                      f"instance: {n} GL steps after 1 warm-up step, {dt:.1f} s on {cores} threads",
            "single_core": {"value": n1 / dt1, "unit": "GL steps/s", "cores": 1,
                            "sample": f"{n1} GL steps after 1 warm-up step, {dt1:.1f} s"},
            "host": {"nproc": os.cpu_count(), "affinity_cores": avail, "cpu_model": cpu_model()}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--bmode", choices=["recompute", "stored"], default="recompute")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-error-run", action="store_true",
                    help="skip the 1000-step error comparison of the mixed modes (profiling runs)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: the PFDD GL step is a single-GPU path here "
              "(the 3D-FFT all-to-all sharding is not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if hip.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    bmode = pa.B_RECOMPUTE if a.bmode == "recompute" else pa.B_STORED
    stored = bmode == pa.B_STORED
    w = C.c1(bmode)
    S, P = w.ns, w.P
    ntri = S * (S + 1) // 2

    # ---- FP64 (value)
    progress("FP64 timed steps")
    ctx = make_ctx(w, pa.FP64, bmode)
    ms64, wall64, wu = timed_steps(ctx, a.steps, a.warmup)
    stages = time_kernels(ctx)
    copy_ms = copy_reference(ctx)
    plan = ctx.step_plan()
    ctx.close()
    sps64 = a.steps / (ms64 / 1e3)

    # dominant kernel of the FP64 step and its roofline
    dom = max(((n, st, k, ms) for n, (st, k, ms) in stages.items() if k > 0 and st != pa.STAGE_FINISH),
              key=lambda t: t[2] * t[3])
    dname, dst, _, dms = dom
    dbytes = stage_bytes(dst, S, P, ntri, True, stored)
    achieved = dbytes / (dms / 1e3) / 1e9
    rname = rocprof_name(dst, plan, True, w.shape, S, stored)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(rname), "kernel": dname,
                "rocprof_kernel": rname, "alg_bytes_per_launch": dbytes, "avg_launch_ms": round(dms, 5)}
    step_bytes = step_alg_bytes(S, P, True, stored)
    step_gbs = step_bytes * sps64 / 1e9
    kernels = {n: {"avg_ms": round(ms, 5), "launches_per_step": k,
                   "GB_s": round(stage_bytes(st, S, P, ntri, True, stored) / (ms / 1e3) / 1e9, 1)}
               for n, (st, k, ms) in stages.items()}
    kernels["k_xi_to_w(copy reference)"] = {"avg_ms": round(copy_ms, 5), "launches_per_step": 0,
                                            "GB_s": round(stage_bytes(pa.STAGE_COPY, S, P, ntri, True, stored)
                                                          / (copy_ms / 1e3) / 1e9, 1)}

    out = {"metric": METRIC, "value": round(sps64, 2), "unit": "GL steps/s", "n_gpus": 1, "steps": a.steps,
           "warmup": wu, "ms_per_step": round(ms64 / a.steps, 5), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic: seeded loops (SplitMix64, seed 20261202+1), sigma_ij ~ U[-0.02,0.02]",
           "config": {"workload": "c1: 128^3 grid x 12 FCC slip systems, 16 seeded circular loops r~U[5,20], "
                                  "FP64 GL step (FFT + B(k) contraction + update)",
                      "grid": [128, 128, 128], "ns": S, "precision": "fp64", "bmode": a.bmode,
                      "parallelism": "single GPU"},
           "roofline": roofline,
           "step_roofline": {"alg_bytes_per_step": step_bytes, "achieved_GB_s": round(step_gbs, 1),
                             "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
           "gridpoints_x_systems_per_s": round(P * S * sps64, 1),
           "kernels_fp64": kernels, "wall_s_fp64": round(wall64, 3)}

    # ---- mixed precision: compensated FP32 master ("mixed"), plain FP32 master ("mixed_lean")
    modes = [] if a.no_mixed else [("mixed", pa.MIXED, False), ("mixed_lean", pa.MIXED_LEAN, True)]
    for key, prec, lean in modes:
        progress(f"{key} timed steps")
        ctxm = make_ctx(w, prec, bmode)
        ms32, _, _ = timed_steps(ctxm, a.steps, a.warmup)
        stages32 = time_kernels(ctxm)
        ctxm.close()
        sps32 = a.steps / (ms32 / 1e3)
        out[key] = {"value": round(sps32, 2), "unit": "GL steps/s", "speedup_vs_fp64": round(sps32 / sps64, 3),
                    "ms_per_step": round(ms32 / a.steps, 5),
                    "master_xi": "plain FP32 complex, 8 B" if lean else "compensated FP32 (hi + lo, im), 12 B",
                    "step_roofline_frac": round(step_alg_bytes(S, P, False, stored, lean) * sps32 / 1e9
                                                / HBM_PEAK_GBS, 4),
                    "kernels": {n: {"avg_ms": round(ms, 5), "launches_per_step": k,
                                    "GB_s": round(stage_bytes(st, S, P, ntri, False, stored, lean)
                                                  / (ms / 1e3) / 1e9, 1)}
                                for n, (st, k, ms) in stages32.items()}}

    # ---- errors of the mixed modes against FP64 over the workload's full 1000 GL steps (fresh
    # contexts, outside the timed region): xi, total slipped area, mean stress
    if modes and not a.no_error_run:
        progress("error runs")

        def run_full(prec):
            c = make_ctx(w, prec, bmode)
            g, _ = c.gl_steps(w.steps)
            x = c.download_xi()
            _, sig = c.strain_stress(w.s_tensor_avepst())
            c.close()
            return x, g, sig

        xi64, g64, sig64 = run_full(pa.FP64)
        area64 = float(xi64.real.sum())
        for key, prec, _ in modes:
            xi32, g32, sig32 = run_full(prec)
            out[key].update({
                "xi_max_abs_err_vs_fp64": float(np.abs(xi32 - xi64).max()),
                "slipped_area_rel_err_vs_fp64": abs(float(xi32.real.sum()) - area64) / abs(area64),
                "gamma_max_rel_err_vs_fp64": float(np.abs(g32 - g64).max() / np.abs(g64).max()),
                "mean_stress_max_abs_err_vs_fp64": float(np.abs(sig32 - sig64).max()),
                "mean_stress_rel_err_vs_fp64": float(np.abs(sig32 - sig64).max() / np.abs(sig64).max()),
                "error_after_steps": w.steps})
            del xi32
        del xi64

    if not a.no_cpu_baseline:
        progress("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
