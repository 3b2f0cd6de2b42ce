"""pfdd_amd -- Python view of libpfdd_gpu.so, the MI355X (gfx950) PFDD
FFT-elasticity + Ginzburg-Landau step.

The product is the C ABI in include/pfdd_gpu.h; this module binds it with
ctypes for tests, the benchmark and scripting.  It never falls back to CPU
code: if the library is missing, load() raises.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libpfdd_gpu.so")

MAX_NS = 16
OK, E_HIP, E_ALLOC, E_SHAPE, E_RADIX, E_NAN, E_UNSUPPORTED, E_STATE, E_LIMIT = 0, -1, -2, -3, -4, -5, -6, -7, -8
NOALLOC_MAX_POINTS = 1 << 21
FP64, MIXED, MIXED_LEAN, REF32 = 0, 1, 2, 3
B_STORED, B_RECOMPUTE = 0, 1
FORWARD, BACKWARD = -1, 1
PLAN_AUTO, PLAN_UNFUSED_ROWS, PLAN_UNFUSED_COLS, PLAN_FUSED_COLS, PLAN_NO_PLANE, PLAN_PLANE = 0, 1, 2, 4, 8, 16
PLAN_ROW_TILE_SHIFT, PLAN_COL_TILE_SHIFT = 8, 12
PLAN_RADIX4_AXIS1, PLAN_RADIX4_AXIS2, PLAN_RADIX4_AXIS3 = 1 << 16, 1 << 17, 1 << 18
PLAN_INTERLEAVED, PLAN_FIELD_MAJOR = 1 << 19, 1 << 20


def plan_row_tile(tile):
    """Plan bits that fix the fused row kernel's tile (a power of two)."""
    return (int(tile).bit_length()) << PLAN_ROW_TILE_SHIFT


def plan_col_tile(tile):
    """Plan bits that fix the fused column kernel's tile (a power of two)."""
    return (int(tile).bit_length()) << PLAN_COL_TILE_SHIFT


def describe_plan(plan):
    names = [n for n, v in (("unfused_rows", 1), ("unfused_cols", 2), ("fused_cols", 4), ("no_plane", 8),
                            ("plane", 16), ("radix4_n1", 1 << 16), ("radix4_n2", 1 << 17),
                            ("radix4_n3", 1 << 18), ("interleaved", 1 << 19), ("field_major", 1 << 20)) if plan & v]
    rt, ct = (plan >> PLAN_ROW_TILE_SHIFT) & 15, (plan >> PLAN_COL_TILE_SHIFT) & 15
    if rt:
        names.append(f"row_tile={1 << (rt - 1)}")
    if ct:
        names.append(f"col_tile={1 << (ct - 1)}")
    return "+".join(names) if names else "auto"


EXPORTED = [
    "pfdd_create", "pfdd_destroy", "pfdd_set_tau", "pfdd_upload_xi", "pfdd_download_xi",
    "pfdd_begin_outer", "pfdd_gl_steps", "pfdd_strain_stress", "pfdd_fft", "pfdd_sync", "pfdd_stream",
    "pfdd_device_bytes", "pfdd_strerror", "pfdd_setfcc", "pfdd_setorient", "pfdd_seteps",
    "pfdd_resolve_tau", "pfdd_splitmix64", "pfdd_uniform01", "pfdd_add_loop", "pfdd_enqueue_stage",
    "pfdd_step_plan", "pfdd_seed_loops", "pfdd_seed_sigma", "pfdd_fft_host_init", "fft_forward", "fft_backward",
    "pfdd_fft_host_status", "pfdd_fft_host_release", "pfdd_set_tau_ramp", "pfdd_set_plan", "pfdd_get_plan",
    "pfdd_autotune", "pfdd_create_ensemble", "pfdd_destroy_ensemble", "pfdd_ensemble_info",
    "pfdd_ensemble_set_tau", "pfdd_ensemble_upload_xi", "pfdd_ensemble_download_xi", "pfdd_ensemble_begin_outer",
    "pfdd_ensemble_gl_steps", "pfdd_ensemble_autotune", "pfdd_ensemble_set_plan", "pfdd_ensemble_sync",
    "pfdd_ensemble_stream", "pfdd_ensemble_device_bytes", "pfdd_time_step_kernels", "pfdd_ensemble_set_mode",
    "pfdd_ensemble_get_mode", "pfdd_set_line_points", "pfdd_line_sums", "pfdd_gl_steps_lines", "pfdd_line_points",
    "pfdd_plastic_converge", "pfdd_fft_host_noalloc",
]
ENS_PHASED, ENS_PERSISTENT = 0, 1
(STAGE_UPDATE, STAGE_FFT_AXIS3, STAGE_FFT_AXIS2, STAGE_FFT_AXIS1, STAGE_CONTRACT, STAGE_STEP, STAGE_COPY,
 STAGE_FINISH) = range(8)


class PfddError(RuntimeError):
    def __init__(self, rc, what):
        self.rc = rc
        super().__init__(f"{what}: {strerror(rc)} ({rc})")


class Params(ctypes.Structure):
    _fields_ = [
        ("n1", ctypes.c_int), ("n2", ctypes.c_int), ("n3", ctypes.c_int), ("ns", ctypes.c_int),
        ("precision", ctypes.c_int), ("bmode", ctypes.c_int),
        ("C11", ctypes.c_double), ("C12", ctypes.c_double), ("C44", ctypes.c_double),
        ("d1", ctypes.c_double), ("d2", ctypes.c_double), ("d3", ctypes.c_double),
        ("cd2", ctypes.c_double * MAX_NS), ("asf2", ctypes.c_double * MAX_NS),
        ("dslip", ctypes.c_double * MAX_NS),
        ("uq2", ctypes.c_double), ("pi_const", ctypes.c_double),
        ("eps", ctypes.c_double * (MAX_NS * 9)),
        ("plan", ctypes.c_int),
    ]


_lib = None


def load():
    """Load libpfdd_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, ci, cd, cl = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long
    L.pfdd_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(vp)]
    L.pfdd_destroy.argtypes = [vp]
    L.pfdd_destroy.restype = None
    L.pfdd_set_tau.argtypes = [vp, vp, ci]
    L.pfdd_set_tau_ramp.argtypes = [vp, vp, vp]
    L.pfdd_upload_xi.argtypes = [vp, vp, vp, ci]
    L.pfdd_download_xi.argtypes = [vp, vp, ci]
    L.pfdd_begin_outer.argtypes = [vp, ci]
    L.pfdd_gl_steps.argtypes = [vp, ci, vp, vp]
    L.pfdd_strain_stress.argtypes = [vp, vp, vp, vp, vp, vp]
    L.pfdd_fft.argtypes = [vp, vp, ci, ci, ci]
    L.pfdd_sync.argtypes = [vp]
    L.pfdd_stream.argtypes = [vp]
    L.pfdd_stream.restype = vp
    L.pfdd_device_bytes.argtypes = [vp]
    L.pfdd_device_bytes.restype = ctypes.c_longlong
    L.pfdd_strerror.argtypes = [ci]
    L.pfdd_strerror.restype = ctypes.c_char_p
    L.pfdd_setfcc.argtypes = [vp, vp]
    L.pfdd_setfcc.restype = None
    L.pfdd_setorient.argtypes = [vp, vp]
    L.pfdd_setorient.restype = None
    L.pfdd_seteps.argtypes = [ci, vp, vp, vp, vp, vp]
    L.pfdd_seteps.restype = None
    L.pfdd_resolve_tau.argtypes = [ci, vp, vp, vp]
    L.pfdd_resolve_tau.restype = None
    L.pfdd_splitmix64.argtypes = [ctypes.POINTER(ctypes.c_ulonglong)]
    L.pfdd_splitmix64.restype = ctypes.c_ulonglong
    L.pfdd_uniform01.argtypes = [ctypes.POINTER(ctypes.c_ulonglong)]
    L.pfdd_uniform01.restype = cd
    L.pfdd_add_loop.argtypes = [vp, ci, ci, ci, ci, vp, vp, cd]
    L.pfdd_add_loop.restype = cl
    L.pfdd_enqueue_stage.argtypes = [vp, ci, ci]
    L.pfdd_step_plan.argtypes = [vp, vp]
    L.pfdd_seed_loops.argtypes = [vp, ci, ci, ci, ci, vp, ci, cd, cd, ctypes.POINTER(ctypes.c_ulonglong), vp]
    L.pfdd_seed_sigma.argtypes = [cd, ctypes.POINTER(ctypes.c_ulonglong), vp]
    L.pfdd_seed_sigma.restype = None
    L.pfdd_fft_host_init.argtypes = [ci, ci, ci]
    L.fft_forward.argtypes = [vp, ci]
    L.fft_forward.restype = None
    L.fft_backward.argtypes = [vp, ci]
    L.fft_backward.restype = None
    L.pfdd_fft_host_status.argtypes = []
    L.pfdd_fft_host_release.argtypes = []
    L.pfdd_fft_host_release.restype = None
    L.pfdd_fft_host_noalloc.argtypes = [vp, ci, ci]
    fp = ctypes.POINTER(ctypes.c_float)
    L.pfdd_time_step_kernels.argtypes = [vp, ci, vp, vp]
    L.pfdd_set_line_points.argtypes = [vp, vp, ci]
    L.pfdd_line_sums.argtypes = [vp, vp]
    L.pfdd_gl_steps_lines.argtypes = [vp, ci, vp, vp, vp]
    L.pfdd_line_points.argtypes = [ci, ci, ci, vp, vp, vp, ci]
    L.pfdd_set_plan.argtypes = [vp, ci]
    L.pfdd_plastic_converge.argtypes = [cd, cd, ci, vp, vp]
    L.pfdd_get_plan.argtypes = [vp]
    L.pfdd_autotune.argtypes = [vp, ci, vp, fp]
    L.pfdd_create_ensemble.argtypes = [vp, ci, ctypes.POINTER(vp)]
    L.pfdd_destroy_ensemble.argtypes = [vp]
    L.pfdd_destroy_ensemble.restype = None
    L.pfdd_ensemble_info.argtypes = [vp, vp, vp, vp]
    L.pfdd_ensemble_set_tau.argtypes = [vp, ci, vp, ci]
    L.pfdd_ensemble_upload_xi.argtypes = [vp, ci, vp, vp, ci]
    L.pfdd_ensemble_download_xi.argtypes = [vp, ci, vp, ci]
    L.pfdd_ensemble_begin_outer.argtypes = [vp, ci]
    L.pfdd_ensemble_gl_steps.argtypes = [vp, ci, vp]
    L.pfdd_ensemble_autotune.argtypes = [vp, ci, vp, vp]
    L.pfdd_ensemble_set_plan.argtypes = [vp, ci, ci]
    L.pfdd_ensemble_sync.argtypes = [vp]
    L.pfdd_ensemble_stream.argtypes = [vp]
    L.pfdd_ensemble_stream.restype = vp
    L.pfdd_ensemble_device_bytes.argtypes = [vp]
    L.pfdd_ensemble_device_bytes.restype = ctypes.c_longlong
    L.pfdd_ensemble_set_mode.argtypes = [vp, ci]
    L.pfdd_ensemble_get_mode.argtypes = [vp, fp, fp]
    _lib = L
    return L


def strerror(rc):
    return load().pfdd_strerror(rc).decode()


def _check(rc, what):
    if rc != OK:
        raise PfddError(rc, what)


def make_params(shape, eps, C11, C12, C44, cd2, asf2, dslip, precision=FP64, bmode=B_RECOMPUTE,
                d=(1.0, 1.0, 1.0), uq2=1.0, pi_const=3.141592654, plan=PLAN_AUTO):
    p = Params()
    p.n1, p.n2, p.n3 = (int(v) for v in shape)
    eps = np.asarray(eps, dtype=np.float64).reshape(-1, 9)
    p.ns = eps.shape[0]
    p.precision, p.bmode = precision, bmode
    p.C11, p.C12, p.C44 = C11, C12, C44
    p.d1, p.d2, p.d3 = d
    for a in range(p.ns):
        p.cd2[a], p.asf2[a], p.dslip[a] = cd2[a], asf2[a], dslip[a]
        for t in range(9):
            p.eps[9 * a + t] = eps[a, t]
    p.uq2, p.pi_const = uq2, pi_const
    p.plan = plan
    return p


class Context:
    """One pfdd_ctx: device state of one PFDD instance on the current GPU."""

    def __init__(self, params):
        L = load()
        self.params = params
        self.ns = params.ns
        self.shape = (params.n1, params.n2, params.n3)
        self.P = params.n1 * params.n2 * params.n3
        h = ctypes.c_void_p()
        _check(L.pfdd_create(ctypes.byref(params), ctypes.byref(h)), "pfdd_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load().pfdd_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tau(self, tau, per_point=False):
        t = np.ascontiguousarray(tau, dtype=np.float64)
        _check(load().pfdd_set_tau(self.h, t.ctypes.data, 1 if per_point else 0), "pfdd_set_tau")

    def set_tau_ramp(self, tau0, rate):
        t = np.ascontiguousarray(tau0, dtype=np.float64)
        r = np.ascontiguousarray(rate, dtype=np.float64)
        _check(load().pfdd_set_tau_ramp(self.h, t.ctypes.data, r.ctypes.data), "pfdd_set_tau_ramp")

    def upload_xi(self, xi, xi_bc=None):
        """xi: complex [nsv, N1, N2, N3] (nsv >= ns; fields past ns must be zero)."""
        x = np.ascontiguousarray(xi, dtype=np.complex128)
        nsv = x.shape[0]
        b = None if xi_bc is None else np.ascontiguousarray(xi_bc, dtype=np.complex128)
        _check(load().pfdd_upload_xi(self.h, x.ctypes.data, None if b is None else b.ctypes.data, nsv),
               "pfdd_upload_xi")

    def download_xi(self, nsv=None):
        nsv = self.ns if nsv is None else nsv
        x = np.zeros((nsv,) + self.shape, dtype=np.complex128)
        _check(load().pfdd_download_xi(self.h, x.ctypes.data, nsv), "pfdd_download_xi")
        return x

    def begin_outer(self, quirk=True):
        _check(load().pfdd_begin_outer(self.h, 1 if quirk else 0), "pfdd_begin_outer")

    def gl_steps(self, n, record=True):
        if not record:
            _check(load().pfdd_gl_steps(self.h, n, None, None), "pfdd_gl_steps")
            return None, None
        g = np.zeros(n)
        ga = np.zeros((n, self.ns))
        _check(load().pfdd_gl_steps(self.h, n, g.ctypes.data, ga.ctypes.data), "pfdd_gl_steps")
        return g, ga

    def set_line_points(self, pts):
        """Point list (flat indices) whose Re xi sums every GL step reports; [] clears it."""
        q = np.ascontiguousarray(pts, dtype=np.int64).reshape(-1)
        _check(load().pfdd_set_line_points(self.h, q.ctypes.data if q.size else None, int(q.size)),
               "pfdd_set_line_points")

    def line_sums(self):
        """Sums of Re xi_a over the point list for the current xi: [ns]."""
        out = np.zeros(self.ns)
        _check(load().pfdd_line_sums(self.h, out.ctypes.data), "pfdd_line_sums")
        return out

    def gl_steps_lines(self, n):
        """gl_steps that also returns the line sums after every step: (gamma, gamma_a, lines [n][ns])."""
        g = np.zeros(n)
        ga = np.zeros((n, self.ns))
        ln = np.zeros((n, self.ns))
        _check(load().pfdd_gl_steps_lines(self.h, n, g.ctypes.data, ga.ctypes.data, ln.ctypes.data),
               "pfdd_gl_steps_lines")
        return g, ga, ln

    def strain_stress(self, avepst_uniform, fields=False):
        ap = np.ascontiguousarray(avepst_uniform, dtype=np.float64).reshape(9)
        aveps, avesig = np.zeros(9), np.zeros(9)
        ef = sf = None
        if fields:
            ef = np.zeros((3, 3) + self.shape)
            sf = np.zeros((3, 3) + self.shape)
        _check(load().pfdd_strain_stress(self.h, ap.ctypes.data, aveps.ctypes.data, avesig.ctypes.data,
                                         None if ef is None else ef.ctypes.data,
                                         None if sf is None else sf.ctypes.data), "pfdd_strain_stress")
        if fields:
            return aveps.reshape(3, 3), avesig.reshape(3, 3), ef, sf
        return aveps.reshape(3, 3), avesig.reshape(3, 3)

    def fft(self, dev_ptr, batch, direction, precision):
        _check(load().pfdd_fft(self.h, dev_ptr, batch, direction, precision), "pfdd_fft")

    def sync(self):
        _check(load().pfdd_sync(self.h), "pfdd_sync")

    def enqueue_stage(self, stage, reps):
        _check(load().pfdd_enqueue_stage(self.h, stage, reps), "pfdd_enqueue_stage")

    def step_plan(self):
        a = (ctypes.c_int * 8)()
        _check(load().pfdd_step_plan(self.h, a), "pfdd_step_plan")
        return list(a)

    def time_step_kernels(self, nsteps=10):
        """Per stage: (ms per step, launches per step) from events between the
        launches of nsteps eager GL steps (the state advances by nsteps)."""
        ms = np.zeros(8, dtype=np.float32)
        n = np.zeros(8, dtype=np.int32)
        _check(load().pfdd_time_step_kernels(self.h, nsteps, ms.ctypes.data, n.ctypes.data),
               "pfdd_time_step_kernels")
        return {st: (float(ms[st]), int(n[st])) for st in range(8) if n[st] > 0}

    def set_plan(self, plan):
        _check(load().pfdd_set_plan(self.h, plan), "pfdd_set_plan")

    @property
    def plan(self):
        return load().pfdd_get_plan(self.h)

    def autotune(self, steps=16):
        """Time the candidate plans, apply the fastest; returns (plan, ms per step)."""
        plan = ctypes.c_int()
        ms = ctypes.c_float()
        _check(load().pfdd_autotune(self.h, steps, ctypes.byref(plan), ctypes.byref(ms)), "pfdd_autotune")
        return plan.value, ms.value

    @property
    def stream(self):
        return load().pfdd_stream(self.h)

    @property
    def device_bytes(self):
        return load().pfdd_device_bytes(self.h)


class Ensemble:
    """One pfdd_ens: many independent instances, grouped into batched classes."""

    def __init__(self, params_list):
        L = load()
        self.params = list(params_list)
        self.count = len(self.params)
        arr = (Params * self.count)(*self.params)
        self._arr = arr
        h = ctypes.c_void_p()
        _check(L.pfdd_create_ensemble(arr, self.count, ctypes.byref(h)), "pfdd_create_ensemble")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load().pfdd_destroy_ensemble(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        """(number of classes, class of each instance, plan of each class)."""
        n = ctypes.c_int()
        _check(load().pfdd_ensemble_info(self.h, ctypes.byref(n), None, None), "pfdd_ensemble_info")
        cls = np.zeros(self.count, dtype=np.int32)
        plans = np.zeros(n.value, dtype=np.int32)
        _check(load().pfdd_ensemble_info(self.h, None, cls.ctypes.data, plans.ctypes.data), "pfdd_ensemble_info")
        return n.value, cls, plans

    def set_tau(self, inst, tau, per_point=False):
        t = np.ascontiguousarray(tau, dtype=np.float64)
        _check(load().pfdd_ensemble_set_tau(self.h, inst, t.ctypes.data, 1 if per_point else 0),
               "pfdd_ensemble_set_tau")

    def upload_xi(self, inst, xi, xi_bc=None):
        x = np.ascontiguousarray(xi, dtype=np.complex128)
        b = None if xi_bc is None else np.ascontiguousarray(xi_bc, dtype=np.complex128)
        _check(load().pfdd_ensemble_upload_xi(self.h, inst, x.ctypes.data, None if b is None else b.ctypes.data,
                                              x.shape[0]), "pfdd_ensemble_upload_xi")

    def download_xi(self, inst, nsv=None):
        p = self.params[inst]
        nsv = p.ns if nsv is None else nsv
        x = np.zeros((nsv, p.n1, p.n2, p.n3), dtype=np.complex128)
        _check(load().pfdd_ensemble_download_xi(self.h, inst, x.ctypes.data, nsv), "pfdd_ensemble_download_xi")
        return x

    def begin_outer(self, quirk=True):
        _check(load().pfdd_ensemble_begin_outer(self.h, 1 if quirk else 0), "pfdd_ensemble_begin_outer")

    def gl_steps(self, n, record=True):
        """n GL steps of every instance; returns gamma [n][count] (or None)."""
        if not record:
            _check(load().pfdd_ensemble_gl_steps(self.h, n, None), "pfdd_ensemble_gl_steps")
            return None
        g = np.zeros((n, self.count))
        _check(load().pfdd_ensemble_gl_steps(self.h, n, g.ctypes.data), "pfdd_ensemble_gl_steps")
        return g

    def autotune(self, steps=16):
        """Per class: (plan, ms per step) of the fastest candidate, which is applied."""
        n = self.info()[0]
        plans = np.zeros(n, dtype=np.int32)
        ms = np.zeros(n, dtype=np.float32)
        _check(load().pfdd_ensemble_autotune(self.h, steps, plans.ctypes.data, ms.ctypes.data),
               "pfdd_ensemble_autotune")
        return plans, ms

    def set_plan(self, cls, plan):
        _check(load().pfdd_ensemble_set_plan(self.h, cls, plan), "pfdd_ensemble_set_plan")

    def set_mode(self, mode):
        _check(load().pfdd_ensemble_set_mode(self.h, mode), "pfdd_ensemble_set_mode")

    def mode(self):
        """(mode, ms per step phased, ms per step persistent) of the last autotune comparison."""
        a, b = ctypes.c_float(), ctypes.c_float()
        m = load().pfdd_ensemble_get_mode(self.h, ctypes.byref(a), ctypes.byref(b))
        if m < 0:
            raise PfddError(m, "pfdd_ensemble_get_mode")
        return m, a.value, b.value

    def sync(self):
        _check(load().pfdd_ensemble_sync(self.h), "pfdd_ensemble_sync")

    @property
    def stream(self):
        return load().pfdd_ensemble_stream(self.h)

    @property
    def device_bytes(self):
        return load().pfdd_ensemble_device_bytes(self.h)
