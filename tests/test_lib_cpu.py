"""CPU-side checks of libpfdd_gpu.so: it loads, exports every entry point the
header declares, its host setup helpers match the compiled reference, and its
argument validation returns the documented codes (no GPU work is issued)."""
import ctypes
import os
import re

import numpy as np
import pytest

import pfdd_amd as pa
import physics as ph
import refpy as R
from pfdd_amd import configs as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "pfdd_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b((?:pfdd_|fft_)[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_function():
    L = pa.load()
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert sorted(pa.EXPORTED) == names


def header_constants():
    """Every `PFDD_NAME = value` of the header's enums, evaluated."""
    txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "pfdd_gpu.h")).read(), flags=re.S)
    out = {}
    for name, expr in re.findall(r"\b(PFDD_[A-Z0-9_]+)\s*=\s*(-?[0-9]+(?:\s*<<\s*[0-9]+)?)", txt):
        out[name[5:]] = eval(expr)
    return out


def test_header_constants_match_package():
    """The package's constants are the header's (precision modes, B modes,
    plan bits, stages, ensemble modes, error codes)."""
    h = header_constants()
    assert len(h) >= 35
    checked = 0
    for name, value in h.items():
        if hasattr(pa, name):
            assert getattr(pa, name) == value, name
            checked += 1
    for name in ("FP64", "MIXED", "MIXED_LEAN", "B_STORED", "B_RECOMPUTE", "PLAN_INTERLEAVED",
                 "PLAN_FIELD_MAJOR", "STAGE_FINISH", "ENS_PERSISTENT"):
        assert name in h and hasattr(pa, name), name
    assert checked >= 30


def test_strerror_names_codes():
    for rc in (pa.OK, pa.E_HIP, pa.E_ALLOC, pa.E_SHAPE, pa.E_RADIX, pa.E_NAN, pa.E_UNSUPPORTED, pa.E_STATE,
               pa.E_LIMIT):
        assert pa.strerror(rc) and pa.strerror(rc) != "unknown error"


def _create_rc(p):
    h = ctypes.c_void_p()
    rc = pa.load().pfdd_create(ctypes.byref(p), ctypes.byref(h))
    assert not h.value
    return rc


def test_create_validates_before_touching_the_device():
    eps = ph.fcc_eps([0, 1, 2])
    base = dict(shape=(32, 32, 32), eps=eps, C11=ph.C11, C12=ph.C12, C44=ph.C44, cd2=[0.1] * 3,
                asf2=ph.asf2_reference(3), dslip=[1.0] * 3)
    assert _create_rc(pa.make_params(**{**base, "shape": (7, 32, 32)})) == pa.E_RADIX
    assert _create_rc(pa.make_params(**{**base, "shape": (32, 1, 32)})) == pa.E_RADIX
    assert _create_rc(pa.make_params(**{**base, "shape": (32, 32, 1024)})) == pa.E_RADIX
    p = pa.make_params(**base)
    p.ns = 0
    assert _create_rc(p) == pa.E_SHAPE
    p = pa.make_params(**base)
    p.ns = 17
    assert _create_rc(p) == pa.E_SHAPE
    assert _create_rc(pa.make_params(**{**base, "C11": float("nan")})) == pa.E_NAN
    assert _create_rc(pa.make_params(**{**base, "dslip": [1.0, 0.0, 1.0]})) == pa.E_NAN
    for bad in (7, 4, -1):
        p = pa.make_params(**base)
        p.precision = bad
        assert _create_rc(p) == pa.E_SHAPE
    # PFDD_REF32: fourn's power-of-two edges and the unfused field-major kernels only
    assert _create_rc(pa.make_params(**{**base, "shape": (32, 48, 32), "precision": pa.REF32})) == pa.E_RADIX
    assert _create_rc(pa.make_params(**{**base, "shape": (32, 12, 32), "precision": pa.REF32})) == pa.E_RADIX
    for plan in (pa.PLAN_INTERLEAVED, pa.PLAN_PLANE, pa.PLAN_FUSED_COLS, pa.plan_row_tile(2)):
        assert _create_rc(pa.make_params(**{**base, "precision": pa.REF32, "plan": plan})) == pa.E_SHAPE


@pytest.mark.parametrize("shape,pp", [((32, 2, 32), None), ((128, 128, 128), None), ((16, 12, 8), (3, 5, 6)),
                                       ((8, 8, 8), (0, 0, 0)), ((8, 8, 8), (9, 0, 0))])
def test_line_points_match_brute_force(shape, pp):
    """pfdd_line_points against the predicate of avestrain (G.c:707) by brute force."""
    ppoint = C.interface_ppoint(shape) if pp is None else pp
    n = C.INTERFACE_N
    want = [(k1 * shape[1] + k2) * shape[2] + k3
            for k1 in range(shape[0]) for k2 in range(shape[1]) for k3 in range(shape[2])
            if k1 == ppoint[0] and abs(n[0] * (k1 - ppoint[0]) + n[1] * (k2 - ppoint[1]) + n[2] * (k3 - ppoint[2])) <= 1.0]
    got = C.line_points(shape, ppoint=ppoint)
    assert got.tolist() == want
    if shape == (128, 128, 128):
        # |sqrt(6)/3 (k3 - 96)| <= 1 holds for k3 = 95, 96, 97 on every k2
        assert len(got) == 3 * 128


def test_line_point_calls_validate_before_touching_the_device():
    L = pa.load()
    assert L.pfdd_set_line_points(None, None, 0) == pa.E_SHAPE
    assert L.pfdd_line_sums(None, None) == pa.E_SHAPE
    assert L.pfdd_gl_steps_lines(None, 1, None, None, None) == pa.E_SHAPE


def test_splitmix64_known_answers():
    """First outputs of SplitMix64 for seed 1234567 (Vigna's reference implementation)."""
    r = C.Rng(1234567)
    assert [r.next64() for _ in range(5)] == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                             4593380528125082431, 16408922859458223821]
    u = C.Rng(1).uniform()
    assert 0.0 <= u < 1.0


def test_setfcc_geometry():
    xn, xb = C.setfcc()
    s3 = float(np.float32(np.sqrt(3)))
    assert np.all(np.abs(np.abs(xn) - 1 / s3) == 0)          # float sqrt3 as G.c:1820
    assert np.abs(np.einsum("ai,ai->a", xn, xb)).max() < 1e-15
    assert np.allclose(np.linalg.norm(xb, axis=1), 1 / np.sqrt(2))
    assert len({tuple(r) for r in np.round(np.hstack([xn, xb]), 6)}) == 12
    assert np.array_equal(C.seteps(xn, xb, [1.0] * 12, np.eye(3)), ph.fcc_eps(range(12)))


@pytest.mark.skipif(not R.available(), reason="compiled reference oracle/_ref not built")
def test_setorient_seteps_resolve_match_reference():
    th_ref = R.setorient()
    rows = [[np.sqrt(3.) / 3., np.sqrt(3.) / 3., np.sqrt(3.) / 3.], [-np.sqrt(2.) / 2., np.sqrt(2.) / 2., 0.0],
            [-np.sqrt(6.) / 6., -np.sqrt(6.) / 6., np.sqrt(6.) / 3.]]
    th = C.setorient(rows)
    assert np.array_equal(th, th_ref[0])
    xn, xb = R.set3D1sys()
    eps_ref, _ = R.seteps(xn, xb, [1.0, 1.0], th_ref)
    eps = C.seteps(xn, xb, [1.0, 1.0], th)
    assert np.array_equal(eps, eps_ref)
    # log.ref:247-254: both eigenstrains are [[0,0,0],[0,0,0],[-1,0,0]] to 6 digits
    assert np.allclose(eps, [[[0, 0, 0], [0, 0, 0], [-1, 0, 0]]] * 2, atol=5e-7)
    sig = np.array([[0.003, 0.01, -0.002], [0.01, 0.0, 0.004], [-0.002, 0.004, -0.006]])
    tau_ref = R.resolSS(np.tile(sig, (R.P, 1, 1)), eps_ref)
    assert np.array_equal(C.resolve_tau(sig, eps), tau_ref[0])
    assert np.array_equal(tau_ref[0], tau_ref[-1])


def test_loop_voxelisation_closed_forms():
    # a loop in a grid plane: the Gauss circle count N(8) = 197
    xi = np.zeros((1, 32, 32, 32), dtype=np.complex128)
    assert C.add_loop(xi, 0, (0, 0, 1), (16, 16, 16), 8.0) == 197
    k = np.nonzero(xi[0].real)[2]
    assert np.all(k == 16)
    # periodic wrap: the same disc centred on a corner has the same count
    xi2 = np.zeros((1, 32, 32, 32), dtype=np.complex128)
    assert C.add_loop(xi2, 0, (0, 0, 1), (0, 0, 0), 8.0) == 197
    # a (111) loop: every point sits on one lattice plane x+y+z = m
    xi3 = np.zeros((1, 40, 40, 40), dtype=np.complex128)
    n = C.add_loop(xi3, 0, (1, 1, 1), (20.2, 19.7, 20.4), 6.0)
    i, j, kk = np.nonzero(xi3[0].real)
    assert n == len(i) > 0
    assert len(set(i + j + kk)) == 1
    pts = np.stack([i, j, kk], 1) - np.array([20.2, 19.7, 20.4])
    nh = np.ones(3) / np.sqrt(3)
    inplane = pts - np.outer(pts @ nh, nh)
    assert np.linalg.norm(inplane, axis=1).max() <= 6.0


def test_workloads_are_reproducible():
    a, b = C.c0(), C.c0()
    assert np.array_equal(a.xi, b.xi) and np.array_equal(a.tau, b.tau)
    assert a.tau[0] == pytest.approx(0.01, abs=1e-15)
    w = C.small_fcc((16, 16, 16))
    assert w.ns == 12 and w.xi.real.sum() > 0


def test_driver_rejects_bad_arguments():
    import subprocess
    exe = os.path.join(ROOT, "pfdd_amd", "pfdd")
    if not os.path.exists(exe):
        pytest.skip("driver not built")
    r = subprocess.run([exe, "nosuchpreset"], capture_output=True, text=True)
    assert r.returncode == 2 and "unknown preset" in r.stderr
    r = subprocess.run([exe, "c0", "frobnicate=3"], capture_output=True, text=True)
    assert r.returncode == 2 and "unknown key" in r.stderr
    r = subprocess.run([exe, "c0", "plan=%d" % (1 << 24)], capture_output=True, text=True)
    assert r.returncode != 0 and "pfdd_create" in r.stderr


def test_driver_restart_rejects_a_short_dump(tmp_path):
    """restart=1 reads xi.dat / xi_bc.dat before any device work and refuses a
    dump shorter than [NS+9][N1][N2][N3][2] doubles."""
    import subprocess
    exe = os.path.join(ROOT, "pfdd_amd", "pfdd")
    if not os.path.exists(exe):
        pytest.skip("driver not built")
    np.zeros(10).tofile(tmp_path / "xi.dat")
    np.zeros(10).tofile(tmp_path / "xi_bc.dat")
    r = subprocess.run([exe, "c0", "steps=2", f"outdir={tmp_path}", "restart=1"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "holds 10 of the 786432 doubles" in r.stderr


def test_fft_seam_requires_init():
    L = pa.load()
    L.pfdd_fft_host_release()
    a = np.zeros(8, dtype=np.complex64)
    L.fft_forward(a.ctypes.data, 1)
    assert L.pfdd_fft_host_status() == pa.E_STATE
    assert L.pfdd_fft_host_noalloc(a.ctypes.data, 1, pa.FORWARD) == pa.E_STATE


def _captured_stdout(fn):
    """Runs fn() with the C-level stdout (fd 1) sent to a temporary file and
    returns (fn's result, the bytes written)."""
    import sys
    import tempfile
    libc = ctypes.CDLL(None)
    sys.stdout.flush()
    libc.fflush(None)
    saved = os.dup(1)
    with tempfile.TemporaryFile() as tmp:
        os.dup2(tmp.fileno(), 1)
        try:
            out = fn()
            libc.fflush(None)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        tmp.seek(0)
        return out, tmp.read()


def _gamma_sequences():
    """gamma sequences that take plasticconverge through its branches: a
    converging series with one outlier (the run counter resets), a series
    that converges for good (the loop ends), gamma == 0 throughout, and
    gamma crossing zero."""
    s = np.arange(400)
    conv = 0.37 + 1e-2 * 0.5 ** (s / 3.0)
    outlier = conv.copy()
    outlier[70] += 1e-6
    zero = np.zeros(60)
    cross = np.concatenate([np.linspace(1e-3, 0.0, 20), np.zeros(40), np.full(40, 1e-13)])
    return [conv, outlier, zero, cross]


@pytest.mark.skipif(not R.available(), reason="oracle/_ref not built")
def test_plastic_converge_matches_reference_log():
    """pfdd_plastic_converge (used by the driver) against the compiled
    plasticconverge (G.c:3732-3785): the same flags, counters and marks, and
    byte-identical log lines, with the line printed after the counter update."""
    L = pa.load()
    ref = R.lib()
    ref.plasticconverge.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    ref.plasticconverge.restype = ctypes.c_int
    ends = 0
    for seq in _gamma_sequences():
        tp = [np.zeros(len(seq), dtype=np.int32) for _ in range(2)]
        cnt = [ctypes.c_int(0), ctypes.c_int(0)]
        logs = [b"", b""]
        last = 0.0
        for it, g in enumerate(seq):
            flags = []
            for k, fn in enumerate((ref.plasticconverge, L.pfdd_plastic_converge)):
                f, out = _captured_stdout(lambda: fn(float(g), last, it, tp[k].ctypes.data, ctypes.byref(cnt[k])))
                flags.append(f)
                logs[k] += out
            assert flags[0] == flags[1], (it, g)
            assert cnt[0].value == cnt[1].value, (it, g)
            last = float(g)
            if flags[0] == 0:
                ends += 1
                break
        assert np.array_equal(tp[0], tp[1])
        assert logs[0] == logs[1]
        assert logs[0].count(b"\n") > 10
    assert ends >= 2
