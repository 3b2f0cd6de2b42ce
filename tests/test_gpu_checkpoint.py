"""Checkpoint/restart of the C driver (G.c:4361-4382, 4540-4547, 4777-4788,
3966-3987): the binary dumps xi.dat / xi_bc.dat / xi1.dat / xi_bc1.dat /
FinalXi_w.dat, a restart from them (the reference's choice == 2) and the
xi-only input of input_xi_from_file, checked bit for bit against a straight
run and against the oracle."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import oraclepy as O
import pfdd_amd as pa
from pfdd_amd import configs as C
from pfdd_amd import hip

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(pa.LIB_PATH), "pfdd")
N = 30   # GL steps per plastic outer step


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if hip.device_count() < 1:
        pytest.skip("no GPU visible")


def say(msg):
    print(f"[gpu-checkpoint] {msg}", flush=True)
    sys.stdout.flush()


def driver(outdir, *keys):
    say("driver " + " ".join(keys))
    r = subprocess.run([EXE, "c0", f"steps={N}", "fields=0", "energy=0", f"outdir={outdir}"] + list(keys),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    return r.stdout


def dump(d, name):
    w = C.c0()
    return np.fromfile(os.path.join(d, name), dtype=np.float64).reshape(w.ns + 9, *w.shape, 2)


def oracle_outer_steps(xi0, nouter):
    """nouter plastic outer steps of N GL steps, each after a prologue with the
    virtualevolv term (as the driver runs them), from xi0 [ns, N1, N2, N3]."""
    w = C.c0()
    p = O.make_params(w.shape, w.eps, C.C11, C.C12, C.C44, [0.1] * w.ns, C.asf2_reference(w.ns), [1.0] * w.ns)
    o = O.Oracle(p)
    o.set_xi(xi0)
    o.set_tau(w.tau)
    for _ in range(nouter):
        o.prologue(True)
        o.gl_steps(N)
    x = o.get_xi()
    o.close()
    return x


def test_restart_equals_straight_run(tmp_path):
    """N steps, dump, restart, N more steps == one run of two plastic outer
    steps (2N steps, the virtualevolv term on the first step of each), bit for
    bit; both match the oracle; the itp == 0 copies and the last outer step's
    FinalXi_w.dat equal the final dump."""
    a, b, c = tmp_path / "a", tmp_path / "b", tmp_path / "c"
    for d in (a, b, c):
        d.mkdir()
    driver(a, "nt=3")
    for name in ("xi.dat", "xi_bc.dat", "xi1.dat", "xi_bc1.dat", "FinalXi_w.dat"):
        assert os.path.getsize(a / name) == 8 * 2 * (3 + 9) * 32 ** 3, name
    assert np.array_equal(dump(a, "xi1.dat"), dump(a, "xi.dat"))
    assert np.array_equal(dump(a, "xi_bc1.dat"), dump(a, "xi_bc.dat"))
    assert np.array_equal(dump(a, "FinalXi_w.dat"), dump(a, "xi.dat"))
    shutil.copy(a / "xi.dat", b / "xi.dat")
    shutil.copy(a / "xi_bc.dat", b / "xi_bc.dat")
    out = driver(b, "nt=3", "restart=1")
    assert "Generating initial data" not in out
    driver(c, "nt=4")
    xb, xc = dump(b, "xi.dat"), dump(c, "xi.dat")
    assert np.array_equal(xb, xc)
    assert np.array_equal(dump(b, "xi_bc.dat"), dump(a, "xi_bc.dat"))
    w = C.c0()
    xo = oracle_outer_steps(w.xi, 2)
    got = xc[:w.ns, ..., 0] + 1j * xc[:w.ns, ..., 1]
    assert np.abs(got - xo).max() <= 1e-12
    # the first run's dump against the oracle too
    xa = dump(a, "xi.dat")
    assert np.abs(xa[:w.ns, ..., 0] + 1j * xa[:w.ns, ..., 1] - oracle_outer_steps(w.xi, 1)).max() <= 1e-12


def test_xi_from_file(tmp_path):
    """xi_in= reads xi alone at it == 0 (input_xi_from_file): a state written
    by the oracle continues on the GPU as it does in the oracle."""
    w = C.c0()
    x1 = oracle_outer_steps(w.xi, 1)
    full = np.zeros((w.ns + 9,) + w.shape + (2,))
    full[:w.ns, ..., 0], full[:w.ns, ..., 1] = x1.real, x1.imag
    path = tmp_path / "FinalXi.dat"
    full.tofile(path)
    driver(tmp_path, "nt=3", f"xi_in={path}")
    x = dump(tmp_path, "xi.dat")
    assert np.abs(x[:w.ns, ..., 0] + 1j * x[:w.ns, ..., 1] - oracle_outer_steps(x1, 1)).max() <= 1e-12

