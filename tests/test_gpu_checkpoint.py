"""GPU tests of checkpoint / resume (gade_checkpoint_save / gade_checkpoint_load and the C driver's -k with -s):
a run stopped after k steps and resumed in a new model (a new process for the driver) from the checkpoint
equals the uninterrupted run bit for bit: psi, the iterations and the tolerances of every later step.  The
checkpoint carries the accumulated time, psi and the solver state that passes from one step to the next
(the BiCGSTAB workspace, as tools/gpu_bicgstab.c:36-47 keeps it, and the extrapolation history).
The reference's -s only shifts the start time and restarts from psi0.vec (tools/dirac_solver.c:203-209)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import dt_of, ROOT
from dirac_hydrogen_code_amd import gade

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "dirac_hydrogen_code_amd", "bin", "dirac_solver")


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def bits(a):
    return np.ascontiguousarray(a).view(np.float64)


def model(d, dt, precision, guess):
    gp = gade.Propagator(d, 4, dt)
    gp.set_param("precision", precision)
    gp.set_param("guess_order", guess)
    return gp


@pytest.mark.parametrize("precision,guess", [(0, 0), (0, 2), (1, 0)])
def test_resume_equals_uninterrupted_run(datadir, tmp_path, precision, guess):
    cfg = "nondipole_small"
    d = datadir(cfg)
    dt = dt_of(cfg)
    K, M = 7, 9
    say(f"checkpoint precision={precision} guess={guess}")
    psi0, _ = gade.read_psi0(d, gade.Propagator(d, 4, dt).n)
    full = model(d, dt, precision, guess)
    x = gade.DeviceArray.from_host(psi0)
    it_f, tol_f, _ = full.cn_run(x, 10.0, 50.0, 2.0, K + M, tol=1e-12, timing=False)
    x_full = x.get()
    first = model(d, dt, precision, guess)
    y = gade.DeviceArray.from_host(psi0)
    first.cn_run(y, 10.0, 50.0, 2.0, K, tol=1e-12, timing=False)
    t = 0.0
    for _ in range(K):
        t = t + dt
    ck = str(tmp_path / "ck.bin")
    first.checkpoint_save(y, K, t, ck)
    first.close()
    second = model(d, dt, precision, guess)
    z = gade.DeviceArray(second.n)
    step, t2 = second.checkpoint_load(z, ck)
    assert step == K and t2 == t
    it_r, tol_r, _ = second.cn_run(z, 10.0, 50.0, 2.0, M, t0=t2, tol=1e-12, timing=False)
    assert np.array_equal(bits(z.get()), bits(x_full))
    assert np.array_equal(it_r, it_f[K:]) and np.array_equal(tol_r, tol_f[K:])
    # a checkpoint of another configuration is refused
    other = model(d, dt, 1 - precision, guess)
    with pytest.raises(gade.GadeError):
        other.checkpoint_load(gade.DeviceArray(other.n), ck)


def test_driver_resume_equals_uninterrupted(datadir, tmp_path):
    """bin/dirac_solver -f -k: 16 steps in one run, against 8 steps, then a new process with -s 9 resuming
    from the checkpoint (dt = 2^-10 exactly in both, so both runs step the same times)."""
    cfg = "dipole"
    src = datadir(cfg)
    runs = {}
    for name in ("full", "split"):
        d = str(tmp_path / name)
        shutil.copytree(src, d)
        runs[name] = d
    base = ["-l", "1", "-i", "10", "-o", "50", "-c", "1", "-e", "1e-12", "-w", "4", "-f"]
    say("driver checkpoint / resume")
    r = subprocess.run([DRIVER, "-d", runs["full"], "-t", repr(2.0 ** -6), "-n", "16", "-k",
                        os.path.join(runs["full"], "ck")] + base, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ck = os.path.join(runs["split"], "ck")
    r = subprocess.run([DRIVER, "-d", runs["split"], "-t", repr(2.0 ** -7), "-n", "8", "-k", ck] + base,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run([DRIVER, "-d", runs["split"], "-t", repr(2.0 ** -6), "-n", "16", "-s", "9", "-k", ck] + base,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "resumed from" in r.stdout
    for step in (12, 16):
        a = open(os.path.join(runs["full"], f"x{step}.out"), "rb").read()
        b = open(os.path.join(runs["split"], f"x{step}.out"), "rb").read()
        assert a == b, step
    sa = [l.split(",")[:2] for l in open(os.path.join(runs["full"], "runstats.txt")).read().splitlines()]
    sb = [l.split(",")[:2] for l in open(os.path.join(runs["split"], "runstats.txt")).read().splitlines()]
    assert len(sa) == len(sb) == 16 and sa[8:] == sb[8:]
    # a resume whose checkpoint holds another step is refused
    r = subprocess.run([DRIVER, "-d", runs["split"], "-t", repr(2.0 ** -6), "-n", "16", "-s", "5", "-k", ck] + base,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "holds step 16" in r.stderr
