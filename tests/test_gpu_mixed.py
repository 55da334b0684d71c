"""GPU tests of the mixed-precision CN step: the inner BiCGSTAB and the preconditioner apply run in complex
FP32, wrapped in FP64 defect correction (the residual r = b - A x, the update and the tolerance test in
FP64), so every step is solved to the FP64 tolerance.

Bars: every step meets the FP64 relative tolerance (1e-12); the trajectory agrees with the plain FP64 form
and with the CPU oracle to 1e-8 relative (final, and per step as a one-step map to 1e-9); the S-norm drift
matches the FP64 form's to 1e-9; batched instances equal their single-instance runs bit for bit, and graph
replays equal one-by-one launches bit for bit."""
import numpy as np
import pytest

import oracle as O
from conftest import dt_of
from dirac_hydrogen_code_amd import gade

pytestmark = pytest.mark.gpu

LMAX = {"tiny": 2, "dipole": 1, "nondipole_small": 4}
TOL = 1e-12


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def bits(a):
    return np.ascontiguousarray(a).view(np.float64)


def model(d, cfg, dt, precision, batch=1, **param):
    gp = gade.Propagator(d, LMAX[cfg], dt, batch=batch)
    gp.set_param("precision", precision)
    for k, v in param.items():
        gp.set_param(k, v)
    return gp


@pytest.mark.parametrize("cfg", ["dipole", "nondipole_small"])
def test_mixed_matches_fp64_and_oracle(datadir, cfg):
    d = datadir(cfg)
    dt = dt_of(cfg)
    nsteps = 40
    say(f"mixed vs fp64 {cfg}")
    out = {}
    for prec in (0, 1):
        gp = model(d, cfg, dt, prec)
        psi0, _ = gade.read_psi0(d, gp.n)
        x = gade.DeviceArray.from_host(psi0)
        it, tols, _ = gp.cn_run(x, 10.0, 50.0, 1.0, nsteps, tol=TOL, timing=False)
        out[prec] = (x.get(), it, tols, gp.snorm(x))
        assert gp.get_param("precision") == prec
    x64, xmp = out[0][0], out[1][0]
    assert np.all(out[1][2] <= TOL), out[1][2].max()  # the FP64 tolerance on every step
    assert out[1][1].max() >= 1                          # the inner solve did work
    err = rel(xmp, x64)
    say(f"  final |psi_mixed - psi_fp64| / |psi_fp64| = {err:.2e}; S-norm fp64 {out[0][3]:.15f} "
        f"mixed {out[1][3]:.15f}; inner iters/step {out[1][1].mean():.2f} vs fp64 {out[0][1].mean():.2f}")
    assert err <= 1e-8
    assert abs(out[1][3] - out[0][3]) <= 1e-9
    # against the oracle (the reference CPU path restated)
    om = O.Model(d, LMAX[cfg], dt, nthreads=4)
    psi0, _ = gade.read_psi0(d, len(x64))
    xo, _, _, _ = om.cn_run(psi0, 10.0, 50.0, 1.0, nsteps, tol=TOL)
    assert rel(xmp, xo) <= 1e-8


def test_mixed_one_step_map(datadir):
    """From the same psi_n (mid-pulse), one mixed step and one FP64 step agree to 1e-9."""
    cfg = "nondipole_small"
    d = datadir(cfg)
    dt = dt_of(cfg)
    say("mixed one-step map")
    gp = model(d, cfg, dt, 0)
    psi0, _ = gade.read_psi0(d, gp.n)
    x = gade.DeviceArray.from_host(psi0)
    gp.cn_run(x, 10.0, 50.0, 1.0, 30, tol=TOL, timing=False)
    psin = x.get()
    worst = 0.0
    for k in range(5):
        t0 = (30 + k) * dt
        res = {}
        for prec in (0, 1):
            g = model(d, cfg, dt, prec)
            y = gade.DeviceArray.from_host(psin)
            it, tols, _ = g.cn_run(y, 10.0, 50.0, 1.0, 1, t0=t0, tol=TOL, timing=False)
            res[prec] = y.get()
            assert tols[0] <= TOL
        worst = max(worst, rel(res[1], res[0]))
        psin = res[0]
    say(f"  worst one-step difference {worst:.2e}")
    assert worst <= 1e-9


def test_mixed_batch_equals_single_and_graph(datadir):
    """Batched mixed instances (different pulses and step counts, one zero-rhs instance) equal their own
    single-instance mixed runs bit for bit; the step graph equals one-by-one launches bit for bit."""
    cfg = "dipole"
    d = datadir(cfg)
    dt = dt_of(cfg)
    say("mixed batch / graph")
    params = [(10.0, 50.0, 1.0, 0.0), (30.0, 40.0, 2.0, 0.7), (3.0, 60.0, 1.0, 2.0)]
    nsteps_b = [20, 12, 20]
    g1 = model(d, cfg, dt, 1)
    psi0, _ = gade.read_psi0(d, g1.n)
    psis = [psi0, psi0, np.zeros_like(psi0)]
    runs = {}
    for graph in (1, 0):
        gp = model(d, cfg, dt, 1, batch=3, graph=graph, window=6)
        x = gade.DeviceArray.from_host(np.concatenate(psis))
        it, tols = gp.cn_run_batch(x, params, 20, nsteps_b=nsteps_b, tol=TOL)
        runs[graph] = (x.get().reshape(3, -1), it, tols)
    assert np.array_equal(bits(runs[1][0]), bits(runs[0][0]))
    assert np.array_equal(runs[1][1], runs[0][1]) and np.array_equal(runs[1][2], runs[0][2])
    xs, it, tols = runs[1]
    assert not np.any(xs[2]) and np.all(it[2] == 0) and np.all(tols[2] == 0)  # zero rhs: x = 0
    for b in range(2):
        g = model(d, cfg, dt, 1, window=6)
        y = gade.DeviceArray.from_host(psis[b])
        it1, tol1 = g.cn_run_batch(y, [params[b]], nsteps_b[b], tol=TOL)
        assert np.array_equal(bits(xs[b]), bits(y.get())), b
        assert np.array_equal(it[b, :nsteps_b[b]], it1[0]) and np.array_equal(tols[b, :nsteps_b[b]], tol1[0])
        assert np.all(tol1 <= TOL)


def test_mixed_handback_and_sweep_count(datadir):
    """One sweep enqueued per step (iter_batch 1) forces steps back to the host; 1..4 sweeps per step give
    the same trajectory bit for bit (extra sweeps are gated no-ops)."""
    cfg = "nondipole_small"
    d = datadir(cfg)
    dt = dt_of(cfg)
    say("mixed hand-back")
    out = []
    for J in (1, 2, 4):
        gp = model(d, cfg, dt, 1, iter_batch=J, window=5)
        psi0, _ = gade.read_psi0(d, gp.n)
        x = gade.DeviceArray.from_host(psi0)
        it, tols, _ = gp.cn_run(x, 10.0, 50.0, 1.0, 12, tol=TOL, timing=False)
        out.append((x.get(), it, tols))
        assert np.all(tols <= TOL)
    for o in out[1:]:
        assert np.array_equal(bits(o[0]), bits(out[0][0]))
        assert np.array_equal(o[1], out[0][1]) and np.array_equal(o[2], out[0][2])
