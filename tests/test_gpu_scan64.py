"""GPU tests of the 64-instance pulse scan on the dipole couplings (NSplines 50, NKappa 4, Nl 1): pulses drawn
from the seeded RNG 0x6ADE (E0 log-uniform in [1, 100], omega uniform in [25, 75], carrier phase uniform in
[0, 2 pi), 1..15 cycles), one dt for all, per-instance step counts so that instances retire at different
steps, in FP64 and in mixed precision.

B = 64 is above the 32 instances that one G-sum launch carries by value, so the host-table G-sums run as
two launches (instances 0..31 and 32..63); the device-driven stepping reads its coefficients from the
window table in one launch.  Both paths are checked.

Bars: every instance equals its own single-instance run bit for bit (psi, iterations, tolerances); the
instance that runs the reference's own pulse (E0 10, omega 50, one cycle, carrier phase 0) for 200 steps
matches the CPU oracle to 1e-9 (FP64) and 1e-8 (mixed); mixed and FP64 scans agree to 1e-8 per instance;
every step meets the FP64 tolerance 1e-12."""
import numpy as np
import pytest

import oracle as O
from conftest import dt_of
from dirac_hydrogen_code_amd import gade
from dirac_hydrogen_code_amd import learned

pytestmark = pytest.mark.gpu

CFG = "dipole"
LMAX = 1
TOL = 1e-12
B = 64
NREF = 200  # steps of the reference-pulse instance


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def bits(a):
    return np.ascontiguousarray(a).view(np.float64)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def scan_setup(dt):
    params, own = learned.scan_pulses(0x6ADE, B, dt)
    params = params.copy()
    params[0] = (10.0, 50.0, 1.0, 0.0)  # the reference's pulse
    nsteps_b = np.array([NREF] + [12 + (7 * b) % 41 for b in range(1, B)], np.int32)
    assert np.all(nsteps_b[1:] <= own[1:])  # every instance stays inside its own pulse
    return params, nsteps_b


def run_batch(d, dt, psi0, params, nsteps_b, precision):
    gp = gade.Propagator(d, LMAX, dt, batch=B)
    gp.set_param("precision", precision)
    x = gade.DeviceArray.from_host(np.tile(psi0, B))
    it, tols = gp.cn_run_batch(x, params, int(nsteps_b.max()), nsteps_b=nsteps_b, tol=TOL)
    return x.get().reshape(B, -1), it, tols


@pytest.mark.parametrize("precision", [0, 1])
def test_scan64_equals_single_instances_and_oracle(datadir, precision):
    d = datadir(CFG)
    dt = dt_of(CFG)
    say(f"scan64 precision={precision}")
    g1 = gade.Propagator(d, LMAX, dt)
    psi0, _ = gade.read_psi0(d, g1.n)
    del g1
    params, nsteps_b = scan_setup(dt)
    xs, it, tols = run_batch(d, dt, psi0, params, nsteps_b, precision)
    for b in range(B):
        nb = int(nsteps_b[b])
        assert np.all(tols[b, :nb] <= TOL), (b, tols[b, :nb].max())
        g = gade.Propagator(d, LMAX, dt)
        g.set_param("precision", precision)
        y = gade.DeviceArray.from_host(psi0)
        it1, tol1 = g.cn_run_batch(y, [params[b]], nb, tol=TOL)
        assert np.array_equal(bits(xs[b]), bits(y.get())), b
        assert np.array_equal(it[b, :nb], it1[0]), b
        assert np.array_equal(tols[b, :nb], tol1[0]), b
    say(f"  64 instances equal their single runs; iterations/step {it[it > 0].mean() if np.any(it) else 0:.2f}")
    om = O.Model(d, LMAX, dt, nthreads=4)
    po, _ = O.read_psi0_vec(d + "/psi0.vec", om.n)
    xo, _, _, _ = om.cn_run(po, 10.0, 50.0, 1.0, NREF, tol=TOL)
    err = rel(xs[0], xo)
    say(f"  reference-pulse instance vs oracle over {NREF} steps: {err:.2e}")
    assert err <= (1e-9 if precision == 0 else 1e-8)


def test_scan64_mixed_agrees_with_fp64(datadir):
    d = datadir(CFG)
    dt = dt_of(CFG)
    say("scan64 mixed vs fp64")
    g1 = gade.Propagator(d, LMAX, dt)
    psi0, _ = gade.read_psi0(d, g1.n)
    del g1
    params, nsteps_b = scan_setup(dt)
    x64, _, _ = run_batch(d, dt, psi0, params, nsteps_b, 0)
    xmp, itmp, tolmp = run_batch(d, dt, psi0, params, nsteps_b, 1)
    worst = max(rel(xmp[b], x64[b]) for b in range(B))
    say(f"  worst instance |psi_mixed - psi_fp64| / |psi_fp64| = {worst:.2e}")
    assert worst <= 1e-8
    for b in range(B):
        assert np.all(tolmp[b, :nsteps_b[b]] <= TOL)


def test_gsums_two_launch_groups_equal_single_applies(datadir):
    """Host-table G-sums for 40 instances (launches of 32 and 8) and the fused apply on the batch: every
    instance's result equals a single-instance model's apply with the same f(t), bit for bit, for both signs
    of the time-dependent part."""
    d = datadir(CFG)
    dt = dt_of(CFG)
    say("gsums over two launch groups (B=40)")
    nb = 40
    gp = gade.Propagator(d, LMAX, dt, batch=nb)
    rng = np.random.default_rng(40)
    n = gp.n
    x = rng.standard_normal(nb * n) + 1j * rng.standard_normal(nb * n)
    params, _ = learned.scan_pulses(0x6ADE, nb, dt)
    ts = (1 + np.arange(nb)) * 3 * dt
    ft = np.array([gade.pulse_cep(*params[b], ts[b]) for b in range(nb)])
    assert np.all(np.abs(ft[32:]).max(axis=1) > 0)  # the second launch group carries non-trivial pulses
    gp.set_time_batch(ft)
    dx = gade.DeviceArray.from_host(x)
    single = gade.Propagator(d, LMAX, dt)
    for cH in (1.0, -1.0):
        dy = gade.DeviceArray(nb * n)
        gp.apply(dx, dy, cH=cH)
        y = dy.get().reshape(nb, n)
        for b in range(nb):
            single.set_time(ft[b])
            sx = gade.DeviceArray.from_host(x[b * n:(b + 1) * n])
            sy = gade.DeviceArray(n)
            single.apply(sx, sy, cH=cH)
            assert np.array_equal(bits(y[b]), bits(sy.get())), (cH, b)
