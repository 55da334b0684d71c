"""GPU parity at the full production size: NSplines 250, NKappa 16, NMu 8, Nl 8 non-dipole couplings
(400 blocks x 510 radial functions, N = 204,000; 15-cycle sin^2 pulse, E0 10, omega 50, dt = T/8000).

The oracle runs here only on what it finishes in seconds at this size (one operator application, one
preconditioner solve, one CN step).  The rest is checked through properties that hold at any size.
Bars:
  * fused operator vs the oracle at a mid-pulse time: max-norm relative 1e-13 (S + H, S - H, S)
  * linearity of the fused operator: A(a x + b y) against a A x + b A y, 1e-13
  * preconditioner solve vs the oracle: 1e-11; P (P^-1 b) against b: 1e-12
  * one CN step from psi_0 at mid-pulse vs the oracle's step, both at rel-tol 1e-12: 1e-10
  * 100 device-driven steps of the timed window: every step at rel-tol 1e-12, S-norm kept to 1e-10
  * time reversal: five steps of the pulse, then five with -dt seeing the same H(t), return psi to 1e-10
  * a batched scan of two pulses: each instance equals its own single-instance run bit for bit
  * 20 mid-pulse steps: the step graph equals the launch-by-launch run bit for bit; the extrapolated guess
    lands within 1e-10 of it and mixed precision within 1e-8 (every step at rel-tol 1e-12)
  * the fused angular + radial form (ang_fused): its applies match the oracle to 1e-13 and the two-kernel
    form bit for bit, and its 20-step graph run equals the two-kernel run bit for bit
"""
import sys

import numpy as np
import pytest

import oracle as O
from dirac_hydrogen_code_amd import gade
from dirac_hydrogen_code_amd.couplings import CONFIGS, generate

pytestmark = pytest.mark.gpu

NAME = "nondipole_ns250_k16_mu8_l8"
LMAX, NTIME, E0, OMEGA, CYCLES = 8, 8000, 10.0, 50.0, 15.0


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def relmax(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def full(tmp_path_factory):
    d = str(tmp_path_factory.mktemp(NAME))
    say(f"generating {NAME}")
    generate(d, **CONFIGS[NAME])
    dt = CYCLES * 2 * np.pi / OMEGA / NTIME
    say("loading the oracle and the device model")
    om = O.Model(d, LMAX, dt, nthreads=8)
    gp = gade.Propagator(d, LMAX, dt)
    assert gp.n == 204000 and gp.blkdim == 510
    psi, _ = gade.read_psi0(d, gp.n)
    return d, dt, om, gp, psi


def test_fullsize_apply_and_linearity(full):
    d, dt, om, gp, psi = full
    say("full-size fused apply")
    rng = np.random.default_rng(11)
    x = rng.standard_normal(gp.n) + 1j * rng.standard_normal(gp.n)
    y = rng.standard_normal(gp.n) + 1j * rng.standard_normal(gp.n)
    f = gade.pulse(E0, OMEGA, CYCLES, NTIME // 2 * dt)
    om.timedep(f)
    gp.set_time(f)
    S, H = om.apply(x, "S"), om.apply(x, "H")
    dx, dy = gade.DeviceArray.from_host(x), gade.DeviceArray(gp.n)
    for cH, want in ((1.0, S + H), (-1.0, S - H), (0.0, S)):
        gp.apply(dx, dy, cH=cH)
        assert relmax(dy.get(), want) <= 1e-13, cH
    # linearity: A(a x + b y) = a A x + b A y
    a, b = 0.75 - 0.5j, -1.25 + 2.0j
    gp.apply(dx, dy, cH=1.0)
    ax = dy.get()
    dyy = gade.DeviceArray.from_host(y)
    gp.apply(dyy, dy, cH=1.0)
    ay = dy.get()
    dz = gade.DeviceArray.from_host(a * x + b * y)
    gp.apply(dz, dy, cH=1.0)
    assert relmax(dy.get(), a * ax + b * ay) <= 1e-13
    # the fused angular + radial form: the same operator, bit for bit
    two = {}
    for fused in (0, 1):
        gp.set_param("ang_fused", fused)
        for cH in (1.0, -1.0, 0.0):
            gp.apply(dx, dy, cH=cH)
            two[(fused, cH)] = dy.get()
    gp.set_param("ang_fused", 0)
    for cH, want in ((1.0, S + H), (-1.0, S - H), (0.0, S)):
        assert relmax(two[(1, cH)], want) <= 1e-13, cH
        assert np.array_equal(two[(1, cH)].view(np.float64), two[(0, cH)].view(np.float64)), cH


def test_fullsize_psolve(full):
    d, dt, om, gp, psi = full
    say("full-size preconditioner solve")
    rng = np.random.default_rng(12)
    b = rng.standard_normal(gp.n) + 1j * rng.standard_normal(gp.n)
    db, dx = gade.DeviceArray.from_host(b), gade.DeviceArray(gp.n)
    gp.psolve(db, dx)
    want = om.psolve(b)
    assert relmax(dx.get(), want) <= 1e-11
    # 32-row levels (one wave per block, two lanes per row): the same solve to rounding
    gp.set_param("psolve_chunk", 32)
    dx32 = gade.DeviceArray(gp.n)
    gp.psolve(db, dx32)
    gp.set_param("psolve_chunk", 16)
    assert relmax(dx32.get(), want) <= 1e-11
    gp.set_time(np.zeros(6, complex))  # with f = 0 the fused operator is exactly P
    dpx = gade.DeviceArray(gp.n)
    gp.apply(dx, dpx, cH=1.0)
    assert relmax(dpx.get(), b) <= 1e-12


def test_fullsize_cn_step_and_window(full):
    d, dt, om, gp, psi = full
    t0 = NTIME // 2 * dt
    say("full-size CN step at mid-pulse")
    dpsi = gade.DeviceArray.from_host(psi)
    gp.cn_step(dpsi, gade.pulse(E0, OMEGA, CYCLES, t0 + dt), tol=1e-12)
    x_ref, _, _, _ = om.cn_run(psi, E0, OMEGA, CYCLES, 1, t0=t0, tol=1e-12)
    err = np.linalg.norm(dpsi.get() - x_ref) / np.linalg.norm(x_ref)
    say(f"  one-step error {err:.3e}")
    assert err <= 1e-10
    say("full-size window of 100 device-driven steps")
    dpsi.upload(psi)
    s0 = gp.snorm(dpsi)
    it, tols, _ = gp.cn_run(dpsi, E0, OMEGA, CYCLES, 100, t0=t0, tol=1e-12)
    assert len(tols) == 100 and np.all(tols < 1e-12)
    assert abs(gp.snorm(dpsi) - s0) <= 1e-10 * s0
    say(f"  iterations/step {np.mean(it):.2f}, S-norm change {abs(gp.snorm(dpsi) - s0):.2e}")


def test_fullsize_time_reversal(full):
    """A model with -dt started at t_k + dt undoes step k (the step from t uses H(t + dt) on both sides)."""
    d, dt, om, gp, _ = full
    say("full-size time reversal, 5 steps each way")
    back_model = gade.Propagator(d, LMAX, -dt)
    t0, k = NTIME // 2 * dt, 5
    rng = np.random.default_rng(13)  # every spectral component, not just the ground state's phase
    psi = rng.standard_normal(gp.n) + 1j * rng.standard_normal(gp.n)
    psi /= np.linalg.norm(psi)
    dpsi = gade.DeviceArray.from_host(psi)
    gp.cn_run(dpsi, E0, OMEGA, CYCLES, k, t0=t0, tol=1e-13)
    x = dpsi.get()
    back_model.cn_run(dpsi, E0, OMEGA, CYCLES, k, t0=t0 + (k + 1) * dt, tol=1e-13)
    err = np.linalg.norm(dpsi.get() - psi) / np.linalg.norm(psi)
    say(f"  moved {np.linalg.norm(x - psi) / np.linalg.norm(psi):.3e}, returned to {err:.3e}")
    assert np.linalg.norm(x - psi) > 1e-3 * np.linalg.norm(psi)
    assert err <= 1e-10


def test_fullsize_batch_equals_single(full):
    """Two pulses (the reference's carrier phase and another) in one batch at the production shape: each
    instance's trajectory, iterations and tolerances equal its single-instance run bit for bit."""
    d, dt, om, gp, psi = full
    say("full-size batch of two pulses, 10 steps")
    t0, k = NTIME // 2 * dt, 10
    params = [(E0, OMEGA, CYCLES, 0.0), (0.5 * E0, 0.9 * OMEGA, CYCLES, 0.7)]
    gb = gade.Propagator(d, LMAX, dt, batch=2)
    xb = gade.DeviceArray.from_host(np.concatenate([psi, psi]))
    it_b, tol_b = gb.cn_run_batch(xb, params, k, t0=t0, tol=1e-12)
    xs = xb.get().reshape(2, -1)
    del gb, xb
    for b in range(2):
        g1 = gade.Propagator(d, LMAX, dt, batch=1)
        x1 = gade.DeviceArray.from_host(psi)
        it1, tol1 = g1.cn_run_batch(x1, [params[b]], k, t0=t0, tol=1e-12)
        assert np.array_equal(xs[b].view(np.float64), x1.get().view(np.float64)), b
        assert np.array_equal(it_b[b], it1[0]) and np.array_equal(tol_b[b], tol1[0]), b
        del g1, x1
    # the reference's pulse through the single-instance model (its BiCGSTAB workspace, which persists across
    # calls as the reference's does, cleared first like a fresh model's): the same trajectory
    gp.reset_workspace()
    x = gade.DeviceArray.from_host(psi)
    it, tols, _ = gp.cn_run(x, E0, OMEGA, CYCLES, k, t0=t0, tol=1e-12, timing=False)
    assert np.array_equal(xs[0].view(np.float64), x.get().view(np.float64))
    assert not np.array_equal(xs[0], xs[1])


def test_fullsize_graph_guess_and_mixed(full):
    """The stepping modes at the production shape, 20 steps from mid-pulse on a fresh model."""
    d, dt, om, gp, psi = full
    say("full-size graph / extrapolation / mixed, 20 steps each")
    gm = gade.Propagator(d, LMAX, dt)
    t0, k = NTIME // 2 * dt, 20

    def run(**param):
        for name, v in param.items():
            gm.set_param(name, v)
        gm.reset_workspace()
        x = gade.DeviceArray.from_host(psi)
        it, tols, _ = gm.cn_run(x, E0, OMEGA, CYCLES, k, t0=t0, tol=1e-12, timing=False)
        assert np.all(tols < 1e-12)
        return x.get(), it

    x_ref, it_ref = run(graph=0, window=4, guess_order=0, precision=0)
    x_g, it_g = run(graph=1)
    assert np.array_equal(x_g.view(np.float64), x_ref.view(np.float64)) and np.array_equal(it_g, it_ref)
    x_f, it_f = run(ang_fused=1)
    assert np.array_equal(x_f.view(np.float64), x_ref.view(np.float64)) and np.array_equal(it_f, it_ref)
    gm.set_param("ang_fused", 0)
    nref = np.linalg.norm(x_ref)
    x_e, it_e = run(graph=0, guess_order=2)
    err_e = np.linalg.norm(x_e - x_ref) / nref
    x_m, it_m = run(guess_order=0, precision=1)
    err_m = np.linalg.norm(x_m - x_ref) / nref
    say(f"  extrapolation {err_e:.2e} ({np.mean(it_e):.2f} it/step), mixed {err_m:.2e}, plain {np.mean(it_ref):.2f}")
    assert err_e <= 1e-10
    assert err_m <= 1e-8
