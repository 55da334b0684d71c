"""GPU tests of the batched pulse scan: B instances sharing the couplings advance in one set of launches per
step, each with its own pulse, BiCGSTAB scalars, convergence state and step count.

Bars: an instance of a batch equals its single-instance run bit for bit (same kernels, same per-instance
reduction order); against the CPU oracle, trajectories agree to 1e-9 relative."""
import numpy as np
import pytest

import oracle as O
from conftest import dt_of
from dirac_hydrogen_code_amd import gade

pytestmark = pytest.mark.gpu

LMAX = {"tiny": 2, "dipole": 1, "nondipole_small": 4}


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def run(d, cfg, dt, psis, params, nsteps, nsteps_b=None, **param):
    B = len(params)
    gp = gade.Propagator(d, LMAX[cfg], dt, batch=B)
    for k, v in param.items():
        gp.set_param(k, v)
    x = gade.DeviceArray.from_host(np.concatenate(psis))
    it, tols = gp.cn_run_batch(x, params, nsteps, nsteps_b=nsteps_b, tol=1e-12)
    return x.get().reshape(B, -1), it, tols


@pytest.mark.parametrize("cfg", ["tiny", "nondipole_small"])
def test_batch_of_identical_pulses_equals_single(datadir, cfg):
    d = datadir(cfg)
    dt = dt_of(cfg)
    say(f"batch identical {cfg}")
    gp1 = gade.Propagator(d, LMAX[cfg], dt)
    psi0, _ = gade.read_psi0(d, gp1.n)
    x1 = gade.DeviceArray.from_host(psi0)
    it1, tol1, _ = gp1.cn_run(x1, 10.0, 50.0, 2.0, 12, tol=1e-12, timing=False)
    xs, it, tols = run(d, cfg, dt, [psi0] * 3, [(10.0, 50.0, 2.0, 0.0)] * 3, 12)
    for b in range(3):
        assert np.array_equal(xs[b].view(np.float64), x1.get().view(np.float64))
        assert np.array_equal(it[b], it1) and np.array_equal(tols[b], tol1)


def test_batch_mixed_pulses_masks_and_zero_rhs(datadir):
    """Different pulses (amplitude, frequency, cycles, carrier phase), different step counts, one instance
    with psi = 0 (zero rhs every step): every instance equals its own single-instance run; the cep = 0
    instance matches the oracle."""
    cfg = "dipole"
    d = datadir(cfg)
    dt = dt_of(cfg)
    say("batch mixed")
    gp1 = gade.Propagator(d, LMAX[cfg], dt)
    psi0, _ = gade.read_psi0(d, gp1.n)
    del gp1
    params = [(10.0, 50.0, 1.0, 0.0), (30.0, 40.0, 2.0, 0.7), (3.0, 60.0, 1.0, 2.0), (10.0, 50.0, 1.0, 0.0)]
    psis = [psi0, psi0, psi0, np.zeros_like(psi0)]
    nsteps_b = [20, 12, 20, 5]
    xs, it, tols = run(d, cfg, dt, psis, params, 20, nsteps_b=nsteps_b)
    for b in range(4):
        xb, itb, tb = run(d, cfg, dt, [psis[b]], [params[b]], nsteps_b[b])
        assert np.array_equal(xs[b].view(np.float64), xb[0].view(np.float64)), b
        assert np.array_equal(it[b, :nsteps_b[b]], itb[0]), b
        assert np.array_equal(tols[b, :nsteps_b[b]], tb[0]), b
    assert not np.any(xs[3]) and np.all(it[3, :5] == 0)
    om = O.Model(d, LMAX[cfg], dt, nthreads=4)
    po, _ = O.read_psi0_vec(d + "/psi0.vec", om.n)
    xo, ito, _, _ = om.cn_run(po, 10.0, 50.0, 1.0, 20, tol=1e-12)
    assert np.linalg.norm(xs[0] - xo) / np.linalg.norm(xo) <= 1e-9


def test_batch_hand_back_gives_same_result(datadir):
    """One enqueued iteration per step: instances needing more are finished by the host (extra gated
    iterations) with results equal to the adaptive run."""
    cfg = "nondipole_small"
    d = datadir(cfg)
    dt = dt_of(cfg)
    say("batch hand-back")
    gp1 = gade.Propagator(d, LMAX[cfg], dt)
    psi0, _ = gade.read_psi0(d, gp1.n)
    del gp1
    params = [(10.0, 50.0, 2.0, 0.0), (25.0, 45.0, 2.0, 1.0)]
    a = run(d, cfg, dt, [psi0] * 2, params, 10)
    b = run(d, cfg, dt, [psi0] * 2, params, 10, iter_batch=1, window=3)
    assert np.array_equal(a[0].view(np.float64), b[0].view(np.float64))
    assert np.array_equal(a[1], b[1]) and a[1].max() >= 1
