"""GPU tests of the captured CN step: with the banded radial kernel every device-driven step is the same
launches (its pulse coefficients and step index are read from device memory), captured once into a
hipGraph and replayed for every step of a window.

Bar: bit-exact against the same steps launched one by one (graph 0), after every call, with host
synchronizations, host-side copies and eager launches between the replays; including stalled steps handed
back to the host (iter_batch 1), extrapolated guesses and batched instances with different step counts."""
import numpy as np
import pytest

from conftest import dt_of
from dirac_hydrogen_code_amd import gade

pytestmark = pytest.mark.gpu

LMAX = {"tiny": 2, "dipole": 1, "nondipole_small": 4}


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def bits(a):
    return np.ascontiguousarray(a).view(np.float64)


@pytest.mark.parametrize("window,batch,guess", [(8, 0, 0), (5, 1, 0), (4, 0, 2)])
def test_graph_replays_equal_eager_steps(datadir, window, batch, guess):
    cfg = "nondipole_small"
    d = datadir(cfg)
    dt = dt_of(cfg)
    say(f"graph window={window} iter_batch={batch} guess={guess}")
    runs = {}
    for graph in (1, 0):
        gp = gade.Propagator(d, LMAX[cfg], dt)
        gp.set_param("radial", 2)
        gp.set_param("device_loop", 1)
        gp.set_param("graph", graph)
        gp.set_param("window", window)
        gp.set_param("iter_batch", batch)
        gp.set_param("guess_order", guess)
        psi0, _ = gade.read_psi0(d, gp.n)
        x = gade.DeviceArray.from_host(psi0)
        snaps = []
        t0 = 0.0
        for call in range(3):  # several calls: replays of one graph across host round trips
            it, tols, _ = gp.cn_run(x, 10.0, 50.0, 1.0, 7, t0=t0, tol=1e-12, timing=True)
            t0 += 7 * dt
            snaps.append((x.get(), it, tols))
            # eager work between the replays: a norm, an operator application and a copy back
            gp.snorm(x)
            y = gade.DeviceArray.from_host(np.zeros(gp.n, complex))
            gp.apply(x, y, 1.0, 0.0)
            del y
        runs[graph] = snaps
        assert gp.get_param("graph") == graph
    for k, (a, b) in enumerate(zip(runs[1], runs[0])):
        assert np.array_equal(bits(a[0]), bits(b[0])), k
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]), k
        assert a[1].max() >= 1


def test_graph_batched_masks(datadir):
    """Batched instances (different pulses and step counts, one zero-rhs instance) through the graph equal
    the one-by-one launches bit for bit."""
    cfg = "dipole"
    d = datadir(cfg)
    dt = dt_of(cfg)
    say("graph batched")
    params = [(10.0, 50.0, 1.0, 0.0), (30.0, 40.0, 2.0, 0.7), (3.0, 60.0, 1.0, 2.0), (10.0, 50.0, 1.0, 0.0)]
    nsteps_b = [20, 12, 20, 5]
    out = {}
    for graph in (1, 0):
        gp = gade.Propagator(d, LMAX[cfg], dt, batch=len(params))
        gp.set_param("graph", graph)
        gp.set_param("window", 6)
        psi0, _ = gade.read_psi0(d, gp.n)
        psis = [psi0, psi0, psi0, np.zeros_like(psi0)]
        x = gade.DeviceArray.from_host(np.concatenate(psis))
        it, tols = gp.cn_run_batch(x, params, 20, nsteps_b=nsteps_b, tol=1e-12)
        out[graph] = (x.get(), it, tols)
    assert np.array_equal(bits(out[1][0]), bits(out[0][0]))
    assert np.array_equal(out[1][1], out[0][1]) and np.array_equal(out[1][2], out[0][2])


@pytest.mark.parametrize("cfg,batch", [("nondipole_small", 1), ("dipole", 3)])
def test_autotune_keeps_trajectory(datadir, cfg, batch):
    """The autotuner times every (window, iter_batch, graph) candidate on probe steps, requires them to agree
    bit for bit (it fails otherwise), keeps the fastest, and restores the solver state: the run after tuning
    equals the same run without it."""
    d = datadir(cfg)
    dt = dt_of(cfg)
    say(f"autotune {cfg} batch={batch}")
    params = [(10.0, 50.0, 1.0, 0.3 * b) for b in range(batch)]
    out = {}
    for tune in (1, 0):
        gp = gade.Propagator(d, LMAX[cfg], dt, batch=batch)
        psi0, _ = gade.read_psi0(d, gp.n)
        x = gade.DeviceArray.from_host(np.concatenate([psi0] * batch))
        it1, tol1 = gp.cn_run_batch(x, params, 5, tol=1e-12)
        if tune:
            chosen, rows = gp.autotune(x, params, 6, t0=5 * dt, reps=1)
            # kernel shapes: 6 tile shapes x 2 output-half groupings x (3 angular groups + the fused form for the
            # 5 shapes with up to 2 outputs per lane), only the angular groups with the FP64 MFMA radial kernel
            # (one shape); then 5 windows x 4 iteration batches x graph on / off
            shapes = 3 if gp.get_param("radial_mfma") and gp.get_param("precision") == 0 else 46
            assert len(rows) == (shapes + 40 if gp.get_param("radial") == 2 else 20)
            assert sum(r["ang_fused"] for r in rows[:shapes]) == (10 if shapes == 46 else 0)
            assert all(r["ms"] > 0 for r in rows)
            c = rows[chosen]
            keys = ("window", "iter_batch", "graph", "dia_rows", "dia_outputs", "ang_lb", "dia_halves", "ang_fused")
            assert tuple(gp.get_param(k) for k in keys) == tuple(c[k] for k in keys)
            say(f"  chose {c} of {len(rows)}")
        it2, tol2 = gp.cn_run_batch(x, params, 10, t0=5 * dt, tol=1e-12)
        out[tune] = (x.get(), it1, it2, tol2)
    assert np.array_equal(bits(out[1][0]), bits(out[0][0]))
    for k in (1, 2, 3):
        assert np.array_equal(out[1][k], out[0][k])


def test_graph_recaptured_when_window_table_grows(datadir):
    """A larger window replaces the device table of pulse coefficients; the cached step graph (same K, same
    psi) must be captured again against the new table: results equal the one-by-one launches."""
    cfg = "dipole"
    d = datadir(cfg)
    dt = dt_of(cfg)
    say("graph window growth")
    params = [(10.0, 50.0, 1.0, 0.0), (30.0, 40.0, 1.0, 1.0)]
    out = {}
    for graph in (1, 0):
        gp = gade.Propagator(d, LMAX[cfg], dt, batch=2)
        gp.set_param("graph", graph)
        gp.set_param("iter_batch", 3)  # the same K in every window
        psi0, _ = gade.read_psi0(d, gp.n)
        x = gade.DeviceArray.from_host(np.concatenate([psi0, psi0]))
        res = []
        t0 = 0.0
        for w in (2, 5, 9, 16):
            gp.set_param("window", w)
            it, tols = gp.cn_run_batch(x, params, w, t0=t0, tol=1e-12)
            t0 += w * dt
            res.append((x.get(), it, tols))
        out[graph] = res
    for a, b in zip(out[1], out[0]):
        assert np.array_equal(bits(a[0]), bits(b[0]))
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
