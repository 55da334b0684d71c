"""GPU test of the learned initial guess on the 64-instance batched scan, in FP64 and in the mixed-precision
form (complex-FP32 inner BiCGSTAB and preconditioner inside FP64 defect correction).

The guess is trained on trajectories of seeded pulses (0x5EED: POD of 400 extrapolation-defect snapshots,
20,000 samples) and judged on the held-out scan 0x6ADE, which shares no pulse with the training set.

Bars: every step of every form meets the FP64 tolerance 1e-12; every mixed-precision instance ends within
1e-8 of the plain FP64 instance (relative); the trained network lowers the FP64 iterations per step to at most
0.8 of the extrapolation's (the gain it is there for) and does not raise the mixed form's inner iterations;
the network takes most steps that have a full history."""
import numpy as np
import pytest

from conftest import dt_of
from dirac_hydrogen_code_amd import gade
from dirac_hydrogen_code_amd import learned

pytestmark = pytest.mark.gpu

CFG = "dipole"
B = 64
TOL = 1e-12


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def test_learned_guess_on_batched_mixed_scan(datadir):
    d = datadir(CFG)
    dt = dt_of(CFG)
    say("learned scan B=64: snapshots, POD, samples, training")
    gp = gade.Propagator(d, 1, dt, batch=B)
    psi0, _ = gade.read_psi0(d, gp.n)
    tr = learned.Trainer(gp, psi0, seed=0x5EED, nsnap=400, nsamples=20000, log=say)
    tr.build_pod()
    tr.collect_samples()
    assert len(tr.valid) >= 20000
    _, v0, v1 = tr.fit(epochs=30, M=256, lr=1e-3, lr_final=1e-4)
    assert abs(v0 - 1.0) <= 1e-5 and v1 < 0.1 * v0
    params, nsteps_b = learned.scan_pulses(0x6ADE, B, dt)
    K = int(min(nsteps_b.min(), 120))
    res = {}
    for prec in (0, 1):
        gp.set_param("precision", prec)
        for name, go in (("plain", 0), ("extrapolation", 2), ("learned", gade.GUESS_LEARNED)):
            gp.set_param("lg_features", 0)
            gp.set_param("guess_order", go)
            x = gade.DeviceArray.from_host(np.tile(psi0, B))
            gp.reset_workspace()
            it, tols = gp.cn_run_batch(x, params, K, tol=TOL)
            key = ("fp64_" if prec == 0 else "mixed_") + name
            assert np.all(tols <= TOL), (key, tols.max())
            res[key] = (x.get().reshape(B, -1), float(it.mean()))
            say(f"  {key}: iterations/step {it.mean():.3f}")
            if go == gade.GUESS_LEARNED:
                uses = gp.lg_uses()
                say(f"  {key}: network on {uses[:, 0].sum()} steps, extrapolation fallback on {uses[:, 1].sum()}")
                assert uses[:, 0].sum() >= 0.5 * uses.sum()
    gp.set_param("precision", 0)
    gp.set_param("guess_order", 0)
    ref = res["fp64_plain"][0]
    for key in ("fp64_learned", "mixed_plain", "mixed_extrapolation", "mixed_learned"):
        worst = max(float(np.linalg.norm(res[key][0][b] - ref[b]) / np.linalg.norm(ref[b])) for b in range(B))
        say(f"  {key}: worst instance |psi - psi_fp64_plain| / |psi_fp64_plain| = {worst:.2e}")
        assert worst <= 1e-8, (key, worst)
    gain = res["fp64_learned"][1] / res["fp64_extrapolation"][1]
    say(f"  FP64 iterations/step, learned / extrapolation = {gain:.3f}")
    assert gain <= 0.8
    assert res["mixed_learned"][1] <= res["mixed_extrapolation"][1] * 1.02
