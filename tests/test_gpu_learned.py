"""GPU tests of the learned initial guess: the device MLP (BF16 MFMA, FP32 accumulation) against the
CPU oracle (oracle/lg_oracle.py, with the same bf16 roundings), POD orthonormality, training, and the
guess inside the solver on held-out pulses (the converged solve at the unchanged tolerance judges it).

Bars: forward outputs within 3e-4 of the output scale; loss within 1e-4 relative and every layer's gradient
within 5e-3 relative (the oracle rounds to bf16 at the same points; the bars are about ten times the observed
agreement, which differs only where accumulation orders differ); POD basis
orthonormal to 1e-12; trained validation loss below the extrapolation's (1.0); every step of the held-out
scan meets the tolerance and the final states agree with the plain guess's to 1e-9."""
import numpy as np
import pytest

import lg_oracle as LG
from conftest import dt_of
from dirac_hydrogen_code_amd import gade
from dirac_hydrogen_code_amd import learned

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def random_params(rng, scale_out=0.3):
    p = np.zeros(LG.NPARAM, np.float32)
    for k, (fi, fo) in enumerate(LG.LAYERS):
        wo, bo = LG.OFFS[k]
        p[wo:wo + fo * fi] = rng.standard_normal(fo * fi) * np.sqrt(2.0 / fi) * (scale_out if k == 3 else 1.0)
        p[bo:bo + fo] = 0.1 * rng.standard_normal(fo)
    return p


def random_samples(rng, M, n=200):
    U, _ = np.linalg.qr(rng.standard_normal((n, LG.POD)) + 1j * rng.standard_normal((n, LG.POD)))
    rows = []
    for _ in range(M):
        Dh = rng.standard_normal((n, 6)) + 1j * rng.standard_normal((n, 6))
        Dh /= np.linalg.norm(Dh, axis=0)
        ehat = 0.3 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
        rows.append(LG.sample_from_vectors(rng.standard_normal(LG.IN) * 2 + 1, ehat, Dh, U))
    return np.array(rows, np.float32)


def net_model(datadir):
    d = datadir("dipole")
    return gade.Propagator(d, 1, dt_of("dipole"))


@pytest.mark.parametrize("M", [16, 48, 160])
def test_mlp_forward_matches_oracle(datadir, M):
    """Up to 64 rows (inference; 16 rows serve up to 16 instances) every tile's k steps are split over four
    waves; more rows (training) use one wave per tile: both within the bar."""
    say(f"mlp forward M={M}")
    rng = np.random.default_rng(11)
    gp = net_model(datadir)
    p = random_params(rng)
    mean = rng.standard_normal(LG.IN).astype(np.float32)
    std = (0.5 + rng.random(LG.IN)).astype(np.float32)
    gp.lg_set_norm(mean, std)
    gp.lg_set_params(p)
    assert np.array_equal(gp.lg_get_params(), p)
    x = (rng.standard_normal((M, LG.IN)) * 2).astype(np.float32)
    dx = gade.DeviceArray.from_host(x.ravel())
    dy = gade.DeviceArray(M * LG.OUT, np.float32)
    gp.lg_forward(dx, M, dy)
    y = dy.get().reshape(M, LG.OUT)
    yo, _ = LG.forward(p, LG.standardize(x, mean, std).astype(np.float32), quantized=True)
    scale = np.abs(yo).max()
    err = np.abs(y - yo).max()
    say(f"  max |y - y_oracle| = {err:.2e} (scale {scale:.2e})")
    assert err <= 3e-4 * scale


def test_mlp_gradient_matches_oracle(datadir):
    say("mlp gradient")
    rng = np.random.default_rng(12)
    gp = net_model(datadir)
    p = random_params(rng)
    gp.lg_set_norm(np.zeros(LG.IN, np.float32), np.ones(LG.IN, np.float32))
    gp.lg_set_params(p)
    M = 32
    S = random_samples(rng, M)
    ds = gade.DeviceArray.from_host(S.ravel())
    didx = gade.DeviceArray.from_host(np.arange(M, dtype=np.int32))
    dg = gade.DeviceArray(LG.NPARAM, np.float32)
    loss = gp.lg_grad(ds, didx, M, dg)
    g = dg.get().astype(np.float64)
    X = LG.standardize(S[:, :LG.IN].astype(np.float64), 0.0, 1.0).astype(np.float32)
    Y, cache = LG.forward(p, X, quantized=True)
    lo, dY = LG.loss_and_dY(Y, S)
    go = LG.backward(p, cache, dY.astype(np.float32), quantized=True)
    say(f"  loss {loss:.6f} vs oracle {lo:.6f}")
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    for k, (fi, fo) in enumerate(LG.LAYERS):
        wo, bo = LG.OFFS[k]
        for a, b, name in ((g[wo:wo + fo * fi], go[wo:wo + fo * fi], "W"), (g[bo:bo + fo], go[bo:bo + fo], "b")):
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            say(f"  layer {k + 1} {name}: relative gradient difference {rel:.2e}")
            assert rel <= 5e-3, (k, name, rel)


def test_learned_guess_pipeline_and_held_out_scan(datadir):
    """POD from solver snapshots, samples, training, then a held-out scan with the learned guess."""
    d = datadir("dipole")
    dt = dt_of("dipole")
    B = 16
    say("learned pipeline: snapshots / POD")
    gp = gade.Propagator(d, 1, dt, batch=B)
    psi0, _ = gade.read_psi0(d, gp.n)
    tr = learned.Trainer(gp, psi0, seed=0x5EED, nsnap=128, nsamples=3000, log=say)
    sv = tr.build_pod()
    assert np.all(np.diff(sv) <= 1e-12 * sv[0])  # descending
    U = gp.lg_basis()
    orth = np.abs(U.conj() @ U.T - np.eye(U.shape[0])).max()
    say(f"  POD orthonormality {orth:.2e}")
    assert orth <= 1e-12
    say("learned pipeline: samples / training")
    tr.collect_samples()
    assert len(tr.valid) >= 1000
    hist, v0, v1 = tr.fit(epochs=12, M=128, lr=1e-3, lr_final=1e-4)
    assert abs(v0 - 1.0) <= 1e-5  # the untrained network is the extrapolation
    assert np.all(np.isfinite(hist))
    assert v1 < v0
    # held-out pulses (another seed): plain, extrapolated and learned guesses
    params, nsteps_b = learned.scan_pulses(0x6ADE, B, dt)
    K = int(min(nsteps_b.min(), 120))
    res = {}
    for name, go in (("plain", 0), ("extrapolation", 2), ("learned", gade.GUESS_LEARNED)):
        gp.set_param("lg_features", 0)
        gp.set_param("guess_order", go)
        x = gade.DeviceArray.from_host(np.tile(psi0, B))
        gp.reset_workspace()
        it, tols = gp.cn_run_batch(x, params, K, tol=1e-12)
        res[name] = (x.get(), it, tols)
        assert np.all(tols <= 1e-12), name
        say(f"  held-out {name}: iterations/step {it.mean():.3f}")
        if go == gade.GUESS_LEARNED:
            uses = gp.lg_uses()
            say(f"  held-out learned: network on {uses[:, 0].sum()} steps, range gate fallback on {uses[:, 1].sum()}")
            # every step with a full history is counted once; the held-out pulses come from the training
            # distribution, so the network takes most of them
            assert np.all(uses.sum(1) <= K) and np.all(uses.sum(1) >= K - 8)
            assert uses[:, 0].sum() >= 0.5 * uses.sum()
    # a tight range gate sends more steps to the extrapolation; the converged result is unchanged
    gp.set_param("lg_zmax", 1)
    x = gade.DeviceArray.from_host(np.tile(psi0, B))
    gp.reset_workspace()
    it, tols = gp.cn_run_batch(x, params, K, tol=1e-12)
    gated = gp.lg_uses()
    say(f"  held-out learned, gate at 1 sigma: network on {gated[:, 0].sum()} steps, fallback on {gated[:, 1].sum()}")
    assert gated[:, 1].sum() > uses[:, 1].sum() and np.all(tols <= 1e-12)
    res["learned_gated"] = (x.get(), it, tols)
    gp.set_param("lg_zmax", 8)
    ref = res["plain"][0]
    for name in ("extrapolation", "learned", "learned_gated"):
        rel = np.linalg.norm(res[name][0] - ref) / np.linalg.norm(ref)
        assert rel <= 1e-9, (name, rel)
    # with the learned guess active the autotuner still leaves the trajectory unchanged (its probes restore the
    # guess's input history with the rest of the solver state); two runs from a reset are identical
    runs = []
    for tune in (1, 0):
        x = gade.DeviceArray.from_host(np.tile(psi0, B))
        gp.reset_workspace()
        gp.cn_run_batch(x, params, 8, tol=1e-12)
        if tune:
            gp.autotune(x, params, 4, t0=8 * dt, reps=1)
        it, tols = gp.cn_run_batch(x, params, 12, t0=8 * dt, tol=1e-12)
        runs.append((x.get(), it, tols))
    assert np.array_equal(runs[0][0].view(np.float64), runs[1][0].view(np.float64))
    assert np.array_equal(runs[0][1], runs[1][1]) and np.array_equal(runs[0][2], runs[1][2])
    gp.set_param("guess_order", 0)


def test_pod_basis_leaves_out_empty_snapshots(datadir):
    """Snapshot columns that are zero (slots of instances whose pulse ended) are left out of the POD: the basis
    spans the top singular vectors of the non-zero columns (numpy SVD), orthonormal to 1e-12."""
    d = datadir("dipole")
    gp = gade.Propagator(d, 1, dt_of("dipole"))
    n = gp.n
    rng = np.random.default_rng(9)
    ncols, npod = 120, 64
    X = np.zeros((ncols, n), np.complex128)
    used = rng.permutation(ncols)[:90]
    # a decaying spectrum so the top-64 subspace is well separated
    W = (rng.standard_normal((90, 90)) + 1j * rng.standard_normal((90, 90))) * (0.7 ** np.arange(90))[None, :]
    Q, _ = np.linalg.qr(rng.standard_normal((n, 90)) + 1j * rng.standard_normal((n, 90)))
    X[used] = (Q @ W).T
    dX = gade.DeviceArray.from_host(X.ravel())
    sv = gp.lg_pod_build(dX, ncols, npod)
    U = gp.lg_basis(npod)
    sref = np.linalg.svd(X[used].T, compute_uv=False)
    assert np.allclose(sv[:90], sref, rtol=1e-10, atol=1e-12 * sref[0]) and np.all(sv[90:] == 0)
    assert np.abs(U.conj() @ U.T - np.eye(npod)).max() <= 1e-12
    Uref = np.linalg.svd(X[used].T, full_matrices=False)[0][:, :npod]
    # same subspace: every principal angle between the two bases is zero (cosines 1)
    cos = np.linalg.svd(Uref.conj().T @ U.T, compute_uv=False)
    assert np.all(np.abs(cos - 1) <= 1e-8), cos.min()
    # the basis moves to another model of the same couplings unchanged (gade_lg_set_basis / gade_lg_basis)
    g2 = gade.Propagator(d, 1, dt_of("dipole"))
    g2.lg_set_basis(U)
    assert np.array_equal(g2.lg_basis(npod).view(np.float64), U.view(np.float64))
