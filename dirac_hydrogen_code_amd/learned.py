"""Training driver of the learned initial guess.

The network, its inputs, the POD basis, training and inference run on the device
(csrc/learned.hip, through the C ABI in include/gade.h); this module sequences them:

1. solver snapshots: a seeded scan of training pulses is stepped with the order-2
   extrapolated guess, and the normalized extrapolation defects of chosen steps are
   kept (``nsnap`` columns);
2. POD: the top ``npod`` left singular vectors of the snapshots (QR + SVD on the device);
3. samples: the scan is stepped again, the network's inputs computed every step and one
   training sample recorded per instance and step (the quadratic form of the guess error);
4. training: Adam on mini-batches with the device forward/backward kernels.

The trained model then steps with ``guess_order`` GUESS_LEARNED; the converged solve at
the unchanged tolerance judges every guess, so only iteration counts change.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import gade


PULSE_RANGES = dict(E0=(1.0, 100.0), omega=(25.0, 75.0), cycles=(1, 15))


def scan_pulses(seed, B, dt, cep=None, ranges=None):
    """Seeded pulse scan: E0 log-uniform in [1, 100], omega uniform in [25, 75], carrier phase uniform in
    [0, 2 pi) (or the given cep for every pulse; 0 is the reference's own phase), cycles uniform in 1..15
    (other ranges through ``ranges``, keys E0, omega, cycles); per-instance step counts ceil(T_b / dt)."""
    r = dict(PULSE_RANGES, **(ranges or {}))
    rng = np.random.default_rng(seed)
    E0 = np.exp(rng.uniform(np.log(r["E0"][0]), np.log(r["E0"][1]), B))
    om = rng.uniform(r["omega"][0], r["omega"][1], B)
    ceps = rng.uniform(0.0, 2 * np.pi, B)
    if cep is not None:
        ceps = np.full(B, float(cep))
    cyc = rng.integers(r["cycles"][0], r["cycles"][1] + 1, B).astype(float)
    params = np.stack([E0, om, cyc, ceps], axis=1)
    nsteps_b = np.ceil(cyc * 2 * np.pi / om / dt).astype(np.int32)
    return params, nsteps_b


def step_scan(gp, psi0, params, nsteps_b, nsteps, after_step=None, tol=1e-12):
    """Steps the batch one step per call from psi0 (every instance), calling after_step(n, x) after step n
    (a true return stops the scan).  Step times accumulate as in a single run.  Returns x and the per-step
    iterations [steps][batch]."""
    B = gp.batch
    x = gade.DeviceArray.from_host(np.tile(psi0, B))
    gp.reset_workspace()
    t = 0.0
    its = []
    for n in range(nsteps):
        nb = (np.asarray(nsteps_b) > n).astype(np.int32)
        if not nb.any():
            break
        it, _ = gp.cn_run_batch(x, params, 1, nsteps_b=nb, t0=t, tol=tol)
        its.append(it[:, 0])
        t = t + gp.dt
        if after_step is not None and after_step(n, x):
            break
    return x, np.array(its)


class Trainer:
    """Builds and trains the learned guess of a batched model ``gp`` (its couplings, dt and batch width)."""

    def __init__(self, gp, psi0, seed=0x5EED, npod=gade.LG_POD, nsnap=400, nsamples=20000, log=None, tol=1e-13,
                 cep=None, ranges=None):
        """``tol``: tolerance of the solves that make the snapshots and samples.  It is tighter than the
        production one, so that a step whose extrapolated guess already meets the production tolerance (it
        ends at 0 iterations there) still gets a solution that differs from the guess: the sample keeps a
        non-zero extrapolation error and the loss learns to leave such steps alone.
        ``cep``: carrier phase of the training pulses (None: seeded per pulse).  A guess for the reference's
        own pulses (fixed phase, cep = 0) is trained with cep = 0: at that phase three of the six f_alpha start
        at zero and the early response is a family of its own, outside what random phases cover.
        ``ranges``: E0 / omega / cycles ranges of the training pulses (scan_pulses).  Short pulses put the
        peak of their envelope within the first steps the samples cover, so the guess also sees the strong-field
        regime of a long pulse's middle."""
        self.gp, self.psi0 = gp, np.asarray(psi0, np.complex128)
        self.seed, self.npod, self.nsnap, self.nsamples, self.tol = seed, npod, nsnap, nsamples, tol
        self.params, self.nsteps_b = scan_pulses(seed, gp.batch, gp.dt, cep, ranges)
        self.log = log or (lambda msg: None)
        self.N1 = gp.nblk * gp.blkdim

    def _mode(self, features):
        self.gp.set_param("guess_order", 2)
        self.gp.set_param("lg_features", 1 if features else 0)

    def build_pod(self, snap_every=20, snap_first=10):
        """Snapshots of the normalized extrapolation defect, then the POD basis; returns singular values.
        snap_every 0: spread the snapshots of each instance over the steps the samples cover."""
        if snap_every <= 0:
            span = min(int(np.median(self.nsteps_b)), self.nsamples // max(1, self.gp.batch))
            snap_every = max(20, (span - snap_first) // max(1, self.nsnap // max(1, self.gp.batch)))
        gp, B = self.gp, self.gp.batch
        self._mode(False)
        snaps = gade.DeviceArray(self.nsnap * self.N1)
        snaps.zero()
        L = gade.lib()
        cols = [0]

        def rec(n, x):
            if n >= snap_first and (n - snap_first) % snap_every == 0 and cols[0] < self.nsnap:
                gade.check(L.gade_lg_record(gp.ptr, x.ptr, None, snaps.ptr, cols[0], self.nsnap))
                cols[0] += B
            return cols[0] >= self.nsnap

        last = snap_first + snap_every * ((self.nsnap + B - 1) // B)
        step_scan(gp, self.psi0, self.params, self.nsteps_b, last + 1, rec, tol=self.tol)
        sv = gp.lg_pod_build(snaps, self.nsnap, self.npod)
        self.log(f"POD: {self.nsnap} snapshots, singular values {sv[0]:.3e} .. {sv[self.npod - 1]:.3e} "
                 f"(basis of {self.npod})")
        return sv

    def collect_samples(self):
        """One training sample per instance and step (valid ones have a full history and a finished
        solve) until nsamples are valid; returns the device buffer and the valid row indices."""
        gp, B = self.gp, self.gp.batch
        self._mode(True)
        max_steps = int(min(self.nsteps_b.max(), self.nsamples // max(1, B) * 3 + 16))
        buf = gade.DeviceArray(max_steps * B * gade.LG_SAMPLE, np.float32)
        buf.zero()
        L = gade.lib()
        count = [0]

        def rec(n, x):
            ptr = C.c_void_p(buf.ptr + n * B * gade.LG_SAMPLE * 4)
            gade.check(L.gade_lg_record(gp.ptr, x.ptr, ptr, None, 0, 0))
            slab = np.zeros(B * gade.LG_SAMPLE, np.float32)
            gade.check(L.gade_memcpy_d2h(slab.ctypes.data, ptr, slab.nbytes))
            count[0] += int(slab.reshape(B, gade.LG_SAMPLE)[:, gade.LG_SAMPLE_VALID].sum())
            return count[0] >= self.nsamples

        step_scan(gp, self.psi0, self.params, self.nsteps_b, max_steps, rec, tol=self.tol)
        host = buf.get().reshape(max_steps * B, gade.LG_SAMPLE)
        valid = np.nonzero(host[:, gade.LG_SAMPLE_VALID] > 0.5)[0]
        nzero = int(np.count_nonzero(host[valid, gade.LG_SAMPLE_BETA] <= 0))
        self.log(f"samples: {len(valid)} valid of {max_steps * B} recorded ({nzero} with a guess already at the "
                 f"solution: no weight in the loss)")
        self.nzero = nzero
        self.samples, self.samples_host, self.valid = buf, host, valid
        return buf, valid

    def standardize(self, rows):
        """Input mean / standard deviation over the given samples (used inputs; padding: 0 / 1).  The POD
        coordinates of the third and fourth differences share one scale: a coordinate's deviation is at least
        1 % of the largest in its block.  The weakest POD modes hardly vary over the training steps, and without
        the floor a state whose history excites them a little (a longer pulse's middle) would read as 100 sigma
        out of range and send every step to the fallback."""
        F = self.samples_host[rows, :gade.LG_IN].astype(np.float64)
        mean = F.mean(0)
        std = F.std(0)
        used = np.arange(gade.LG_IN) < gade.LG_FEATURES
        std = np.where(used & (std > 1e-6 * (np.abs(mean) + 1)), std, 1.0)
        z0 = gade.LG_FEATURES - 4 * gade.LG_POD  # the z3 block, then the z4 block, 2 * LG_POD reals each
        for b in (z0, z0 + 2 * gade.LG_POD):
            blk = std[b:b + 2 * gade.LG_POD]
            std[b:b + 2 * gade.LG_POD] = np.maximum(blk, 1e-2 * blk.max())
        mean = np.where(used, mean, 0.0)
        self.gp.lg_set_norm(mean, std)
        self.mean, self.std = mean, std
        return mean, std

    def export_to(self, other):
        """Installs the trained guess (inputs' standardization, network, POD basis) into another model of the
        same couplings, e.g. a single-instance model for the run that uses it."""
        other.lg_set_norm(self.mean, self.std)
        other.lg_set_params(self.gp.lg_get_params())
        other.lg_set_basis(self.gp.lg_basis(self.npod))

    def fit(self, epochs=30, M=256, lr=1e-3, lr_final=1e-4, holdout=0.1, init_seed=1):
        """Adam over the training rows (a held-out fraction is kept for validation); returns the per-step
        loss history and the validation loss before and after."""
        rng = np.random.default_rng(self.seed + 1)
        rows = rng.permutation(self.valid)
        nval = max(M, int(len(rows) * holdout) // M * M)
        val, train = rows[:nval], rows[nval:]
        self.standardize(train)
        self.gp.lg_init(init_seed)
        dval = gade.DeviceArray.from_host(val.astype(np.int32))
        v0 = self.validation_loss(dval, nval, M)
        steps_per_epoch = len(train) // M
        hist = []
        for e in range(epochs):
            perm = rng.permutation(train)[:steps_per_epoch * M].astype(np.int32)
            didx = gade.DeviceArray.from_host(perm)
            frac = e / max(1, epochs - 1)
            lr_e = float(lr * (lr_final / lr) ** frac)
            hist.append(self.gp.lg_train(self.samples, didx, M, steps_per_epoch, lr_e))
        v1 = self.validation_loss(dval, nval, M)
        self.log(f"training: {epochs} epochs x {steps_per_epoch} steps of {M}; validation loss {v0:.4f} -> {v1:.4f}")
        return np.concatenate(hist), v0, v1

    def validation_loss(self, dval, nval, M):
        tot = 0.0
        for k in range(nval // M):
            sub = gade.DeviceArray(M, np.int32)
            gade.check(gade.lib().gade_memcpy_d2d(sub.ptr, C.c_void_p(dval.ptr + k * M * 4), M * 4))
            tot += self.gp.lg_grad(self.samples, sub, M)
        return tot / (nval // M)
