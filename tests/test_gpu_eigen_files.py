"""GPU tests of the eigen-projection outputs: the evl<kappa>.mat / evc<kappa>.mat files of
src/dumpEigvals.cpp:239-252 written from this library's eigenpairs, the psiev amplitudes psi_r^H S v_j of
src/hamiltonian.h:1246-1331 (eigProj, as savePsievs writes them), and the C driver's -x post-processing.

The oracle's eigenpairs are scipy's eigh (LAPACK) of the same real matrices with the same sign convention
(largest-magnitude entry positive).
Bars: energies to 1e-10 of the spectrum's scale; each eigenvector within 1e-12 * scale / gap of the oracle's in
the S-norm (gap: distance to the nearest other energy; first-order perturbation theory bounds both solvers'
errors by a few eps * scale / gap, and the observed agreement is within 0.6 of 1e-13 * scale / gap);
evc^T S evc = I to 1e-10;
psiev_j equal to the oracle's projection within |psi_r|_S * 1e-12 * scale / gap_j (Cauchy-Schwarz on the
eigenvector bar); sum_j |psiev_j|^2 = psi_r^H S psi_r to
1e-10 (completeness); the driver's psiev file holds the same numbers to 1e-15 relative (17-digit text)."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import dt_of, ROOT
from dirac_hydrogen_code_amd import gade

pytestmark = pytest.mark.gpu

LMAX = {"tiny": 2, "dipole": 1, "nondipole_small": 4}
DRIVER = os.path.join(ROOT, "dirac_hydrogen_code_amd", "bin", "dirac_solver")


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def dense_read(path, dtype):
    """dense_write layout: int64 rows, int64 cols, column-major data."""
    raw = open(path, "rb").read()
    rows, cols = np.frombuffer(raw, np.int64, 2)
    return np.frombuffer(raw, dtype, rows * cols, 16).reshape(cols, rows).T


@pytest.mark.parametrize("cfg", ["tiny", "nondipole_small"])
def test_eigen_files_match_oracle(datadir, tmp_path, cfg):
    d = datadir(cfg)
    say(f"evl/evc files {cfg}")
    gp = gade.Propagator(d, LMAX[cfg], dt_of(cfg))
    gp.eig_write(str(tmp_path))
    kap, ev = gp.eig_values()
    ref = O.eigenpairs(d)
    assert sorted(int(k) for k in kap) == sorted(ref)
    for k, E_dev in zip(kap, ev):
        E = dense_read(str(tmp_path / f"evl{int(k)}.mat"), np.float64)
        V = dense_read(str(tmp_path / f"evc{int(k)}.mat"), np.complex128)
        n = gp.blkdim
        assert E.shape == (n, 1) and V.shape == (n, n)
        assert np.array_equal(E[:, 0], E_dev)
        Er, Vr, S = ref[int(k)]
        scale = np.abs(Er).max()
        assert np.abs(E[:, 0] - Er).max() <= 1e-10 * scale
        assert np.all(V.imag == 0)
        assert np.abs(V.real.T @ S @ V.real - np.eye(n)).max() <= 1e-10
        gap = np.minimum(np.diff(Er, prepend=-np.inf), np.diff(Er, append=np.inf))
        dv = V.real - Vr
        dS = np.sqrt(np.einsum("ij,ij->j", dv, S @ dv))
        assert np.all(dS <= 1e-12 * scale / gap), (int(k), (dS * gap / scale).max())


@pytest.mark.parametrize("cfg", ["tiny", "nondipole_small"])
def test_psiev_matches_oracle(datadir, cfg):
    d = datadir(cfg)
    say(f"psiev {cfg}")
    gp = gade.Propagator(d, LMAX[cfg], dt_of(cfg))
    ref = O.eigenpairs(d)
    nblk, _, _, _, perm = O._read_csr(os.path.join(d, "H.csr"))
    psi0, _ = gade.read_psi0(d, gp.n)
    rng = np.random.default_rng(3)
    for psi in (psi0, rng.standard_normal(gp.n) + 1j * rng.standard_normal(gp.n)):
        got = gp.psiev(gade.DeviceArray.from_host(psi))
        blocks = psi.reshape(gp.nblk, gp.blkdim)
        for r in range(gp.nblk):
            E, V, S = ref[O.lib().or_ik(int(perm[r]))]
            want = np.conj(blocks[r]) @ (S @ V)
            nrm = np.vdot(blocks[r], S @ blocks[r]).real
            gap = np.minimum(np.diff(E, prepend=-np.inf), np.diff(E, append=np.inf))
            bar = np.sqrt(nrm) * 1e-12 * np.abs(E).max() / gap
            assert np.all(np.abs(got[r] - want) <= bar), (r, (np.abs(got[r] - want) / bar).max())
            assert abs(np.sum(np.abs(got[r]) ** 2) - nrm) <= 1e-10 * max(nrm, 1e-300), r


def test_driver_psiev_postprocessing(datadir, tmp_path):
    """bin/dirac_solver -f runs 12 steps writing x12.out; -x x12.out writes evl/evc and psiev_x12.out, whose
    numbers equal gade_psiev of the same state (permuted back into block order)."""
    cfg = "nondipole_small"
    d = str(tmp_path / "run")
    shutil.copytree(datadir(cfg), d)
    dt = dt_of(cfg)
    say("driver -x psiev")
    base = [DRIVER, "-d", d, "-t", repr(12 * dt), "-n", "12", "-l", "4", "-i", "10", "-o", "50", "-c", "2", "-f"]
    res = subprocess.run(base + ["-w", "12"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    res = subprocess.run(base + ["-x", os.path.join(d, "x12.out")], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    text = open(os.path.join(d, "psiev_x12.out")).read()
    raw = open(os.path.join(d, "x12.out"), "rb").read()
    n = int(np.frombuffer(raw, np.int32, 1)[0])
    xorig = np.frombuffer(raw, np.complex128, n, 8)
    gp = gade.Propagator(d, 4, dt)
    _, perm = gade.read_psi0(d, gp.n)
    got = gp.psiev(gade.DeviceArray.from_host(xorig[perm]))
    kap, ev = gp.eig_values()
    nblk, _, _, _, hperm = O._read_csr(os.path.join(d, "H.csr"))
    order = np.argsort(hperm, kind="stable")
    num = r"[-+0-9.eEinfa]+"
    recs = re.findall(r"psiev\[(\d+)\] = npy\.array\(\(\n\((.*?)\),\n\)\)", text)
    evls = re.findall(r"evl\[(\d+)\] = npy\.array\(\(\n(.*?)\n\)\)", text, flags=re.S)
    assert len(recs) == len(evls) == nblk
    for (th, body), (th2, ebody) in zip(recs, evls):
        r = order[int(th)]
        vals = np.array([complex(float(a), float(b)) for a, b in re.findall(rf"\(({num}),({num})\)", body)])
        assert vals.shape == (gp.blkdim,)
        assert np.abs(vals - got[r]).max() <= 1e-15 * np.abs(got[r]).max() + 1e-300
        E = np.array([float(v) for v in re.findall(rf"\(({num})\),", ebody)])
        k = O.lib().or_ik(int(hperm[r]))
        assert np.array_equal(E, ev[list(kap).index(k)])
    for k in kap:
        assert os.path.exists(os.path.join(d, f"evl{int(k)}.mat")) and os.path.exists(os.path.join(d, f"evc{int(k)}.mat"))
