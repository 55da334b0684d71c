"""GPU parity: the HIP path (libgade) against the CPU oracle on identical inputs.

Bars (stated per test):
  * structure and index maps (Hfull Ap/Ai/Ai_sub_map): bit-exact
  * operator applications: max-norm relative 1e-13
  * preconditioner solves: 1e-11 relative (different elimination order)
  * one CN step from the same psi_n, both at rel-tol 1e-12: ||psi_gpu - psi_ref|| / ||psi_ref|| <= 1e-10
"""
import sys

import numpy as np
import pytest

import oracle as O
from conftest import dt_of
from dirac_hydrogen_code_amd import gade

pytestmark = pytest.mark.gpu

LMAX = {"tiny": 2, "dipole": 1, "nondipole_small": 4}
CFGS = ["tiny", "nondipole_small", "dipole"]


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def relmax(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def pair(datadir, cfg, ntime=200):
    d = datadir(cfg)
    dt = dt_of(cfg, ntime)
    return d, dt, O.Model(d, LMAX[cfg], dt, nthreads=4), gade.Propagator(d, LMAX[cfg], dt)


# radial phase kernels: general ELL, banded on the vector units (the default), banded on the FP64 matrix cores,
# banded with the angular phase inside (ang_fused)
RADIAL = {"ell": (1, 0, 0), "dia": (2, 0, 0), "mfma": (2, 1, 0), "fused": (2, 0, 1)}


def use_radial(gp, kind):
    radial, mfma, fused = RADIAL[kind]
    gp.set_param("radial", radial)
    if radial == 2:
        gp.set_param("radial_mfma", mfma)
        gp.set_param("ang_fused", fused)
    assert gp.get_param("radial") == radial and (radial == 1 or gp.get_param("radial_mfma") == mfma)
    assert radial == 1 or gp.get_param("ang_fused") == fused


@pytest.mark.parametrize("kind", sorted(RADIAL))
@pytest.mark.parametrize("cfg", CFGS)
def test_fused_apply_matches_oracle(datadir, cfg, kind):
    d, dt, om, gp = pair(datadir, cfg)
    use_radial(gp, kind)
    say(f"fused apply {cfg} n={gp.n} radial={kind}")
    rng = np.random.default_rng(0)
    x = rng.standard_normal(gp.n) + 1j * rng.standard_normal(gp.n)
    y0 = rng.standard_normal(gp.n) + 1j * rng.standard_normal(gp.n)
    dx, dy = gade.DeviceArray.from_host(x), gade.DeviceArray(gp.n)
    for f in (gade.pulse(10, 50, 1, 0.01), rng.standard_normal(6) + 0j, np.zeros(6, complex)):
        om.timedep(f)
        gp.set_time(f)
        S, H = om.apply(x, "S"), om.apply(x, "H")
        for cH, want in ((1.0, S + H), (-1.0, S - H), (0.0, S)):
            gp.apply(dx, dy, cH=cH)
            assert relmax(dy.get(), want) <= 1e-13, (cfg, cH)
        dy.upload(y0)
        gp.apply(dx, dy, cH=1.0, beta=0.5)
        assert relmax(dy.get(), S + H + 0.5 * y0) <= 1e-13


@pytest.mark.parametrize("chunk", [16, 32])
@pytest.mark.parametrize("cfg", CFGS)
def test_psolve_matches_oracle(datadir, cfg, chunk):
    """The level-scheduled block-LU solve (16- or 32-row levels) against the oracle's banded solve."""
    d, dt, om, gp = pair(datadir, cfg)
    gp.set_param("psolve_chunk", chunk)
    say(f"psolve {cfg} chunk {chunk}")
    rng = np.random.default_rng(1)
    b = rng.standard_normal(gp.n) + 1j * rng.standard_normal(gp.n)
    want = om.psolve(b)
    db, dx = gade.DeviceArray.from_host(b), gade.DeviceArray(gp.n)
    gp.psolve(db, dx)
    assert relmax(dx.get(), want) <= 1e-11
    # and it inverts P = S + ihdt H0: with f = 0 the fused operator is exactly P
    gp.set_time(np.zeros(6, complex))
    dpx = gade.DeviceArray(gp.n)
    gp.apply(dx, dpx, cH=1.0)
    assert relmax(dpx.get(), b) <= 1e-12


@pytest.mark.parametrize("cfg", CFGS)
def test_bicgstab_matches_oracle(datadir, cfg):
    d, dt, om, gp = pair(datadir, cfg)
    say(f"bicgstab {cfg}")
    psi, _ = gade.read_psi0(d, gp.n)
    rng = np.random.default_rng(2)
    x0 = psi + 1e-3 * (rng.standard_normal(gp.n) + 1j * rng.standard_normal(gp.n))
    f = gade.pulse(10, 50, 1, 37 * dt)
    om.timedep(f)
    gp.set_time(f)
    rhs = om.apply(x0, "S") - om.apply(x0, "H")
    xo, it_o, te_o, na_o = om.bicgstab(rhs, x0, tol=1e-12)
    drhs, dx = gade.DeviceArray.from_host(rhs), gade.DeviceArray.from_host(x0)
    it_g, te_g, na_g = gp.bicgstab(drhs, dx, tol=1e-12)
    assert te_g < 1e-12
    assert abs(it_g - it_o) <= 1, (it_g, it_o)
    xg = dx.get()
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) <= 1e-10
    # true residual of the GPU answer, measured by the oracle operator
    r = rhs - om.apply(xg, "HS")
    assert np.linalg.norm(r) <= 1e-11 * np.linalg.norm(rhs)


@pytest.mark.parametrize("cfg", ["dipole", "nondipole_small"])
def test_cn_one_step_map_matches_oracle(datadir, cfg):
    """Each GPU CN step from the oracle's psi_n lands within 1e-10 of the oracle's psi_{n+1}."""
    d, dt, om, gp = pair(datadir, cfg)
    say(f"cn one-step map {cfg}")
    psi, _ = gade.read_psi0(d, gp.n)
    E0, w, cyc = 10.0, 50.0, 1.0 if cfg == "dipole" else 2.0
    x_ref = psi.copy()
    dpsi = gade.DeviceArray(gp.n)
    t = 0.0
    worst = 0.0
    for step in range(20):
        t = t + dt
        f = gade.pulse(E0, w, cyc, t)
        dpsi.upload(x_ref)
        gp.cn_step(dpsi, f, tol=1e-12)
        x_new, _, _, _ = om.cn_run(x_ref, E0, w, cyc, 1, t0=t - dt, tol=1e-12)
        err = np.linalg.norm(dpsi.get() - x_new) / np.linalg.norm(x_new)
        worst = max(worst, err)
        x_ref = x_new
    say(f"  worst one-step error {worst:.3e}")
    assert worst <= 1e-10


def test_cn_trajectory_dipole(datadir):
    """200 steps of the dipole configuration: drift vs the oracle, S-norm, per-step tolerance."""
    cfg = "dipole"
    d, dt, om, gp = pair(datadir, cfg)
    say("cn trajectory dipole 200 steps")
    psi, _ = gade.read_psi0(d, gp.n)
    xo, it_o, tol_o, _ = om.cn_run(psi, 10.0, 50.0, 1.0, 200, tol=1e-12)
    dpsi = gade.DeviceArray.from_host(psi)
    it_g, tol_g, usec = gp.cn_run(dpsi, 10.0, 50.0, 1.0, 200, tol=1e-12)
    xg = dpsi.get()
    drift = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    say(f"  drift {drift:.3e} iters gpu {it_g.mean():.2f} oracle {it_o.mean():.2f}")
    assert drift <= 1e-9
    assert np.all(tol_g < 1e-12)
    assert abs(gp.snorm(dpsi) - 1.0) < 1e-10
    assert np.abs(it_g - it_o).max() <= 1


@pytest.mark.parametrize("cfg", CFGS)
def test_assembled_structure_bitexact(datadir, cfg):
    d, dt, om, gp = pair(datadir, cfg)
    say(f"assembled structure {cfg}")
    f = gade.pulse(10, 50, 1, 0.02)
    gp.set_time(f)
    gp.assembled_timedep()
    Ap, Ai, sm, Ax = gp.assembled_get()
    oAp, oAi, osm = om.hfull_structure()
    assert np.array_equal(Ap, oAp)
    assert np.array_equal(Ai, oAi)
    assert np.array_equal(sm, osm)
    om.timedep(f)
    assert relmax(Ax, om.hfull_values()) <= 1e-13
    rng = np.random.default_rng(3)
    x = rng.standard_normal(gp.n) + 1j * rng.standard_normal(gp.n)
    dx, dy = gade.DeviceArray.from_host(x), gade.DeviceArray(gp.n)
    gp.assembled_spmv(dx, dy, alpha=-1.0, beta=0.0, with_S=True)  # (S - H) x as the rhs of the reference
    assert relmax(dy.get(), om.apply(x, "S") - om.apply(x, "H")) <= 1e-13


def test_tile_width_invariance(datadir):
    """Every radial tile width gives bit-identical operator results (per-row accumulation order is fixed)."""
    cfg = "nondipole_small"
    d, dt, om, gp = pair(datadir, cfg)
    say("tile width invariance")
    rng = np.random.default_rng(9)
    x = rng.standard_normal(gp.n) + 1j * rng.standard_normal(gp.n)
    dx, dy = gade.DeviceArray.from_host(x), gade.DeviceArray(gp.n)
    gp.set_time(gade.pulse(10, 50, 2, 0.03))
    use_radial(gp, "ell")
    outs = []
    for r in (1, 2, 4, 8):
        gp.set_param("tile_r", r)
        assert gp.get_param("tile_r") == r
        gp.apply(dx, dy, cH=1.0)
        outs.append(dy.get())
    for o in outs[1:]:
        assert np.array_equal(o.view(np.float64), outs[0].view(np.float64))


def test_radial_kernels_agree_and_switch(datadir):
    """The banded radial kernels (matrix cores and vector units) and the ELL kernel give the same operator
    (1e-14), the banded kernel with the angular phase inside (ang_fused) equals the two-kernel form bit for bit,
    and switching the kernel mid-trajectory re-derives the coefficients of the current f(t) in the new
    layout."""
    cfg = "nondipole_small"
    d, dt, om, gp = pair(datadir, cfg)
    say("radial kernels agree")
    rng = np.random.default_rng(11)
    x = rng.standard_normal(gp.n) + 1j * rng.standard_normal(gp.n)
    dx, dy = gade.DeviceArray.from_host(x), gade.DeviceArray(gp.n)
    f = gade.pulse(10, 50, 2, 0.07)
    gp.set_time(f)
    res = {}
    for kind in ("dia", "ell", "mfma", "fused", "dia", "mfma"):
        use_radial(gp, kind)
        for cH in (1.0, -1.0, 0.0):
            gp.apply(dx, dy, cH=cH)
            res.setdefault((kind, cH), []).append(dy.get())
    for cH in (1.0, -1.0, 0.0):
        a, b, c = res[("dia", cH)], res[("ell", cH)][0], res[("mfma", cH)]
        assert np.array_equal(a[0].view(np.float64), a[1].view(np.float64))
        assert np.array_equal(a[0].view(np.float64), res[("fused", cH)][0].view(np.float64))
        assert np.array_equal(c[0].view(np.float64), c[1].view(np.float64))
        assert relmax(a[0], b) <= 1e-14, cH
        assert relmax(c[0], b) <= 1e-14, cH
    om.timedep(f)
    assert relmax(res[("dia", 1.0)][0], om.apply(x, "S") + om.apply(x, "H")) <= 1e-13


@pytest.mark.parametrize("cfg,window,batch", [("nondipole_small", 8, 0), ("nondipole_small", 3, 1),
                                              ("nondipole_small", 1, 0), ("dipole", 8, 0), ("tiny", 4, 1)])
def test_device_loop_matches_host_loop(datadir, cfg, window, batch):
    """The device-driven stepping (decisions on the device, one host round trip per window, the rhs and
    A x0 from one paired apply) gives the host loop's trajectory (separate applies) bit for bit, including
    steps it has to hand back to the host (iter_batch 1); dipole: a single multipole grid."""
    d, dt, om, gp = pair(datadir, cfg)
    psi0, _ = gade.read_psi0(d, gp.n)
    say(f"device loop window={window} batch={batch}")
    out = {}
    for dev in (0, 1):
        gp.set_param("device_loop", dev)
        gp.set_param("window", window)
        gp.set_param("iter_batch", batch)
        gp.reset_workspace()
        dpsi = gade.DeviceArray.from_host(psi0)
        it, tols, us = gp.cn_run(dpsi, 10.0, 50.0, 1.0, 12, t0=0.0, tol=1e-12, timing=True)
        out[dev] = (dpsi.get(), it, tols)
    assert np.array_equal(out[0][0].view(np.float64), out[1][0].view(np.float64))
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])
    assert out[1][1].max() >= 1  # the window exercised real iterations


def test_device_loop_zero_rhs(datadir):
    """psi = 0 gives rhs = 0: both loops return x = 0 with 0 iterations and tolerance 0."""
    cfg = "tiny"
    d, dt, om, gp = pair(datadir, cfg)
    say("device loop zero rhs")
    for dev in (0, 1):
        gp.set_param("device_loop", dev)
        dpsi = gade.DeviceArray.from_host(np.zeros(gp.n, complex))
        it, tols, _ = gp.cn_run(dpsi, 10.0, 50.0, 1.0, 3, t0=0.0, tol=1e-12, timing=False)
        assert np.all(it == 0) and np.all(tols == 0)
        assert not np.any(dpsi.get())


def test_psolve_chunk_trajectory(datadir):
    """32-row preconditioner levels: the device-driven trajectory (FP64 and mixed, batched) matches the 16-row
    one to solver tolerance (the level inverses round differently) and meets the tolerance every step."""
    cfg = "dipole"
    d = datadir(cfg)
    dt = dt_of(cfg)
    say("psolve chunk trajectory")
    params = [(10.0, 50.0, 1.0, 0.0), (25.0, 40.0, 1.0, 1.3)]
    out = {}
    for chunk in (16, 32):
        for prec in (0, 1):
            gp = gade.Propagator(d, LMAX[cfg], dt, batch=2)
            gp.set_param("psolve_chunk", chunk)
            gp.set_param("precision", prec)
            psi0, _ = gade.read_psi0(d, gp.n)
            x = gade.DeviceArray.from_host(np.tile(psi0, 2))
            it, tols = gp.cn_run_batch(x, params, 30, tol=1e-12)
            assert gp.get_param("psolve_chunk") == chunk
            assert np.all(tols <= 1e-12)
            out[(chunk, prec)] = x.get()
    for prec in (0, 1):
        a, b = out[(32, prec)], out[(16, prec)]
        assert np.linalg.norm(a - b) / np.linalg.norm(b) <= 1e-10, prec


def test_iteration_budget_reached(datadir):
    """An unreachable tolerance with a small iteration budget: every solve stops at maxit (break-before-
    increment count = maxit, tolerance above the request) and the device-driven loop equals the host loop bit for
    bit; the mixed form stops at the same inner budget."""
    cfg = "nondipole_small"
    d, dt, om, gp = pair(datadir, cfg)
    psi0, _ = gade.read_psi0(d, gp.n)
    say("iteration budget")
    out = {}
    for dev in (0, 1):
        gp.set_param("device_loop", dev)
        gp.reset_workspace()
        x = gade.DeviceArray.from_host(psi0)
        it, tols, _ = gp.cn_run(x, 10.0, 50.0, 1.0, 5, tol=1e-18, maxit=2, timing=False)
        out[dev] = (x.get(), it, tols)
        assert np.all(it == 2) and np.all(tols > 1e-18), (dev, it, tols)
    assert np.array_equal(out[0][0].view(np.float64), out[1][0].view(np.float64))
    assert np.array_equal(out[0][2], out[1][2])
    gp.set_param("device_loop", 1)
    gp.set_param("precision", 1)
    gp.reset_workspace()
    x = gade.DeviceArray.from_host(psi0)
    it, tols, _ = gp.cn_run(x, 10.0, 50.0, 1.0, 5, tol=1e-18, maxit=3, timing=False)
    assert np.all(it == 3) and np.all(tols > 1e-18), (it, tols)
    gp.set_param("precision", 0)


@pytest.mark.parametrize("cfg", ["tiny", "nondipole_small"])
def test_populations_match_oracle(datadir, cfg):
    """A11: device eigenpairs (per distinct kappa) and eigen-projection populations against the oracle's
    scipy eigh + dense projection: per-block classes to 1e-10 of the block norm, energies to 1e-10."""
    d, dt, om, gp = pair(datadir, cfg)
    say(f"populations {cfg}")
    rng = np.random.default_rng(21)
    psi0, _ = gade.read_psi0(d, gp.n)
    x = rng.standard_normal(gp.n) + 1j * rng.standard_normal(gp.n)
    kap, ev = gp.eig_values()
    for v in (psi0, x):
        ref = O.populations(d, v)
        got = gp.populations(gade.DeviceArray.from_host(v), per_block=True)
        scale = np.maximum(ref["per_block"][:, 3], 1e-300)[:, None]
        assert np.abs(got["per_block"] - ref["per_block"]).max(axis=None) <= 1e-10 * max(ref["norm"], 1.0)
        assert np.all(np.abs(got["per_block"] - ref["per_block"]) <= 1e-9 * scale + 1e-14)
        for key in ("norm", "bound", "ionised"):
            assert abs(got[key] - ref[key]) <= 1e-10 * ref["norm"], key
    for k, E in zip(kap, ev):
        Eref = ref["energies"][int(k)]
        assert np.abs(E - Eref).max() <= 1e-10 * np.abs(Eref).max()


@pytest.mark.parametrize("order", [1, 2, 3])
def test_extrapolation_polynomial_exact(datadir, order):
    """The initial-guess extrapolation of order p reproduces any vector polynomial of degree <= p in t at the next
    step (equally spaced history), to rounding."""
    cfg = "tiny"
    d, dt, om, gp = pair(datadir, cfg)
    say(f"extrapolation order {order}")
    rng = np.random.default_rng(order)
    coef = [rng.standard_normal(gp.n) + 1j * rng.standard_normal(gp.n) for _ in range(order + 1)]
    traj = lambda t: sum(c * t ** k for k, c in enumerate(coef))
    gp.reset_workspace()
    d_x = gade.DeviceArray(gp.n)
    h = 0.37
    for k in range(order + 1):
        d_x.upload(traj(k * h))
        gp.extrapolate(d_x, order)
    want = traj((order + 1) * h)
    assert relmax(d_x.get(), want) <= 1e-12


def test_extrapolated_guess_same_solution_fewer_iterations(datadir):
    """guess_order 2: the device and host loops agree bit for bit, the trajectory equals the plain one to the
    solver tolerance, and the total iteration count does not grow."""
    cfg = "nondipole_small"
    d, dt, om, gp = pair(datadir, cfg)
    psi0, _ = gade.read_psi0(d, gp.n)
    say("extrapolated guess")
    out = {}
    for go, dev in ((0, 1), (2, 0), (2, 1)):
        gp.set_param("guess_order", go)
        gp.set_param("device_loop", dev)
        gp.reset_workspace()
        dpsi = gade.DeviceArray.from_host(psi0)
        it, tols, _ = gp.cn_run(dpsi, 10.0, 50.0, 2.0, 40, t0=0.0, tol=1e-12, timing=False)
        out[(go, dev)] = (dpsi.get(), it, tols)
        assert tols.max() < 1e-12
    gp.set_param("guess_order", 0)
    a, b = out[(2, 0)], out[(2, 1)]
    assert np.array_equal(a[0].view(np.float64), b[0].view(np.float64)) and np.array_equal(a[1], b[1])
    p0, p2 = out[(0, 1)][0], out[(2, 1)][0]
    assert np.linalg.norm(p2 - p0) / np.linalg.norm(p0) <= 1e-10
    say(f"  iterations plain {out[(0, 1)][1].sum()} extrapolated {out[(2, 1)][1].sum()}")
    assert out[(2, 1)][1].sum() <= out[(0, 1)][1].sum()


@pytest.mark.parametrize("cfg,batch", [("nondipole_small", 1), ("dipole", 3)])
def test_banded_tile_shapes_bit_identical(datadir, cfg, batch):
    """Every banded tile shape (rows per tile, outputs per lane, one or both output halves per workgroup),
    angular multipole group and angular form (coupling lists, Hall rows, or the lists inside the radial kernel:
    ang_fused) sums each output in the same order: applies (both signs, FP64) and whole device-driven
    trajectories (paired apply, FP32 inner solves of the mixed form) are bit-identical across shapes."""
    d = datadir(cfg)
    dt = dt_of(cfg)
    say(f"tile shapes {cfg} batch={batch}")
    rng = np.random.default_rng(3)
    out = {}
    # (rows per tile, outputs per lane, angular group, coupling lists, output halves per workgroup, fused)
    shapes = ((8, 2, 4, 1, 1, 0), (8, 2, 4, 0, 2, 0), (4, 2, 2, 1, 2, 0), (4, 2, 2, 0, 1, 0), (4, 4, 8, 1, 1, 0),
              (4, 4, 8, 0, 2, 0), (8, 1, 4, 1, 2, 0), (16, 1, 2, 0, 2, 0), (4, 1, 8, 1, 1, 0), (16, 1, 4, 1, 1, 0),
              (8, 2, 4, 1, 2, 1), (8, 1, 4, 1, 1, 1), (4, 2, 2, 1, 1, 1), (4, 1, 4, 1, 2, 1), (16, 1, 4, 1, 2, 1))
    for shape in shapes:
        gp = gade.Propagator(d, LMAX[cfg], dt, batch=batch)
        if gp.get_param("radial") != 2:
            pytest.skip("radial pattern not banded")
        gp.set_param("dia_outputs", 1)  # (rows, 1) is a valid shape for every row count
        gp.set_param("dia_rows", shape[0])
        gp.set_param("dia_outputs", shape[1])
        gp.set_param("ang_lb", shape[2])
        gp.set_param("ang_list", shape[3])
        gp.set_param("dia_halves", shape[4])
        gp.set_param("ang_fused", shape[5])
        assert gp.get_param("dia_halves") == shape[4] and gp.get_param("ang_fused") == shape[5]
        gp.set_time_batch(np.array([gade.pulse(10.0 + 5 * b, 50.0, 1.0, 0.013) for b in range(batch)]))
        n = gp.n * batch
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n) if not out else out["x"]
        out["x"] = x
        dx = gade.DeviceArray.from_host(x)
        ys = []
        for cH in (1.0, -1.0):
            dy = gade.DeviceArray(n)
            gp.apply(dx, dy, cH, 0.0)
            ys.append(dy.get())
        psi0, _ = gade.read_psi0(d, gp.n)
        params = [(10.0, 50.0, 1.0, 0.4 * b) for b in range(batch)]
        traj = []
        for prec in (0, 1):
            gp.set_param("precision", prec)
            p = gade.DeviceArray.from_host(np.tile(psi0, batch))
            gp.reset_workspace()
            it, tols = gp.cn_run_batch(p, params, 8, tol=1e-12)
            traj.append((p.get(), it, tols))
        out[shape] = (ys, traj)
    ref = out[shapes[0]]
    for shape in shapes[1:]:
        for a, b in zip(out[shape][0], ref[0]):
            assert np.array_equal(a.view(np.float64), b.view(np.float64)), shape
        for a, b in zip(out[shape][1], ref[1]):
            assert np.array_equal(a[0].view(np.float64), b[0].view(np.float64)), shape
            assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]), shape


@pytest.mark.parametrize("cfg", CFGS)
def test_mfma_radial_steps_agree(datadir, cfg):
    """The matrix-core radial kernel is available (the vector-unit kernel is the default FP64 form); CN steps
    with it and with the vector-unit kernel (device-driven stepping with the paired rhs / A x0 apply) agree to
    1e-11 over 12 steps, every step at the tolerance."""
    d, dt, om, gp = pair(datadir, cfg)
    assert gp.get_param("radial") == 2 and gp.get_param("radial_mfma") == 0
    say(f"mfma vs dia steps {cfg}")
    psi0, _ = gade.read_psi0(d, gp.n)
    out = []
    for mfma in (1, 0):
        g = gade.Propagator(d, LMAX[cfg], dt)
        g.set_param("radial_mfma", mfma)
        x = gade.DeviceArray.from_host(psi0)
        it, tols, _ = g.cn_run(x, 10.0, 50.0, 1.0, 12, tol=1e-12, timing=False)
        assert np.all(tols <= 1e-12)
        out.append(x.get())
    assert np.linalg.norm(out[0] - out[1]) <= 1e-11 * np.linalg.norm(out[1])

