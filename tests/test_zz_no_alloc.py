"""GPU test of gpu_bicgstab_noalloc, the BiCGSTAB entry point that allocates nothing (its workspace is the
library's fixed device array), against gpu_bicgstab with a fresh workspace, through the C ABI with the same
callbacks: the fused operator (gade_apply) as spmv and the block-LU preconditioner (gade_psolve) as psolve.

Two calls in a row, the second on a smaller system (the dipole couplings, N = 4400, then the tiny ones,
N = 360): solution, iterations and tol_error equal bit for bit (same kernels in the same order); both solve to
the tolerance.  A system beyond the array's 2^18 rows is refused with a message and nothing launched."""
import ctypes as C

import numpy as np
import pytest

from conftest import dt_of
from dirac_hydrogen_code_amd import gade

pytestmark = pytest.mark.gpu

SPMV = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double)
PSOLVE = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p)


class DenseVec(C.Structure):
    """include/gpu_sparse.h gpu_dense_vec_t"""
    _fields_ = [("desc", C.c_void_p), ("desc_local", C.c_void_p), ("dim", C.c_int), ("local_dim", C.c_int),
                ("local_offset", C.c_int), ("x", C.c_void_p)]


class Workspace(C.Structure):
    """include/gpu_bicgstab.h gpu_solver_workspace_t"""
    _fields_ = [(n, DenseVec) for n in ("r", "r0", "v", "p", "y", "z", "s", "t")]


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def solvers():
    L = gade.lib()
    # csr_data_t alpha, beta by value: each complex double is two SSE eightbytes, i.e. four double arguments
    L.gpu_bicgstab.argtypes = [SPMV, C.c_void_p, C.POINTER(DenseVec), C.POINTER(DenseVec), C.c_int, C.c_int, C.c_int,
                               PSOLVE, C.c_void_p, C.POINTER(Workspace), C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.gpu_bicgstab.restype = None
    L.gpu_bicgstab_noalloc.argtypes = [SPMV, C.c_void_p, C.POINTER(DenseVec), C.POINTER(DenseVec), C.c_int, C.c_int,
                                       C.c_int, PSOLVE, C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.gpu_bicgstab_noalloc.restype = C.c_int
    L.gpu_free_vec.argtypes = [C.POINTER(DenseVec)]
    L.gpu_free_vec.restype = None
    return L


def callbacks(gp, seen):
    L = gade.lib()

    def spmv(mat, x, out, ar, ai, br, bi):
        seen.append((ar, ai, br, bi))
        xv, ov = DenseVec.from_address(x), DenseVec.from_address(out)
        gade.check(L.gade_apply(gp.ptr, xv.x, ov.x, 1.0, br))  # out = (S + H) x + beta out, alpha = 1

    def psolve(p, rhs, x):
        rv, xv = DenseVec.from_address(rhs), DenseVec.from_address(x)
        dst = xv.x + 16 * xv.local_offset if xv.desc_local else xv.x
        gade.check(L.gade_psolve(gp.ptr, rv.x, dst))

    return SPMV(spmv), PSOLVE(psolve)


def vec(arr):
    d = gade.DeviceArray.from_host(arr)
    v = DenseVec()
    v.x, v.dim = d.ptr, d.n
    return d, v


def test_noalloc_equals_bicgstab_on_two_calls(datadir):
    L = solvers()
    for cfg, lmax in (("dipole", 1), ("tiny", 2)):
        d = datadir(cfg)
        dt = dt_of(cfg)
        gp = gade.Propagator(d, lmax, dt)
        n = gp.n
        say(f"noalloc vs bicgstab, {cfg} (n = {n})")
        gp.set_time(gade.pulse(10.0, 50.0, 1.0, 40 * dt))
        psi0, _ = gade.read_psi0(d, n)
        dpsi = gade.DeviceArray.from_host(psi0)
        drhs = gade.DeviceArray(n)
        gp.apply(dpsi, drhs, cH=-1.0)  # rhs = (S - H) psi
        rhs_v = DenseVec()
        rhs_v.x, rhs_v.dim = drhs.ptr, n
        seen = []
        spmv, psolve = callbacks(gp, seen)
        out = {}
        for name in ("bicgstab", "noalloc"):
            dx, xv = vec(psi0)
            it, te = C.c_int(500), C.c_double(1e-12)
            if name == "bicgstab":
                w = Workspace()
                L.gpu_bicgstab(spmv, None, C.byref(rhs_v), C.byref(xv), n, n, 0, psolve, None, C.byref(w),
                               C.byref(it), C.byref(te))
                for f in ("r", "r0", "v", "p", "y", "z", "s", "t"):
                    L.gpu_free_vec(C.byref(getattr(w, f)))
            else:
                rc = L.gpu_bicgstab_noalloc(spmv, None, C.byref(rhs_v), C.byref(xv), n, n, 0, psolve, None,
                                            C.byref(it), C.byref(te))
                assert rc == 0, L.gade_last_error().decode()
            out[name] = (dx.get(), it.value, te.value)
        assert all(s == (1.0, 0.0, 0.0, 0.0) for s in seen)  # alpha = 1, beta = 0, as gpu_bicgstab passes them
        (xa, ia, ta), (xb, ib, tb) = out["bicgstab"], out["noalloc"]
        say(f"  iterations {ia} / {ib}, tol_error {ta:.3e} / {tb:.3e}")
        assert np.array_equal(xa.view(np.float64), xb.view(np.float64))
        assert ia == ib and ta == tb and ta <= 1e-12
        # the solve is a solve: (S + H) x = rhs to the tolerance
        dy = gade.DeviceArray(n)
        gp.apply(gade.DeviceArray.from_host(xb), dy, cH=1.0)
        r = np.linalg.norm(dy.get() - drhs.get()) / np.linalg.norm(drhs.get())
        assert r <= 1e-11, r


def test_noalloc_refuses_oversize_system():
    L = solvers()
    big = (1 << 18) + 1
    it, te = C.c_int(10), C.c_double(1e-12)
    rhs, x = DenseVec(), DenseVec()
    rc = L.gpu_bicgstab_noalloc(SPMV(lambda *a: None), None, C.byref(rhs), C.byref(x), big, big, 0,
                                PSOLVE(lambda *a: None), None, C.byref(it), C.byref(te))
    assert rc == -1
    assert "262144" in L.gade_last_error().decode()
