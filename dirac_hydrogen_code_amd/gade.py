"""Python host mirror of the GaDE CN hot path over libgade.so (ctypes).

The interface follows the reference's driver loop (tools/dirac_solver.c):
load the coupling files, set f(t) per step, form rhs = (S - H) psi and run
BiCGSTAB on (S + H).  Everything numeric runs in libgade's HIP kernels; this
module only moves pointers.  It raises if the HIP library is not built or no
GPU is visible -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgade.so")

# every entry point declared in include/gade.h
EXPORTS = (
    "gade_last_error", "gade_malloc", "gade_free", "gade_memcpy_h2d", "gade_memcpy_d2h", "gade_memset",
    "gade_memcpy_d2d", "gade_device_sync", "gade_model_create", "gade_model_load", "gade_model_destroy", "gade_model_info",
    "gade_set_time", "gade_apply", "gade_precond_setup", "gade_psolve", "gade_bicgstab", "gade_reset_workspace",
    "gade_cn_step", "gade_cn_run", "gade_pulse", "gade_ik", "gade_imu", "gade_snorm", "gade_assembled_setup",
    "gade_assembled_timedep", "gade_assembled_spmv", "gade_assembled_get", "gade_profile", "gade_profile_get",
    "gade_set_param", "gade_get_param", "gade_eig_setup", "gade_eig_values", "gade_populations",
    "gade_extrapolate", "gade_model_create_batch", "gade_model_load_batch", "gade_model_batch", "gade_cn_run_batch",
    "gade_set_time_batch", "gade_pulse_cep", "gade_autotune", "gade_lg_record", "gade_lg_pod_build",
    "gade_lg_basis", "gade_lg_set_norm", "gade_lg_init", "gade_lg_set_params", "gade_lg_get_params",
    "gade_lg_forward", "gade_lg_grad", "gade_lg_train", "gade_lg_set_basis", "gade_lg_uses",
    "gade_psiev", "gade_eig_write", "gade_checkpoint_save", "gade_checkpoint_load", "gade_eig_dense",
)

# learned initial guess (include/gade.h)
GUESS_LEARNED = 4
LG_IN, LG_HIDDEN, LG_OUT, LG_POD, LG_SAMPLE, LG_PARAMS = 288, 256, 144, 64, 1232, 242576
LG_FEATURES = 17 + 4 * LG_POD                               # used inputs (the rest is zero padding)
LG_SAMPLE_BETA = LG_IN + 36 + 6 + 12 * LG_POD + 2 * LG_POD  # index of |e2|^2 / s^2 (the loss's normalizer)
LG_SAMPLE_VALID = LG_SAMPLE_BETA + 1                             # index of a sample's valid flag

GADE_TUNE_MAX = 96
GADE_TUNE_COLS = 9

_lib = None


class GadeError(RuntimeError):
    pass


def lib():
    """Load libgade.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GadeError(f"{LIB_PATH} is missing: build it with `make -C {HERE}` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, i32, f64, sz = C.c_void_p, C.c_int, C.c_double, C.c_size_t
        L.gade_last_error.restype = C.c_char_p
        sig = {
            "gade_malloc": [C.POINTER(vp), sz], "gade_free": [vp], "gade_memcpy_h2d": [vp, vp, sz],
            "gade_memcpy_d2h": [vp, vp, sz], "gade_memset": [vp, i32, sz], "gade_memcpy_d2d": [vp, vp, sz], "gade_eig_setup": [vp], "gade_extrapolate": [vp, vp, i32],
            "gade_model_load_batch": [C.c_char_p, i32, i32, f64, f64, vp, vp],
            "gade_cn_run_batch": [vp, vp, vp, vp, f64, i32, f64, i32, vp, vp],
            "gade_autotune": [vp, vp, vp, f64, i32, f64, i32, i32, vp, C.POINTER(i32), C.POINTER(i32)],
            "gade_lg_record": [vp, vp, vp, vp, i32, i32], "gade_lg_pod_build": [vp, vp, i32, i32, vp],
            "gade_lg_basis": [vp, vp], "gade_lg_set_basis": [vp, vp, i32], "gade_lg_set_norm": [vp, vp, vp], "gade_lg_init": [vp, C.c_uint],
            "gade_lg_set_params": [vp, vp], "gade_lg_get_params": [vp, vp], "gade_lg_forward": [vp, vp, i32, vp], "gade_lg_uses": [vp, vp],
            "gade_lg_grad": [vp, vp, vp, i32, C.POINTER(C.c_float), vp],
            "gade_lg_train": [vp, vp, vp, i32, i32, C.c_float, vp], "gade_set_time_batch": [vp, vp], "gade_eig_values": [vp, vp, vp, vp],
            "gade_populations": [vp, vp, C.POINTER(f64), C.POINTER(f64), C.POINTER(f64), vp], "gade_device_sync": [],
            "gade_model_load": [C.c_char_p, i32, f64, f64, vp, C.POINTER(vp)], "gade_model_destroy": [vp],
            "gade_model_info": [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), C.POINTER(i32),
                                C.POINTER(C.c_longlong), C.POINTER(i32)],
            "gade_set_time": [vp, vp], "gade_apply": [vp, vp, vp, f64, f64], "gade_precond_setup": [vp],
            "gade_psolve": [vp, vp, vp],
            "gade_bicgstab": [vp, vp, vp, C.POINTER(i32), C.POINTER(f64), C.POINTER(i32)],
            "gade_reset_workspace": [vp],
            "gade_cn_step": [vp, vp, vp, f64, i32, C.POINTER(i32), C.POINTER(f64)],
            "gade_cn_run": [vp, vp, f64, f64, f64, f64, i32, f64, i32, vp, vp, vp],
            "gade_snorm": [vp, vp, C.POINTER(f64), vp], "gade_assembled_setup": [vp],
            "gade_assembled_timedep": [vp], "gade_assembled_spmv": [vp, vp, vp, f64, f64, i32],
            "gade_assembled_get": [vp, vp, vp, vp, vp], "gade_profile": [vp, i32],
            "gade_profile_get": [vp, C.c_char_p, C.POINTER(f64), C.POINTER(C.c_longlong)],
            "gade_set_param": [vp, C.c_char_p, i32], "gade_get_param": [vp, C.c_char_p, C.POINTER(i32)],
            "gade_psiev": [vp, vp, vp], "gade_eig_write": [vp, C.c_char_p],
            "gade_checkpoint_save": [vp, vp, i32, f64, C.c_char_p],
            "gade_checkpoint_load": [vp, vp, C.POINTER(i32), C.POINTER(f64), C.c_char_p],
            "gade_eig_dense": [i32, vp, vp, vp, vp],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.gade_pulse.argtypes = [f64, f64, f64, f64, vp]
        L.gade_pulse.restype = None
        L.gade_pulse_cep.argtypes = [f64, f64, f64, f64, f64, vp]
        L.gade_pulse_cep.restype = None
        L.gade_ik.argtypes = [i32]
        L.gade_ik.restype = i32
        L.gade_imu.argtypes = [i32]
        L.gade_imu.restype = f64
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise GadeError(f"libgade error {rc}: {lib().gade_last_error().decode()}")


def pulse_cep(E0, omega, cycles, cep, t):
    """f_alpha(t) with carrier phase -pi/2 + cep as complex128[6]."""
    out = np.zeros(6, np.complex128)
    lib().gade_pulse_cep(E0, omega, cycles, cep, t, out.ctypes.data)
    return out


def pulse(E0, omega, cycles, t):
    """f_alpha(t), alpha=0..5 (tools/solver_common.c:46-65) as complex128[6]."""
    out = np.zeros(6, np.complex128)
    lib().gade_pulse(E0, omega, cycles, t, out.ctypes.data)
    return out


PROF = {"gsums": 1, "angular": 2, "radial": 4, "psolve": 8, "blas1": 16, "assemble": 32, "csr_spmv": 64, "radial_sum": 128,
        "radial_pair": 4}
PROF_ALL = 0x7fffffff


class DeviceArray:
    """A complex128 vector in HBM owned by libgade's allocator."""

    def __init__(self, n, dtype=np.complex128):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().gade_malloc(C.byref(p), self.nbytes))
        self.ptr = p.value

    @classmethod
    def from_host(cls, arr):
        arr = np.ascontiguousarray(arr)
        d = cls(arr.size, arr.dtype)
        d.upload(arr)
        return d

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, self.dtype)
        assert arr.size == self.n
        check(lib().gade_memcpy_h2d(self.ptr, arr.ctypes.data, self.nbytes))

    def get(self):
        out = np.empty(self.n, self.dtype)
        check(lib().gade_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def zero(self):
        check(lib().gade_memset(self.ptr, 0, self.nbytes))

    def free(self):
        if getattr(self, "ptr", None):
            lib().gade_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Propagator:
    """Device model of one coupling directory: fused A-apply, block-LU, BiCGSTAB, CN."""

    def __init__(self, directory, lmax, dt, h=1.0, batch=1):
        self.directory = directory
        p = C.c_void_p()
        check(lib().gade_model_load_batch(directory.encode(), lmax, batch, h, dt, None, C.byref(p)))
        self.ptr = p.value
        self.batch = batch
        nb, bd, hn, nz, nk = (C.c_int() for _ in range(5))
        nf = C.c_longlong()
        check(lib().gade_model_info(self.ptr, C.byref(nb), C.byref(bd), C.byref(hn), C.byref(nz), C.byref(nf),
                                    C.byref(nk)))
        self.nblk, self.blkdim, self.hall_nnz, self.nnz_b, self.nnz_full, self.nkappa = (
            nb.value, bd.value, hn.value, nz.value, nf.value, nk.value)
        self.n = self.nblk * self.blkdim
        self.lmax = lmax
        self.dt = dt

    def close(self):
        if getattr(self, "ptr", None):
            lib().gade_model_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def cn_run_batch(self, psi: DeviceArray, params, nsteps, nsteps_b=None, t0=0.0, tol=1e-12, maxit=500):
        """Batched pulse scan: params [batch][4] = (E0, omega, cycles, cep); psi holds batch*n entries.
        Returns iters, tols of shape [batch][nsteps] (entries past an instance's own count stay 0)."""
        params = np.ascontiguousarray(params, np.float64).reshape(self.batch, 4)
        iters = np.zeros((self.batch, nsteps), np.int32)
        tols = np.zeros((self.batch, nsteps), np.float64)
        nb = None if nsteps_b is None else np.ascontiguousarray(nsteps_b, np.int32)
        check(lib().gade_cn_run_batch(self.ptr, psi.ptr, params.ctypes.data, None if nb is None else nb.ctypes.data,
                                      t0, nsteps, tol, maxit, iters.ctypes.data, tols.ctypes.data))
        return iters, tols

    def autotune(self, psi: DeviceArray, params, nsteps, t0=0.0, tol=1e-12, maxit=500, reps=2):
        """Time every kernel shape and stepping candidate on nsteps probe steps from psi (unchanged), keep the
        fastest; every candidate must reproduce the same trajectory bit for bit.
        Returns (chosen, [dict(window, iter_batch, graph, dia_rows, dia_outputs, ang_lb, ms, dia_halves,
        ang_fused)])."""
        params = np.ascontiguousarray(params, np.float64).reshape(self.batch, 4)
        rep = np.zeros((GADE_TUNE_MAX, GADE_TUNE_COLS), np.float64)
        nc, ch = C.c_int(), C.c_int()
        check(lib().gade_autotune(self.ptr, psi.ptr, params.ctypes.data, t0, nsteps, tol, maxit, reps,
                                  rep.ctypes.data, C.byref(nc), C.byref(ch)))
        rows = [dict(window=int(r[0]), iter_batch=int(r[1]), graph=int(r[2]), dia_rows=int(r[3]),
                     dia_outputs=int(r[4]), ang_lb=int(r[5]), ms=float(r[6]), dia_halves=int(r[7]),
                     ang_fused=int(r[8]))
                for r in rep[:nc.value]]
        return ch.value, rows

    # ---- learned initial guess ----
    def lg_record(self, psi: DeviceArray, samples=None, snaps=None, snap_col=0, snap_cols=0):
        check(lib().gade_lg_record(self.ptr, psi.ptr, samples.ptr if samples is not None else None,
                                   snaps.ptr if snaps is not None else None, snap_col, snap_cols))

    def lg_pod_build(self, snaps: DeviceArray, nsnap, npod=LG_POD):
        sv = np.zeros(nsnap, np.float64)
        check(lib().gade_lg_pod_build(self.ptr, snaps.ptr, nsnap, npod, sv.ctypes.data))
        return sv

    def lg_basis(self, npod=LG_POD):
        U = np.zeros((npod, self.nblk * self.blkdim), np.complex128)
        check(lib().gade_lg_basis(self.ptr, U.ctypes.data))
        return U

    def lg_set_basis(self, U):
        U = np.ascontiguousarray(U, np.complex128)
        assert U.shape[1] == self.nblk * self.blkdim
        check(lib().gade_lg_set_basis(self.ptr, U.ctypes.data, U.shape[0]))

    def lg_set_norm(self, mean, std):
        mean = np.ascontiguousarray(mean, np.float32)
        std = np.ascontiguousarray(std, np.float32)
        check(lib().gade_lg_set_norm(self.ptr, mean.ctypes.data, std.ctypes.data))

    def lg_init(self, seed=0):
        check(lib().gade_lg_init(self.ptr, seed))

    def lg_set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == LG_PARAMS
        check(lib().gade_lg_set_params(self.ptr, p.ctypes.data))

    def lg_get_params(self):
        p = np.zeros(LG_PARAMS, np.float32)
        check(lib().gade_lg_get_params(self.ptr, p.ctypes.data))
        return p

    def lg_uses(self):
        """Per instance: steps since the last workspace reset that took the network, and those that the input
        range gate (set_param "lg_zmax") sent to the extrapolation."""
        c = np.zeros((self.batch, 2), np.int32)
        check(lib().gade_lg_uses(self.ptr, c.ctypes.data))
        return c

    def lg_forward(self, x: DeviceArray, M, y: DeviceArray):
        check(lib().gade_lg_forward(self.ptr, x.ptr, M, y.ptr))

    def lg_grad(self, samples: DeviceArray, idx: DeviceArray, M, grad: DeviceArray = None):
        loss = C.c_float()
        check(lib().gade_lg_grad(self.ptr, samples.ptr, idx.ptr, M, C.byref(loss),
                                 grad.ptr if grad is not None else None))
        return loss.value

    def lg_train(self, samples: DeviceArray, idx: DeviceArray, M, nsteps, lr):
        hist = np.zeros(nsteps, np.float32)
        check(lib().gade_lg_train(self.ptr, samples.ptr, idx.ptr, M, nsteps, lr, hist.ctypes.data))
        return hist

    def set_time_batch(self, ft):
        ft = np.ascontiguousarray(ft, np.complex128).reshape(self.batch, 6)
        check(lib().gade_set_time_batch(self.ptr, ft.ctypes.data))

    def set_time(self, ft):
        ft = np.ascontiguousarray(ft, np.complex128)
        check(lib().gade_set_time(self.ptr, ft.ctypes.data))

    def apply(self, x: DeviceArray, y: DeviceArray, cH=1.0, beta=0.0):
        check(lib().gade_apply(self.ptr, x.ptr, y.ptr, cH, beta))

    def psolve(self, b: DeviceArray, x: DeviceArray):
        check(lib().gade_psolve(self.ptr, b.ptr, x.ptr))

    def bicgstab(self, rhs: DeviceArray, x: DeviceArray, tol=1e-12, maxit=500):
        it, te, na = C.c_int(maxit), C.c_double(tol), C.c_int(0)
        check(lib().gade_bicgstab(self.ptr, rhs.ptr, x.ptr, C.byref(it), C.byref(te), C.byref(na)))
        return it.value, te.value, na.value

    def reset_workspace(self):
        check(lib().gade_reset_workspace(self.ptr))

    def cn_step(self, psi: DeviceArray, ft, tol=1e-12, maxit=500):
        ft = np.ascontiguousarray(ft, np.complex128)
        it, te = C.c_int(0), C.c_double(0)
        check(lib().gade_cn_step(self.ptr, psi.ptr, ft.ctypes.data, tol, maxit, C.byref(it), C.byref(te)))
        return it.value, te.value

    def cn_run(self, psi: DeviceArray, E0, omega, cycles, nsteps, t0=0.0, tol=1e-12, maxit=500, timing=True):
        iters = np.zeros(nsteps, np.int32)
        tols = np.zeros(nsteps, np.float64)
        usec = np.zeros(nsteps, np.float64)
        check(lib().gade_cn_run(self.ptr, psi.ptr, E0, omega, cycles, t0, nsteps, tol, maxit, iters.ctypes.data,
                                tols.ctypes.data, usec.ctypes.data if timing else None))
        return iters, tols, usec

    def snorm(self, x: DeviceArray, populations=False):
        nrm = C.c_double()
        pop = np.zeros(self.nblk) if populations else None
        check(lib().gade_snorm(self.ptr, x.ptr, C.byref(nrm), pop.ctypes.data if populations else None))
        return (nrm.value, pop) if populations else nrm.value

    def extrapolate(self, psi: DeviceArray, order):
        check(lib().gade_extrapolate(self.ptr, psi.ptr, order))

    def populations(self, x: DeviceArray, per_block=False):
        """Eigen-projection populations: dict(norm, bound, ionised[, per_block [nblk][4]])."""
        nrm, b, i = C.c_double(), C.c_double(), C.c_double()
        pb = np.zeros((self.nblk, 4)) if per_block else None
        check(lib().gade_populations(self.ptr, x.ptr, C.byref(nrm), C.byref(b), C.byref(i),
                                     pb.ctypes.data if per_block else None))
        out = dict(norm=nrm.value, bound=b.value, ionised=i.value)
        if per_block:
            out["per_block"] = pb
        return out

    def psiev(self, x: DeviceArray):
        """Eigen-state amplitudes psi_r^H S v_j per block r and eigenpair j of its kappa: complex [nblk][blkdim]."""
        out = np.zeros((self.nblk, self.blkdim), np.complex128)
        check(lib().gade_psiev(self.ptr, x.ptr, out.ctypes.data))
        return out

    def eig_write(self, directory):
        """evl<kappa>.mat / evc<kappa>.mat of every distinct kappa into directory."""
        check(lib().gade_eig_write(self.ptr, directory.encode()))

    def checkpoint_save(self, x: DeviceArray, step, time, path):
        check(lib().gade_checkpoint_save(self.ptr, x.ptr, int(step), float(time), path.encode()))

    def checkpoint_load(self, x: DeviceArray, path):
        """Restores psi and the solver state; returns (step, time)."""
        st, t = C.c_int(), C.c_double()
        check(lib().gade_checkpoint_load(self.ptr, x.ptr, C.byref(st), C.byref(t), path.encode()))
        return st.value, t.value

    def eig_values(self):
        """(kappa[nkappa], energies[nkappa][blkdim]) of H0(k) v = E S(k) v."""
        check(lib().gade_eig_setup(self.ptr))
        kap = np.zeros(self.nkappa, np.int32)
        ev = np.zeros((self.nkappa, self.blkdim))
        nk = C.c_int()
        check(lib().gade_eig_values(self.ptr, C.byref(nk), kap.ctypes.data, ev.ctypes.data))
        return kap[:nk.value], ev[:nk.value]

    # assembled reference formulation
    def assembled_timedep(self):
        check(lib().gade_assembled_timedep(self.ptr))

    def assembled_spmv(self, x: DeviceArray, y: DeviceArray, alpha=1.0, beta=0.0, with_S=False):
        check(lib().gade_assembled_spmv(self.ptr, x.ptr, y.ptr, alpha, beta, int(with_S)))

    def assembled_get(self, submap=True, values=True):
        Ap = np.zeros(self.n + 1, np.int32)
        Ai = np.zeros(self.nnz_full, np.int32)
        sm = np.zeros(self.nnz_full, np.int32) if submap else None
        Ax = np.zeros(self.nnz_full, np.complex128) if values else None
        check(lib().gade_assembled_get(self.ptr, Ap.ctypes.data, Ai.ctypes.data,
                                       sm.ctypes.data if submap else None, Ax.ctypes.data if values else None))
        return Ap, Ai, sm, Ax

    def set_param(self, name, value):
        check(lib().gade_set_param(self.ptr, name.encode(), int(value)))

    def get_param(self, name):
        v = C.c_int()
        check(lib().gade_get_param(self.ptr, name.encode(), C.byref(v)))
        return v.value

    def profile(self, families="all"):
        """Enable per-family event timing: 'all', a list of family names, or a false value to stop."""
        if not families:
            mask = 0
        elif families == "all" or families is True:
            mask = PROF_ALL
        else:
            mask = 0
            for f in ([families] if isinstance(families, str) else families):
                mask |= PROF[f]
        check(lib().gade_profile(self.ptr, mask))

    def profile_get(self, name):
        t, c = C.c_double(), C.c_longlong()
        check(lib().gade_profile_get(self.ptr, name.encode(), C.byref(t), C.byref(c)))
        return t.value, c.value


def read_psi0(directory, n):
    """psi0.vec: int32 n, complex128[n], int32 perm[n] (tools/dirac_solver.c:68-108)."""
    raw = open(os.path.join(directory, "psi0.vec"), "rb").read()
    nn = int(np.frombuffer(raw, np.int32, 1)[0])
    if nn != n:
        raise GadeError(f"psi0.vec has {nn} entries, the model has {n}")
    psi = np.frombuffer(raw, np.complex128, n, 4).copy()
    perm = np.frombuffer(raw, np.int32, n, 4 + 16 * n).copy()
    return psi, perm
