"""ctypes bindings for the CPU oracle (liboracle.so) and oracle/_ref/libref.so.

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product path never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None

cplx_p = np.ctypeslib.ndpointer(np.complex128, flags="C_CONTIGUOUS")
int_p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
dbl_p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
long_p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: run `make -C {HERE}`")
        L = C.CDLL(path)
        L.or_ik.restype = C.c_int
        L.or_ik.argtypes = [C.c_int]
        L.or_imu.restype = C.c_double
        L.or_imu.argtypes = [C.c_int]
        L.or_pulse.argtypes = [C.c_double] * 4 + [cplx_p]
        L.or_locate.restype = C.c_int
        L.or_locate.argtypes = [int_p, C.c_int, C.c_int]
        L.or_unblock_nnz.restype = C.c_longlong
        L.or_unblock_nnz.argtypes = [C.c_int, int_p, int_p, C.c_int]
        L.or_unblock.argtypes = [C.c_int, int_p, int_p, C.c_int, int_p, int_p, int_p, int_p, int_p]
        L.or_model_load.restype = C.c_void_p
        L.or_model_load.argtypes = [C.c_char_p, C.c_int, C.c_double, C.c_double, C.c_int]
        L.or_model_free.argtypes = [C.c_void_p]
        L.or_model_set_threads.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.or_model_info.argtypes = [C.c_void_p] + [C.POINTER(C.c_int)] * 3 + [C.POINTER(C.c_longlong),
                                                                              C.POINTER(C.c_int)]
        for name in ("or_model_hfull_Ap", "or_model_hfull_Ai", "or_model_hfull_submap",
                     "or_model_S_Ap", "or_model_S_Ai"):
            getattr(L, name).restype = C.POINTER(C.c_int)
            getattr(L, name).argtypes = [C.c_void_p]
        for name in ("or_model_hfull_Ax", "or_model_S_Ax"):
            getattr(L, name).restype = C.POINTER(C.c_double)
            getattr(L, name).argtypes = [C.c_void_p]
        L.or_model_row_col.argtypes = [C.c_void_p, int_p, int_p, int_p, int_p, cplx_p, cplx_p]
        L.or_model_timedep.argtypes = [C.c_void_p, cplx_p]
        for name in ("or_model_spmv_HS", "or_model_spmv_H", "or_model_spmv_S"):
            getattr(L, name).argtypes = [C.c_void_p, cplx_p, cplx_p]
        L.or_model_matfree_H.argtypes = [C.c_void_p, cplx_p, cplx_p, cplx_p]
        L.or_model_precond_setup.restype = C.c_int
        L.or_model_precond_setup.argtypes = [C.c_void_p]
        L.or_model_psolve.argtypes = [C.c_void_p, cplx_p, cplx_p]
        L.or_bicgstab.argtypes = [C.c_void_p, cplx_p, cplx_p, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                  C.POINTER(C.c_int)]
        L.or_reset_workspace.argtypes = [C.c_void_p]
        L.or_cn_run.restype = C.c_int
        L.or_cn_run.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                C.c_double, C.c_int, cplx_p, int_p, dbl_p, long_p]
        L.or_vec_read.restype = C.c_int
        L.or_vec_read.argtypes = [C.c_char_p, C.c_int, cplx_p, int_p]
        _lib = L
    return _lib


def ref():
    """oracle/_ref/libref.so built from the reference sources, or None if absent."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.ik.restype = C.c_int
        R.ik.argtypes = [C.c_int]
        R.imu.restype = C.c_double
        R.imu.argtypes = [C.c_int]
        R.sorted_list_locate.restype = C.c_int
        R.sorted_list_locate.argtypes = [int_p, C.c_int, C.c_int]
        _ref = R
    return _ref


def pulse(E0, omega, cycles, t):
    out = np.zeros(6, np.complex128)
    lib().or_pulse(E0, omega, cycles, t, out)
    return out


class Model:
    """The reference CPU path on one input directory."""

    def __init__(self, directory, lmax, dt, h=1.0, nthreads=1, all_cores=False):
        """nthreads: OpenMP threads of the per-step assembly (the reference's only threaded part); with
        all_cores the SpMV rows and preconditioner blocks use them too (same results bit for bit)."""
        L = lib()
        self.ptr = L.or_model_load(directory.encode(), lmax, h, dt, nthreads)
        if not self.ptr:
            raise RuntimeError(f"oracle failed to load {directory}")
        L.or_model_set_threads(self.ptr, nthreads, 1 if all_cores else 0)
        nb, bd, hn, nz = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        nf = C.c_longlong()
        L.or_model_info(self.ptr, C.byref(nb), C.byref(bd), C.byref(hn), C.byref(nf), C.byref(nz))
        self.nblk, self.blkdim, self.hall_nnz, self.nnz_full, self.nnz_b = nb.value, bd.value, hn.value, nf.value, nz.value
        self.n = self.nblk * self.blkdim
        self.lmax = lmax
        self.dt = dt
        self._pre = False

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().or_model_free(self.ptr)
            self.ptr = None

    def _iarr(self, fn, n):
        return np.ctypeslib.as_array(fn(self.ptr), shape=(n,)).copy()

    def hfull_structure(self):
        L = lib()
        return (self._iarr(L.or_model_hfull_Ap, self.n + 1), self._iarr(L.or_model_hfull_Ai, self.nnz_full),
                self._iarr(L.or_model_hfull_submap, self.nnz_full))

    def hfull_values(self):
        raw = np.ctypeslib.as_array(lib().or_model_hfull_Ax(self.ptr), shape=(2 * self.nnz_full,))
        return raw.view(np.complex128).copy()

    def S_matrix(self):
        L = lib()
        nnz = self.nblk * self.nnz_b
        Ap = self._iarr(L.or_model_S_Ap, self.n + 1)
        Ai = self._iarr(L.or_model_S_Ai, nnz)
        Ax = np.ctypeslib.as_array(L.or_model_S_Ax(self.ptr), shape=(2 * nnz,)).view(np.complex128).copy()
        return Ap, Ai, Ax

    def row_col(self):
        h = self.hall_nnz
        a, b, ki, kj = (np.zeros(h, np.int32) for _ in range(4))
        H0 = np.zeros(h * self.lmax, np.complex128)
        H1 = np.zeros(h * self.lmax, np.complex128)
        lib().or_model_row_col(self.ptr, a, b, ki, kj, H0, H1)
        return a, b, ki, kj, H0, H1

    def timedep(self, ft):
        lib().or_model_timedep(self.ptr, np.ascontiguousarray(ft, np.complex128))

    def apply(self, x, which="HS"):
        out = np.zeros(self.n, np.complex128)
        getattr(lib(), f"or_model_spmv_{which}")(self.ptr, np.ascontiguousarray(x, np.complex128), out)
        return out

    def matfree_H(self, ft, x):
        out = np.zeros(self.n, np.complex128)
        lib().or_model_matfree_H(self.ptr, np.ascontiguousarray(ft, np.complex128),
                                 np.ascontiguousarray(x, np.complex128), out)
        return out

    def precond(self):
        if not self._pre:
            if lib().or_model_precond_setup(self.ptr):
                raise RuntimeError("oracle preconditioner setup failed")
            self._pre = True

    def psolve(self, b):
        self.precond()
        x = np.zeros(self.n, np.complex128)
        lib().or_model_psolve(self.ptr, np.ascontiguousarray(b, np.complex128), x)
        return x

    def bicgstab(self, rhs, x, tol=1e-12, maxit=500):
        self.precond()
        it, te, na = C.c_int(maxit), C.c_double(tol), C.c_int(0)
        x = np.ascontiguousarray(x, np.complex128).copy()
        lib().or_bicgstab(self.ptr, np.ascontiguousarray(rhs, np.complex128), x, C.byref(it), C.byref(te),
                          C.byref(na))
        return x, it.value, te.value, na.value

    def reset_workspace(self):
        lib().or_reset_workspace(self.ptr)

    def cn_run(self, psi, E0, omega, cycles, nsteps, t0=0.0, tol=1e-12, maxit=500):
        self.precond()
        x = np.ascontiguousarray(psi, np.complex128).copy()
        iters = np.zeros(nsteps, np.int32)
        tols = np.zeros(nsteps, np.float64)
        usec = np.zeros(nsteps, np.int64)
        lib().or_cn_run(self.ptr, E0, omega, cycles, t0, nsteps, tol, maxit, x, iters, tols, usec)
        return x, iters, tols, usec


def read_psi0_vec(path, n):
    out = np.zeros(n, np.complex128)
    perm = np.zeros(n, np.int32)
    if lib().or_vec_read(path.encode(), n, out, perm):
        raise RuntimeError(f"cannot read {path}")
    return out, perm


def _read_csr(path):
    """The .csr layout of tools/csr.c:262-318: int32 dim, nnz, Ap[dim+1], Ai[nnz], complex128 Ax[nnz],
    then optionally int32 perm[dim], npart, part[npart+1]."""
    raw = open(path, "rb").read()
    n, nnz = np.frombuffer(raw, np.int32, 2)
    off = 8
    Ap = np.frombuffer(raw, np.int32, n + 1, off)
    off += 4 * (n + 1)
    Ai = np.frombuffer(raw, np.int32, nnz, off)
    off += 4 * nnz
    Ax = np.frombuffer(raw, np.complex128, nnz, off)
    off += 16 * nnz
    perm = np.frombuffer(raw, np.int32, n, off) if len(raw) - off >= 4 * n else None
    return int(n), Ap, Ai, Ax, perm


def _dense(path):
    n, Ap, Ai, Ax, _ = _read_csr(path)
    M = np.zeros((n, n), np.complex128)
    rows = np.repeat(np.arange(n), np.diff(Ap))
    M[rows, Ai] = Ax
    return M


def populations(directory, psi, csq=137.035999084 ** 2):
    """Eigen-projection populations (src/hamiltonian.h:872-943, 1268-1307; src/process_output.cpp:332-355).

    For each block r of psi (Hall order, kappa from the H.csr permutation and ik of tools/solver_common.c:30-37):
    S = s0 + k s1 + k^2 s2, H0 = h00 + k h01 + k^2 h02 + k^3 h03 (tools/solver.c:409-458), generalized
    eigenpairs H0 v = E S v with v^H S v = 1 (LAPACK through scipy), population_j = |v_j^H S psi_r|^2.
    Returns dict(norm, bound, ionised, negative, per_block [nblk][4], energies {kappa: E})."""
    import scipy.linalg as sl
    s = [_dense(os.path.join(directory, f"s{i}.csr")) for i in range(3)]
    h = [_dense(os.path.join(directory, f"h0{i}.csr")) for i in range(4)]
    nblk, _, _, _, perm = _read_csr(os.path.join(directory, "H.csr"))
    perm = np.arange(nblk) if perm is None else perm
    n = s[0].shape[0]
    psi = np.asarray(psi).reshape(nblk, n)
    cache = {}
    out = np.zeros((nblk, 4))
    for r in range(nblk):
        k = lib().or_ik(int(perm[r]))
        if k not in cache:
            S = s[0] + k * s[1] + k * k * s[2]
            H = h[0] + k * h[1] + k * k * h[2] + k ** 3 * h[3]
            E, V = sl.eigh(H, S)
            cache[k] = (E, V, S)
        E, V, S = cache[k]
        w = S @ psi[r]
        p = np.abs(V.conj().T @ w) ** 2
        out[r] = [p[E < -csq].sum(), p[(E >= -csq) & (E < 0)].sum(), p[E >= 0].sum(), np.vdot(psi[r], w).real]
    return dict(norm=out[:, 3].sum(), bound=out[:, 1].sum(), ionised=out[:, 2].sum(), negative=out[:, 0].sum(),
                per_block=out, energies={k: v[0] for k, v in cache.items()})


def eigenpairs(directory):
    """Per kappa of the H.csr blocks: the generalized eigenpairs of the real parts of H0(k), S(k)
    (src/hamiltonian.h:906-943 takes .real() of both) from scipy's eigh, E ascending, v^T S v = 1, each
    eigenvector's largest-magnitude entry made positive.  Returns {kappa: (E, V, S)} with V[:, j] = v_j."""
    import scipy.linalg as sl
    s = [_dense(os.path.join(directory, f"s{i}.csr")).real for i in range(3)]
    h = [_dense(os.path.join(directory, f"h0{i}.csr")).real for i in range(4)]
    nblk, _, _, _, perm = _read_csr(os.path.join(directory, "H.csr"))
    perm = np.arange(nblk) if perm is None else perm
    out = {}
    for r in range(nblk):
        k = lib().or_ik(int(perm[r]))
        if k in out:
            continue
        S = s[0] + k * s[1] + k * k * s[2]
        H = h[0] + k * h[1] + k * k * h[2] + k ** 3 * h[3]
        E, V = sl.eigh(H, S)
        sign = np.sign(V[np.argmax(np.abs(V), axis=0), np.arange(V.shape[1])])
        out[k] = (E, V * sign, S)
    return out

