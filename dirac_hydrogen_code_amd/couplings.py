"""Coupling-file generator: writes the file set the CN propagator reads.

The propagator consumes the files that the reference's offline generator
``dumpCouplings`` (src/dumpBdpCouplings.cpp:129-328) and preprocessor
``tools/main.py`` produce:

* ``H0l{l}.csr``, ``H1l{l}.csr``   angular couplings, ul=+1 / ul=-1
  (src/dumpBdpCouplings.cpp:207-217 -> src/spnrbasis.cpp:395-485)
* ``g{n}a{alpha}l{l}.csr``        radial interaction integrals, one shared pattern
  (src/dumpBdpCouplings.cpp:239-265)
* ``h0{n}.csr`` (n<4), ``s{n}.csr`` (n<3)   field-free H0 and overlap, kappa-polynomial
  (src/dumpBdpCouplings.cpp:289-305)
* ``psi0``                         ground state, int64 rows, int64 cols, complex128[]
  (src/dumpBdpCouplings.cpp:84-108, 307-324)
* ``H.csr``, ``psi0.vec``          block pattern + perm trailer, permuted psi0
  (tools/main.py:52-171)

The angular matrices follow the reference's formulas exactly (Wigner-3j
selection rules and phases of spnrbasis::bdpK/L/M/N), with a Wigner-3j of our
own (Racah's formula in exact integer arithmetic).  The radial matrices are this
project's own quadrature of the reference's construction: a degree-7 B-spline
dual-kinetic-balance basis on the same knots (src/dumpBdpCouplings.cpp:194-198,
src/splinehandler.cpp:26-52, first and last spline dropped) with the P / Q
kappa-parts of src/dkbbasis.cpp (dkbCache, bdpalphmat), overlap and field-free
Dirac H0 by Gauss-Legendre quadrature, and interaction integrals P^T W C_alpha Q
with the spherical Bessel kernels of src/potential.h:470-497 (one wave number per
alpha pair).  The field-free spectrum is the Dirac hydrogen spectrum (1s to
1e-11); the values are not compared with files the reference wrote (it ships
none, and dumpCouplings cannot be built here).
"""
from __future__ import annotations

import json
import math
import os
from fractions import Fraction
from functools import lru_cache

import numpy as np

SOL = 137.035999084  # speed of light in a.u. (tools/solver_common.h:37)


# --------------------------------------------------------------------------
# angular index maps  (tools/solver_common.c:30-44, src/spnrbasis.cpp:610-652)
# --------------------------------------------------------------------------
def ik(i: int) -> int:
    ii = i // 4
    abskappa = int(0.5 * (math.sqrt(8.0 * ii + 1.0) - 1.0)) + 1
    sgnmod = max(4, abskappa * 4)
    sgnfloor = 2.0 * abskappa * (abskappa - 1)
    q = (i - sgnfloor) / sgnmod
    sgnkappa = int(q >= 0.5) - int(q < 0.5)
    return abskappa * sgnkappa


def imu(i: int) -> float:
    abskappa = abs(ik(i))
    kmod = max(2, 2 * abskappa)
    return (i % kmod) - abskappa + 0.5


def n_angular(nkappa: int) -> int:
    """spnrbasis::aqn (src/spnrbasis.cpp:87-93)."""
    return 4 if nkappa == 1 else 2 * (nkappa + 1) * nkappa


# --------------------------------------------------------------------------
# Wigner 3j, Racah's formula on doubled quantum numbers, exact arithmetic
# --------------------------------------------------------------------------
@lru_cache(maxsize=None)
def _fact(n: int) -> int:
    return math.factorial(n)


@lru_cache(maxsize=None)
def wigner3j_2(tj1, tj2, tj3, tm1, tm2, tm3) -> float:
    """Wigner 3j symbol with every argument given as twice its value."""
    if tm1 + tm2 + tm3 != 0:
        return 0.0
    for tj, tm in ((tj1, tm1), (tj2, tm2), (tj3, tm3)):
        if tj < 0 or abs(tm) > tj or (tj + tm) % 2:
            return 0.0
    if (tj1 + tj2 + tj3) % 2 or tj3 < abs(tj1 - tj2) or tj3 > tj1 + tj2:
        return 0.0
    a1 = (tj1 + tj2 - tj3) // 2
    a2 = (tj1 - tj2 + tj3) // 2
    a3 = (-tj1 + tj2 + tj3) // 2
    b = (tj1 + tj2 + tj3) // 2 + 1
    delta = Fraction(_fact(a1) * _fact(a2) * _fact(a3), _fact(b))
    pref = (_fact((tj1 + tm1) // 2) * _fact((tj1 - tm1) // 2) * _fact((tj2 + tm2) // 2)
            * _fact((tj2 - tm2) // 2) * _fact((tj3 + tm3) // 2) * _fact((tj3 - tm3) // 2))
    # summation bounds
    kmin = max(0, (tj2 - tj3 - tm1) // 2, (tj1 - tj3 + tm2) // 2)
    kmax = min(a1, (tj1 - tm1) // 2, (tj2 + tm2) // 2)
    s = Fraction(0)
    for k in range(kmin, kmax + 1):
        den = (_fact(k) * _fact((tj3 - tj2 + tm1) // 2 + k) * _fact((tj3 - tj1 - tm2) // 2 + k)
               * _fact(a1 - k) * _fact((tj1 - tm1) // 2 - k) * _fact((tj2 + tm2) // 2 - k))
        s += Fraction((-1) ** k, den)
    if s == 0:
        return 0.0
    sq = delta * pref * s * s
    sign = (-1) ** ((tj1 - tj2 - tm3) // 2) * (1 if s > 0 else -1)
    return sign * math.sqrt(float(sq))


def wigner3j(j1, j2, j3, m1, m2, m3) -> float:
    return wigner3j_2(int(round(2 * j1)), int(round(2 * j2)), int(round(2 * j3)),
                      int(round(2 * m1)), int(round(2 * m2)), int(round(2 * m3)))


def _ipow(n: int) -> complex:
    """i**n with C's remainder semantics, as src/spnrbasis.cpp:486-507."""
    r = int(math.fmod(n, 4))
    return {-3: 1j, -2: -1, -1: -1j, 0: 1, 1: 1j, 2: -1, 3: -1j}[r]


def _sgn(v: float) -> int:
    return (v > 0) - (v < 0)


def _bdp_term(ki, kj, mui, muj, l, kind):
    """spnrbasis::bdpK/L/M/N (src/spnrbasis.cpp:509-608)."""
    if kind in ("K", "L"):
        a, b = abs(ki + 0.5), abs(kj - 0.5)
    else:
        a, b = abs(ki - 0.5), abs(kj + 0.5)
    if not (abs(a - b) <= l <= a + b - 1):
        return 0j
    la, lb = a - 0.5, b - 0.5
    if kind == "K":
        if int(2 * muj) != int(2 * mui) - 2:
            return 0j
        ph, rad = _sgn(ki), abs(ki + 0.5 - mui) * abs(kj - 0.5 - muj)
        m1, m2 = -mui + 0.5, muj + 0.5
    elif kind == "L":
        if int(2 * muj) != int(2 * mui) + 2:
            return 0j
        ph, rad = _sgn(kj), abs(ki + 0.5 + mui) * abs(kj - 0.5 + muj)
        m1, m2 = -mui - 0.5, muj - 0.5
    elif kind == "M":
        if int(2 * muj) != int(2 * mui) - 2:
            return 0j
        ph, rad = _sgn(ki), abs(ki - 0.5 + mui) * abs(kj + 0.5 + muj)
        m1, m2 = -mui + 0.5, muj + 0.5
    else:  # "N"
        if int(2 * muj) != int(2 * mui) + 2:
            return 0j
        ph, rad = _sgn(kj), abs(ki - 0.5 - mui) * abs(kj + 0.5 - muj)
        m1, m2 = -mui - 0.5, muj - 0.5
    z = complex(ph * _ipow(int(2 * mui + l)))
    v = z * (2.0 * l + 1.0)
    v = v * math.sqrt(rad)
    v = v * wigner3j(la, lb, l, 0, 0, 0)
    v = v * wigner3j(la, lb, l, m1, m2, 0)
    return v


def angular_matrix(nkappa: int, nmu: int, ul: int, l: int):
    """spnrbasis::bdpalphsigmaXmat(ul, l) (src/spnrbasis.cpp:395-485) as CSR arrays."""
    n = n_angular(nkappa)
    kap = [ik(i) for i in range(n)]
    mu = [imu(i) for i in range(n)]
    scale = 1j if l % 2 == 1 else -1.0
    kinds = ("K", "L") if ul == 1 else ("M", "N")
    # only |mu_i - mu_j| == 1 can couple; bucket columns by 2*mu
    by_mu = {}
    for j in range(n):
        by_mu.setdefault(int(2 * mu[j]), []).append(j)
    Ap = [0]
    Ai, Ax = [], []
    for i in range(n):
        row = []
        if nmu == 0 or abs(mu[i]) < nmu:
            cands = sorted(by_mu.get(int(2 * mu[i]) - 2, []) + by_mu.get(int(2 * mu[i]) + 2, []))
            for j in cands:
                if nmu != 0 and not abs(mu[j]) < nmu:
                    continue
                t1 = _bdp_term(kap[i], kap[j], mu[i], mu[j], l, kinds[0])
                t2 = _bdp_term(kap[i], kap[j], mu[i], mu[j], l, kinds[1])
                if t1 != 0 or t2 != 0:
                    row.append((j, scale * (t1 + t2)))
        for j, v in row:
            Ai.append(j)
            Ax.append(v)
        Ap.append(len(Ai))
    return (np.array(Ap, np.int32), np.array(Ai, np.int32), np.array(Ax, np.complex128), n)


# --------------------------------------------------------------------------
# radial part: degree-7 B-spline DKB-style basis
# --------------------------------------------------------------------------
def _gauss_legendre_points(t, nq):
    xg, wg = np.polynomial.legendre.leggauss(nq)
    a, b = t[:-1], t[1:]
    x = (0.5 * (b - a)[:, None] * (xg[None, :] + 1.0) + a[:, None]).ravel()
    w = (0.5 * (b - a)[:, None] * wg[None, :]).ravel()
    return x, w


def _bspline_cache(t, deg, x):
    from scipy.interpolate import BSpline
    knots = np.concatenate([np.full(deg, t[0]), t, np.full(deg, t[-1])])
    nspl = len(knots) - 1 - deg
    eye = np.eye(nspl)
    spl = BSpline(knots, eye, deg, extrapolate=False)
    B = np.nan_to_num(spl(x))
    dB = np.nan_to_num(spl.derivative(1)(x))
    ddB = np.nan_to_num(spl.derivative(2)(x))
    # drop first and last spline (src/splinehandler.cpp:441-443)
    return B[:, 1:-1], dB[:, 1:-1], ddB[:, 1:-1]


def radial_pattern(ns: int, half_band: int = 7):
    """Four banded quadrants of half-bandwidth 7 over the 2*ns radial functions."""
    n = 2 * ns
    Ap = [0]
    Ai = []
    for a in range(n):
        ia = a % ns
        for q in range(2):
            for jb in range(max(0, ia - half_band), min(ns, ia + half_band + 1)):
                Ai.append(q * ns + jb)
        Ap.append(len(Ai))
    return np.array(Ap, np.int32), np.array(Ai, np.int32)


def dkb_parts(nsplines, rbox, nq=12):
    """Quadrature points x, weights w and the kappa-polynomial parts of the radial basis functions at x:
    P(kappa) = P0 + kappa P1 and Q(kappa) = Q0 + kappa Q1 (columns: the 2 (nsplines - 1) functions), with
    their r-derivatives, as the reference's dual kinetically balanced basis (src/dkbbasis.cpp: dkbCache, with P
    the ul = 1 and Q the ul = -1 part):
      large functions: P = B,                    Q = (B' + kappa B/r)/(2c)
      small functions: P = (-B' + kappa B/r)/(2c), Q = B
    Returns dict(x, w, P0, P1, Q0, Q1, dP0, dP1, dQ0, dQ1)."""
    t = np.array([rbox / nsplines * i for i in range(nsplines + 1)])
    x, w = _gauss_legendre_points(t, nq)
    B, dB, ddB = _bspline_cache(t, 7, x)
    c = SOL
    inv_r = 1.0 / x
    z0 = np.zeros_like(B)
    Br = B * inv_r[:, None]
    dBr = dB * inv_r[:, None] - B * (inv_r ** 2)[:, None]
    return dict(x=x, w=w,
                P0=np.hstack([B, -dB / (2 * c)]), P1=np.hstack([z0, Br / (2 * c)]),
                Q0=np.hstack([dB / (2 * c), B]), Q1=np.hstack([Br / (2 * c), z0]),
                dP0=np.hstack([dB, -ddB / (2 * c)]), dP1=np.hstack([z0, dBr / (2 * c)]),
                dQ0=np.hstack([ddB / (2 * c), dB]), dQ1=np.hstack([dBr / (2 * c), z0]))


def bessel_kernel(alpha, l, omega, cycles, x):
    """C_alpha(l)(r) = j_l(k_alpha r), one wave number per pair alpha = 2j, 2j+1 (src/potential.h:470-497):
    k = 2 pi/(c T) + omega/c, -(2 pi/(c T) - omega/c), omega/c; j_l(-z) = (-1)^l j_l(z)."""
    from scipy.special import spherical_jn
    c = SOL
    T = cycles * 2 * math.pi / omega
    k = [2 * math.pi / (c * T) + omega / c, -(2 * math.pi / (c * T) - omega / c), omega / c][alpha // 2]
    return spherical_jn(l, abs(k) * x) * (np.sign(k) ** l if k != 0 else 1.0)


def radial_matrices(nsplines, rbox, Z, omega, cycles, nl, nq=12):
    """Overlap s0..s2, field-free h00..h03 and interaction g[n][alpha][l] (dense)."""
    d = dkb_parts(nsplines, rbox, nq)
    x, w = d["x"], d["w"]
    P0, P1, Q0, Q1 = d["P0"], d["P1"], d["Q0"], d["Q1"]
    dP0, dP1, dQ0, dQ1 = d["dP0"], d["dP1"], d["dQ0"], d["dQ1"]
    c = SOL
    inv_r = 1.0 / x
    # (H below is [[V, c(-d/dr + k/r)], [c(d/dr + k/r), V - 2c^2]])
    V = -Z * inv_r

    def ip(f, g, weight=None):
        ww = w if weight is None else w * weight
        return f.T @ (ww[:, None] * g)

    s = [ip(P0, P0) + ip(Q0, Q0),
         ip(P0, P1) + ip(P1, P0) + ip(Q0, Q1) + ip(Q1, Q0),
         ip(P1, P1) + ip(Q1, Q1)]
    # H phi = (V P - c Q' + kappa c Q / r,  c P' + kappa c P / r + (V - 2c^2) Q)
    A = [V[:, None] * P0 - c * dQ0,
         V[:, None] * P1 - c * dQ1 + c * Q0 * inv_r[:, None],
         c * Q1 * inv_r[:, None]]
    C = [c * dP0 + (V - 2 * c * c)[:, None] * Q0,
         c * dP1 + c * P0 * inv_r[:, None] + (V - 2 * c * c)[:, None] * Q1,
         c * P1 * inv_r[:, None]]
    Pp, Qp = [P0, P1], [Q0, Q1]
    h = [np.zeros((P0.shape[1],) * 2) for _ in range(4)]
    for p in range(2):
        for q in range(3):
            h[p + q] += ip(Pp[p], A[q]) + ip(Qp[p], C[q])
    s = [0.5 * (m + m.T) for m in s]
    h = [0.5 * (m + m.T) for m in h]
    # interaction integrals g^n_{alpha,l} = P^T W C_alpha(l) Q over the four kappa parts (src/dkbbasis.cpp:
    # 780-820), with the Bessel kernels of bessel_kernel: <P(k_r)| C |Q(k_c)> = g0 + k_r g1 + k_c g2 + k_r k_c g3
    g = [[[None] * nl for _ in range(6)] for _ in range(4)]
    for alpha in range(6):
        for l in range(nl):
            jl = bessel_kernel(alpha, l, omega, cycles, x)
            g[0][alpha][l] = ip(P0, Q0, jl)
            g[1][alpha][l] = ip(P1, Q0, jl)
            g[2][alpha][l] = ip(P0, Q1, jl)
            g[3][alpha][l] = ip(P1, Q1, jl)
    return s, h, g


# --------------------------------------------------------------------------
# file formats
# --------------------------------------------------------------------------
def csr_write(fname, Ap, Ai, Ax, perm=None, part=None):
    """.csr layout read by tools/csr.c:262-318 (int32 dim, nnz, Ap, Ai, complex128 Ax, [perm, npart, part])."""
    Ap = np.asarray(Ap, np.int32)
    with open(fname, "wb") as f:
        np.array([len(Ap) - 1, Ap[-1]], np.int32).tofile(f)
        Ap.tofile(f)
        np.asarray(Ai, np.int32).tofile(f)
        np.asarray(Ax, np.complex128).tofile(f)
        if perm is not None:
            np.asarray(perm, np.int32).tofile(f)
            np.asarray(part, np.int32)
            np.array([len(part) - 1], np.int32).tofile(f)
            np.asarray(part, np.int32).tofile(f)


def csr_read(fname):
    """Returns dict(Ap, Ai, Ax, n, perm, part)."""
    raw = open(fname, "rb").read()
    n, nnz = np.frombuffer(raw, np.int32, 2)
    off = 8
    Ap = np.frombuffer(raw, np.int32, n + 1, off).copy(); off += 4 * (n + 1)
    if Ap[n] != nnz:
        raise ValueError(f"wrong file format (nnz) in {fname}")
    Ai = np.frombuffer(raw, np.int32, nnz, off).copy(); off += 4 * nnz
    Ax = np.frombuffer(raw, np.complex128, nnz, off).copy(); off += 16 * nnz
    perm = part = None
    if len(raw) - off >= 4 * n:
        perm = np.frombuffer(raw, np.int32, n, off).copy(); off += 4 * n
        npart = int(np.frombuffer(raw, np.int32, 1, off)[0]); off += 4
        part = np.frombuffer(raw, np.int32, npart + 1, off).copy()
    return dict(Ap=Ap, Ai=Ai, Ax=Ax, n=int(n), perm=perm, part=part)


def _dense_to_pattern(M, Ap, Ai):
    rows = np.repeat(np.arange(len(Ap) - 1), np.diff(Ap))
    return M[rows, Ai].astype(np.complex128)


# --------------------------------------------------------------------------
# ground state and tools/main.py preprocessing
# --------------------------------------------------------------------------
def ground_state(s, h, Z):
    """Eigenvector of (H0(kappa=-1), S(kappa=-1)) nearest -Z^2/2 (src/dumpBdpCouplings.cpp:307-322)."""
    from scipy.linalg import eigh
    kap = -1
    S = s[0] + kap * s[1] + kap * kap * s[2]
    H = h[0] + kap * h[1] + kap ** 2 * h[2] + kap ** 3 * h[3]
    ev, vec = eigh(H, S)
    idx = int(np.argmin(np.abs(ev + Z * Z / 2.0)))
    v = vec[:, idx]
    v = v / math.sqrt(float(v @ S @ v))
    if v[np.argmax(np.abs(v))] < 0:
        v = -v
    return float(ev[idx]), v


def preprocess(outdir, lmax):
    """tools/main.py:52-171 with one partition: H.csr (Hall + nzc perm) and psi0.vec."""
    from scipy import sparse

    def rd(fname):
        d = csr_read(os.path.join(outdir, fname))
        return sparse.csr_matrix((d["Ax"], d["Ai"], d["Ap"]), shape=(d["n"], d["n"]))

    res = None
    for l in range(lmax):
        m = rd(f"H0l{l}.csr")
        res = m if res is None else res + m
        res = res + rd(f"H1l{l}.csr")
    res_orig = res
    nzc = sum(res).nonzero()[1]
    res = res[nzc[:, None], nzc]
    res = res + sparse.identity(res.shape[0], format="csr")
    res_orig = res_orig + sparse.identity(res_orig.shape[0], format="csr")
    part_Ap = np.array([0, len(nzc)], np.int32)
    perm = np.arange(0, res.shape[0])
    g0 = rd("g0a0l0.csr")
    blkdim = g0.shape[1]
    res = res_orig[nzc[:, None], nzc]
    raw = open(os.path.join(outdir, "psi0"), "rb").read()
    rows, cols = np.frombuffer(raw, np.int64, 2)
    psi0 = np.frombuffer(raw, np.complex128, rows * cols, 16)
    perm_full = np.tile(perm.reshape((-1, 1)), (1, blkdim))
    dof_blk = np.indices((1, blkdim))[1]
    psi0_perm = (perm_full * blkdim + dof_blk).reshape((1, -1)).astype(np.int32)
    psi0 = psi0[psi0_perm]
    with open(os.path.join(outdir, "psi0.vec"), "wb") as f:
        np.array(psi0_perm.shape[1], np.int32).tofile(f)
        psi0.tofile(f)
        psi0_perm.tofile(f)
    res.sort_indices()
    csr_write(os.path.join(outdir, "H.csr"), res.indptr, res.indices, res.data,
              perm=nzc.astype(np.int32), part=part_Ap)


def generate(outdir, nsplines=50, nkappa=4, nmu=0, nl=1, rbox=30.0, Z=1,
             intensity=10, omega=50.0, cycles=1, verbose=False):
    """Write the complete input file set for one configuration into ``outdir``."""
    os.makedirs(outdir, exist_ok=True)
    for l in range(nl):
        for ul, name in ((1, "H0l"), (-1, "H1l")):
            Ap, Ai, Ax, n = angular_matrix(nkappa, nmu, ul, l)
            csr_write(os.path.join(outdir, f"{name}{l}.csr"), Ap, Ai, Ax)
    s, h, g = radial_matrices(nsplines, rbox, Z, omega, cycles, nl)
    ns = nsplines + 5
    Ap, Ai = radial_pattern(ns)
    for n in range(4):
        for alpha in range(6):
            for l in range(nl):
                csr_write(os.path.join(outdir, f"g{n}a{alpha}l{l}.csr"), Ap, Ai,
                          _dense_to_pattern(g[n][alpha][l], Ap, Ai))
    for n in range(4):
        csr_write(os.path.join(outdir, f"h0{n}.csr"), Ap, Ai, _dense_to_pattern(h[n], Ap, Ai))
    for n in range(3):
        csr_write(os.path.join(outdir, f"s{n}.csr"), Ap, Ai, _dense_to_pattern(s[n], Ap, Ai))
    e0, v = ground_state(s, h, Z)
    nang = n_angular(nkappa)
    psi = np.zeros(nang * 2 * ns, np.complex128)
    psi[: 2 * ns] = v  # angular index 0 is kappa=-1, mu=-1/2
    with open(os.path.join(outdir, "psi0"), "wb") as f:
        np.array([len(psi), 1], np.int64).tofile(f)
        psi.tofile(f)
    preprocess(outdir, nl)
    params = dict(Nsplines=nsplines, Nkappa=nkappa, Nmu=nmu, Nl=nl, rBox=rbox, Z=Z,
                  Intensity=intensity, Omega=omega, Cycles=cycles, e0=e0)
    with open(os.path.join(outdir, "params.json"), "w") as f:
        json.dump(params, f, indent=1)
    if verbose:
        print(f"generated {outdir}: e0={e0:.10f}")
    return params


CONFIGS = {
    # dipole, NSplines 50, NKappa 4, Nl 1, 1 cycle
    "dipole_ns50_k4_l1": dict(nsplines=50, nkappa=4, nmu=0, nl=1, rbox=30.0, Z=1,
                              intensity=10, omega=50.0, cycles=1),
    # nondipole production size, 15 cycles
    "nondipole_ns250_k16_mu8_l8": dict(nsplines=250, nkappa=16, nmu=8, nl=8, rbox=30.0, Z=1,
                                       intensity=10, omega=50.0, cycles=15),
}


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("outdir")
    ap.add_argument("--config", choices=sorted(CONFIGS), default="dipole_ns50_k4_l1")
    a = ap.parse_args()
    generate(a.outdir, verbose=True, **CONFIGS[a.config])
