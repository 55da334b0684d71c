"""CPU statement of the learned initial guess's network, loss and gradients (numpy).

TEST INFRASTRUCTURE ONLY: the checker of the device MLP kernels (forward, loss,
backward), used by tests/.  The product path never imports it.

The learned guess has no counterpart in the reference; this file defines its
arithmetic independently of the HIP code:
  * inputs: standardized raw features, rounded to bf16, zero beyond the used ones;
  * 3 hidden layers Z = A W^T + b, A = bf16(GELU(Z)) (exact erf GELU), output
    Y = A W4^T + b4, with bf16 weights and products accumulated in float
    (``quantized=True``), or everything in float64 (``quantized=False``, the form the
    finite-difference check pins);
  * loss per sample: L = (beta - 2 w.a + w^T G w - 2 Re(c^H q) + |c|^2 + 2 Re(c^H P w)) / beta,
    with w = Y[:6] and c_j = Y[6+2j] + i Y[7+2j]; the batch loss is the mean.
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT, POD, SAMPLE = 288, 256, 144, 64, 1232
F_END = 17 + 4 * POD  # used inputs
S_G, S_A = IN, IN + 36
S_P = S_A + 6
S_Q = S_P + POD * 12
S_BETA = S_Q + POD * 2
S_VALID = S_BETA + 1
LAYERS = [(IN, HID), (HID, HID), (HID, HID), (HID, OUT)]


def offsets():
    """(W offset, b offset) of each layer in the flat parameter vector."""
    out, o = [], 0
    for fi, fo in LAYERS:
        out.append((o, o + fo * fi))
        o += fo * fi + fo
    return out, o


OFFS, NPARAM = offsets()


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(x.shape)


def _erf(x):
    from scipy.special import erf
    return erf(x)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + _erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def layer_params(p, k):
    wo, bo = OFFS[k]
    fi, fo = LAYERS[k]
    return p[wo:wo + fo * fi].reshape(fo, fi), p[bo:bo + fo]


def standardize(x, mean, std):
    X = (np.asarray(x, np.float64) - mean) / std
    X[:, F_END:] = 0.0
    return X


def forward(p, X, quantized=True):
    """X: standardized inputs [M][IN].  Returns Y [M][OUT] and the cache for backward."""
    dt = np.float32 if quantized else np.float64
    q = bf16 if quantized else (lambda v: np.asarray(v, np.float64))
    A = q(np.asarray(X, dt))
    cache = {"A": [A], "Z": []}
    for k in range(4):
        W, b = layer_params(p, k)
        Wq = q(np.asarray(W, dt))
        Z = (A.astype(np.float64) @ Wq.astype(np.float64).T + b).astype(dt)
        cache["Z"].append(Z)
        if k < 3:
            A = q(gelu(Z.astype(np.float64)).astype(dt))
            cache["A"].append(A)
        else:
            Y = Z
    cache["W"] = [q(np.asarray(layer_params(p, k)[0], dt)) for k in range(4)]
    return Y, cache


def unpack(s):
    """Quadratic-form pieces of samples s [M][SAMPLE]."""
    s = np.asarray(s, np.float64)
    G = s[:, S_G:S_A].reshape(-1, 6, 6)
    a = s[:, S_A:S_P]
    Pr = s[:, S_P:S_Q].reshape(-1, POD, 6, 2)
    P = Pr[..., 0] + 1j * Pr[..., 1]
    qr = s[:, S_Q:S_BETA].reshape(-1, POD, 2)
    qv = qr[..., 0] + 1j * qr[..., 1]
    beta = s[:, S_BETA]
    return G, a, P, qv, beta


def loss_and_dY(Y, s):
    """Mean loss over the rows and dLoss/dY (float64)."""
    Y = np.asarray(Y, np.float64)
    M = Y.shape[0]
    G, a, P, qv, beta = unpack(s)
    w = Y[:, :6]
    c = Y[:, 6:6 + 2 * POD:2] + 1j * Y[:, 7:7 + 2 * POD:2]
    Pw = np.einsum("mjk,mk->mj", P, w)
    L = (beta - 2 * np.sum(w * a, 1) + np.einsum("mk,mkl,ml->m", w, G, w) - 2 * np.real(np.sum(np.conj(c) * qv, 1))
         + np.sum(np.abs(c) ** 2, 1) + 2 * np.real(np.sum(np.conj(c) * Pw, 1))) / beta
    dY = np.zeros_like(Y)
    dw = 2 * (-a + np.einsum("mkl,ml->mk", G, w) + np.real(np.einsum("mjk,mj->mk", np.conj(P), c))) / beta[:, None]
    dc = 2 * (c - qv + Pw) / beta[:, None]
    dY[:, :6] = dw
    dY[:, 6:6 + 2 * POD:2] = dc.real
    dY[:, 7:7 + 2 * POD:2] = dc.imag
    return float(L.mean()), dY / M


def backward(p, cache, dY, quantized=True):
    """Gradient of the loss w.r.t. the flat parameters, given dLoss/dY."""
    q = bf16 if quantized else (lambda v: np.asarray(v, np.float64))
    dt = np.float32 if quantized else np.float64
    g = np.zeros(NPARAM, np.float64)
    dZ = np.asarray(dY, dt)
    for k in range(3, -1, -1):
        wo, bo = OFFS[k]
        fi, fo = LAYERS[k]
        Gq = q(dZ).astype(np.float64)
        Ain = cache["A"][k].astype(np.float64)
        g[wo:wo + fo * fi] = (Gq.T @ Ain).ravel()
        g[bo:bo + fo] = np.asarray(dZ, np.float64).sum(0)
        if k > 0:
            dA = (Gq @ cache["W"][k].astype(np.float64)).astype(dt)
            dZ = (dA.astype(np.float64) * gelu_d(cache["Z"][k - 1].astype(np.float64))).astype(dt)
    return g


def direct_loss(ehat, Dh, U, w, c):
    """The loss from full vectors: |ehat - Dh w - U c|^2 / |ehat|^2."""
    e = ehat - Dh @ w - U @ c
    return float(np.vdot(e, e).real / np.vdot(ehat, ehat).real)


def sample_from_vectors(feat, ehat, Dh, U):
    """A sample row from full vectors (Dh columns unit-norm, U orthonormal columns)."""
    s = np.zeros(SAMPLE, np.float64)
    s[:IN] = feat
    s[S_G:S_A] = np.real(Dh.conj().T @ Dh).ravel()
    s[S_A:S_P] = np.real(Dh.conj().T @ ehat)
    P = U.conj().T @ Dh
    s[S_P:S_Q] = np.stack([P.real, P.imag], -1).ravel()
    qv = U.conj().T @ ehat
    s[S_Q:S_BETA] = np.stack([qv.real, qv.imag], -1).ravel()
    s[S_BETA] = np.vdot(ehat, ehat).real
    s[S_VALID] = 1.0
    return s
