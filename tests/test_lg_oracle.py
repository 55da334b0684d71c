"""CPU pins of the learned guess's oracle (oracle/lg_oracle.py): bf16 rounding, the quadratic-form loss
against the loss computed from full vectors, and the backpropagated gradient against central finite
differences (float64 form)."""
import numpy as np
import pytest

import lg_oracle as LG


def test_bf16_rounding_known_values():
    x = np.array([1.0, -2.5, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7, 3.0e38, 1e-40], np.float32)
    r = LG.bf16(x)
    # ties to even: 1 + 2^-8 -> 1, 1 + 3*2^-8 -> 1 + 2^-6; exactly representable values stay
    assert r[0] == 1.0 and r[1] == -2.5 and r[2] == 1.0 and r[3] == np.float32(1.0 + 2 ** -6)
    assert r[4] == np.float32(1.0 + 2 ** -7)
    assert np.all(np.isfinite(r[:6]))
    # 8 significant bits kept: the low 16 bits of the float32 pattern are zero
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)


def random_problem(rng, n=300, M=4):
    U, _ = np.linalg.qr(rng.standard_normal((n, LG.POD)) + 1j * rng.standard_normal((n, LG.POD)))
    rows, vecs = [], []
    for _ in range(M):
        Dh = rng.standard_normal((n, 6)) + 1j * rng.standard_normal((n, 6))
        Dh /= np.linalg.norm(Dh, axis=0)
        ehat = 0.3 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
        feat = rng.standard_normal(LG.IN)
        rows.append(LG.sample_from_vectors(feat, ehat, Dh, U))
        vecs.append((ehat, Dh))
    return U, np.array(rows), vecs


def test_quadratic_form_equals_direct_loss():
    rng = np.random.default_rng(5)
    U, S, vecs = random_problem(rng)
    Y = rng.standard_normal((len(S), LG.OUT)) * 0.1
    L, _ = LG.loss_and_dY(Y, S)
    direct = []
    for m, (ehat, Dh) in enumerate(vecs):
        w = Y[m, :6]
        c = Y[m, 6:6 + 2 * LG.POD:2] + 1j * Y[m, 7:7 + 2 * LG.POD:2]
        direct.append(LG.direct_loss(ehat, Dh, U, w, c))
    assert abs(L - np.mean(direct)) <= 1e-12 * max(1.0, abs(L))
    # zero outputs reproduce the extrapolation: loss 1
    L0, _ = LG.loss_and_dY(np.zeros_like(Y), S)
    assert abs(L0 - 1.0) <= 1e-12


def test_dY_matches_finite_differences():
    rng = np.random.default_rng(6)
    U, S, _ = random_problem(rng, M=3)
    Y = rng.standard_normal((3, LG.OUT)) * 0.1
    _, dY = LG.loss_and_dY(Y, S)
    h = 1e-6
    for (m, j) in [(0, 0), (1, 5), (2, 6), (0, 7), (1, 100), (2, 133), (0, 140)]:
        Yp, Ym = Y.copy(), Y.copy()
        Yp[m, j] += h
        Ym[m, j] -= h
        fd = (LG.loss_and_dY(Yp, S)[0] - LG.loss_and_dY(Ym, S)[0]) / (2 * h)
        assert abs(fd - dY[m, j]) <= 1e-7 * max(1.0, abs(fd)), (m, j, fd, dY[m, j])


def test_backprop_matches_finite_differences():
    """float64 network: every layer's weight and bias gradient against central differences."""
    rng = np.random.default_rng(7)
    U, S, _ = random_problem(rng, M=4)
    p = np.zeros(LG.NPARAM)
    for k, (fi, fo) in enumerate(LG.LAYERS):
        wo, bo = LG.OFFS[k]
        p[wo:wo + fo * fi] = rng.standard_normal(fo * fi) * np.sqrt(2.0 / fi) * (0.3 if k == 3 else 1.0)
        p[bo:bo + fo] = 0.1 * rng.standard_normal(fo)
    mean, std = np.zeros(LG.IN), np.ones(LG.IN)
    X = LG.standardize(S[:, :LG.IN], mean, std)

    def loss(pp):
        Y, _ = LG.forward(pp, X, quantized=False)
        return LG.loss_and_dY(Y, S)[0]

    Y, cache = LG.forward(p, X, quantized=False)
    _, dY = LG.loss_and_dY(Y, S)
    g = LG.backward(p, cache, dY, quantized=False)
    h = 1e-6
    picks = []
    for k, (fi, fo) in enumerate(LG.LAYERS):
        wo, bo = LG.OFFS[k]
        picks += [wo + int(rng.integers(fo * fi)) for _ in range(4)] + [bo + int(rng.integers(fo)) for _ in range(2)]
    for i in picks:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (loss(pp) - loss(pm)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(fd)) + 1e-9, (i, fd, g[i])
