"""Pins of the FP64 radiance-cache oracle (oracle/nrc_oracle.cpp), CPU only.

* gradients == central finite differences of the loss with the denominators
  frozen at the evaluation point (the loss the library differentiates);
* corner indices == the published hash / dense-index formula recomputed here
  in Python integers;
* trilinear weights sum to 1 and reproduce the sample's grid coordinate;
* the forward output is linear in the last layer and zero for a zero input
  layer (structure of the network).
"""
import numpy as np
import pytest

import oracle

LOG2 = 8
RES = [2, 3, 3, 4, 5, 6, 8, 10, 12, 15, 19, 24, 30, 38, 48, 60]
LAYOUT = {"resolution": RES, "log2_table": LOG2, "origin": [-1.0, -0.5, 0.25], "inv_extent": 0.25}
OFF = [0, 8192, 12288, 16384, 20480]


def _net(seed=0, grid_scale=1.0):
    rng = np.random.default_rng(seed)
    mlp = np.zeros(21504)
    for L, (o, i) in enumerate([(64, 128), (64, 64), (64, 64), (64, 64), (3, 64)]):
        lim = np.sqrt(6.0 / (o + i))
        mlp[OFF[L]:OFF[L] + o * i] = rng.uniform(-lim, lim, o * i)
    grid = rng.uniform(-grid_scale, grid_scale, (16 << LOG2) * 2)
    return mlp, grid


def _inputs(n, seed=1):
    rng = np.random.default_rng(seed)
    pos = rng.uniform([-1, -0.5, 0.25], [3, 3.5, 4.25], (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    nr = rng.normal(size=(n, 3))
    nr /= np.linalg.norm(nr, axis=1, keepdims=True)
    rough = rng.uniform(0, 1, n).astype(np.float32)
    target = rng.uniform(0.0, 3.0, (n, 3)).astype(np.float32)
    return pos, d.astype(np.float32), nr.astype(np.float32), rough, target


def _frozen_loss(oc, inp, den):
    pos, d, nr, r, t = inp
    y = oc.forward(pos, d, nr, r)
    return float(np.sum((y - t) ** 2 / den) / (3 * len(pos)))


def test_gradients_match_finite_differences():
    mlp, grid = _net(0)
    inp = _inputs(3)
    oc = oracle.OracleCache(LAYOUT, mlp, grid)
    loss, gw, gg = oc.gradients(*inp)
    y0 = oc.forward(*inp[:4])
    den = y0 ** 2 + 0.01
    assert abs(loss - _frozen_loss(oc, inp, den)) <= 1e-12 * max(1.0, abs(loss))
    rng = np.random.default_rng(5)
    # MLP weights: a spread over every layer, including the grid-feature and bias columns of layer 0
    picks = list(rng.integers(0, 8192, 12)) + [OFF[0] + 7 * 128 + 3, OFF[0] + 11 * 128 + 104] + \
        list(rng.integers(8192, 20480, 12)) + [OFF[4] + k for k in (0, 64 + 5, 128 + 63)]
    for i in picks:
        h = 1e-4 * max(1.0, abs(mlp[i]))
        m2 = mlp.copy()
        m2[i] += h
        lp = _frozen_loss(oracle.OracleCache(LAYOUT, m2, grid), inp, den)
        m2[i] -= 2 * h
        lm = _frozen_loss(oracle.OracleCache(LAYOUT, m2, grid), inp, den)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gw[i]) <= 1e-6 * max(abs(gw[i]), 1e-4), (i, fd, gw[i])
    # grid entries the samples touch
    touched = np.nonzero(gg)[0]
    assert len(touched) > 50
    for i in rng.choice(touched, 25, replace=False):
        h = 1e-4
        g2 = grid.copy()
        g2[i] += h
        lp = _frozen_loss(oracle.OracleCache(LAYOUT, mlp, g2), inp, den)
        g2[i] -= 2 * h
        lm = _frozen_loss(oracle.OracleCache(LAYOUT, mlp, g2), inp, den)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gg[i]) <= 1e-6 * max(abs(gg[i]), 1e-4), (i, fd, gg[i])
    # rows 3..15 of the last layer receive no gradient
    assert not np.any(gw[OFF[4] + 3 * 64:])


def _expected_corners(p, level):
    """Published formula in Python integers (independent of the C++ code)."""
    origin = np.float32(LAYOUT["origin"])
    inv = np.float32(LAYOUT["inv_extent"])
    r = RES[level]
    T = 1 << LOG2
    pn = np.clip((np.float32(p) - origin) * inv, np.float32(0), np.float32(1)).astype(np.float32)
    g = (pn * np.float32(r)).astype(np.float32)
    c = np.minimum(np.floor(g).astype(np.int64), r - 1)
    f = (g - c.astype(np.float32)).astype(np.float64)
    idx, w = [], []
    for k in range(8):
        x, y, z = int(c[0]) + (k & 1), int(c[1]) + ((k >> 1) & 1), int(c[2]) + ((k >> 2) & 1)
        if (r + 1) ** 3 <= T:
            i = x + (r + 1) * (y + (r + 1) * z)
        else:
            i = (x ^ ((y * 2654435761) & 0xFFFFFFFF) ^ ((z * 805459861) & 0xFFFFFFFF)) & (T - 1)
        idx.append((level << LOG2) + i)
        wx = f[0] if k & 1 else 1 - f[0]
        wy = f[1] if k & 2 else 1 - f[1]
        wz = f[2] if k & 4 else 1 - f[2]
        w.append(wx * wy * wz)
    return np.array(idx), np.array(w), g


def test_corner_indices_and_trilinear_weights():
    mlp, grid = _net(0)
    oc = oracle.OracleCache(LAYOUT, mlp, grid)
    pos = _inputs(40, seed=3)[0]
    pos[0] = [-1.0, -0.5, 0.25]  # box corner: cell 0, frac 0
    pos[1] = [3.0, 3.5, 4.25]    # far corner: clamped to the last cell, frac 1
    pos[2] = [-5.0, 9.0, 1.0]    # outside: clamped
    dense_seen = hashed_seen = False
    for p in pos:
        for level in range(16):
            idx, w = oc.corners(p, level)
            eidx, ew, g = _expected_corners(p, level)
            assert np.array_equal(idx, eidx), (p, level)
            assert np.allclose(w, ew, rtol=0, atol=1e-15)
            assert abs(w.sum() - 1.0) < 1e-12
            r = RES[level]
            if (r + 1) ** 3 <= (1 << LOG2):
                dense_seen = True
                loc = idx - (level << LOG2)
                xyz = np.stack([loc % (r + 1), (loc // (r + 1)) % (r + 1), loc // (r + 1) ** 2], 1)
                # trilinear weights reproduce the lattice coordinate of the sample
                assert np.allclose(w @ xyz, np.minimum(g.astype(np.float64), r), atol=1e-5)
            else:
                hashed_seen = True
                assert np.all(idx - (level << LOG2) < (1 << LOG2))
    assert dense_seen and hashed_seen


def test_forward_structure():
    mlp, grid = _net(2)
    inp = _inputs(16, seed=4)
    y = oracle.OracleCache(LAYOUT, mlp, grid).forward(*inp[:4])
    m2 = mlp.copy()
    m2[OFF[4]:] *= 2.5
    y2 = oracle.OracleCache(LAYOUT, m2, grid).forward(*inp[:4])
    assert np.allclose(y2, 2.5 * y, rtol=1e-12, atol=1e-15)
    m3 = mlp.copy()
    m3[OFF[0]:OFF[1]] = 0.0
    assert not np.any(oracle.OracleCache(LAYOUT, m3, grid).forward(*inp[:4]))


def test_one_blob_peaks_at_the_encoded_bin():
    """A network whose first layer picks one input feature: with roughness v
    the 8-bin one-blob is largest at bin floor(v*8) and equals exp(-d^2/2)."""
    mlp = np.zeros(21504)
    grid = np.zeros((16 << LOG2) * 2)
    # h1[0] = feature k; h2..h4 pass unit 0 through; y0 = h4[0]
    for L in (1, 2, 3):
        mlp[OFF[L] + 0] = 1.0
    mlp[OFF[4] + 0] = 1.0
    inp = _inputs(1, seed=9)
    for v in (0.03, 0.31, 0.53, 0.77, 0.99):
        vals = []
        for b in range(8):
            m = mlp.copy()
            m[OFF[0] + 96 + b] = 1.0
            y = oracle.OracleCache(LAYOUT, m, grid).forward(inp[0], inp[1], inp[2], np.float32([v]))[0, 0]
            vals.append(y)
            d = v * 8 - (b + 0.5)
            assert abs(y - np.exp(-0.5 * d * d)) < 1e-6
        assert int(np.argmax(vals)) == min(int(v * 8), 7)
