"""Pins of the oracle's lab-frame moments (ComputeMomentsLF,
src/Radiation.cpp:270-301) by closed forms and invariants, not by retyping
the transform:

* identity tetrad and metric: LF = IF and F = |F|;
* a rotation tetrad: E is unchanged, F and P rotate as a vector and a tensor,
  |F| is invariant;
* a boost-like tetrad followed by its inverse returns the IF tensor
  (TransformLFtoIF of TransformIFtoLF, src/SpecialMath.cpp:85-105);
* F = sqrt(|gamma_ij F^i F^j|) for a diagonal metric;
* cells inside the excision radius are 0.
"""
import numpy as np


def _grid(O, n=6):
    return O.make_grid(n, n, (-1, -1), (1, 1), 0.5)


def _moments(g, seed=3):
    r = np.random.default_rng(seed)
    m = r.standard_normal((6, g.N))
    m[0] = np.abs(m[0]) + 2.0
    return m


def _tetrad_field(g, e):
    return np.repeat(np.asarray(e, float).reshape(9, 1), g.N, axis=1)


def test_identity_frame(O):
    g = _grid(O)
    mom = _moments(g)
    lf = O.moments_lf(g, mom, _tetrad_field(g, np.eye(3)), np.stack([np.ones(g.N), np.zeros(g.N), np.ones(g.N)]))
    assert np.array_equal(lf[:6], mom)
    assert np.allclose(lf[6], np.hypot(mom[1], mom[2]), rtol=1e-15, atol=0)


def test_rotation_frame(O):
    g = _grid(O)
    mom = _moments(g, 5)
    th = 0.7
    c, s = np.cos(th), np.sin(th)
    e = [[1, 0, 0], [0, c, -s], [0, s, c]]
    lf = O.moments_lf(g, mom, _tetrad_field(g, e), np.stack([np.ones(g.N), np.zeros(g.N), np.ones(g.N)]))
    E, Fx, Fy, Pxx, Pxy, Pyy = mom
    assert np.allclose(lf[0], E, rtol=1e-14)
    assert np.allclose(lf[1], c * Fx - s * Fy, atol=1e-13)
    assert np.allclose(lf[2], s * Fx + c * Fy, atol=1e-13)
    # P' = R P R^T
    assert np.allclose(lf[3], c * c * Pxx - 2 * c * s * Pxy + s * s * Pyy, atol=1e-13)
    assert np.allclose(lf[4], c * s * (Pxx - Pyy) + (c * c - s * s) * Pxy, atol=1e-13)
    assert np.allclose(lf[5], s * s * Pxx + 2 * c * s * Pxy + c * c * Pyy, atol=1e-13)
    assert np.allclose(lf[6], np.hypot(Fx, Fy), rtol=1e-13)


def test_boost_round_trip(O):
    g = _grid(O)
    mom = _moments(g, 7)
    b, gam = 0.4, 1.0 / np.sqrt(1 - 0.4 ** 2)
    e = np.array([[gam, gam * b, 0.0], [gam * b, gam, 0.0], [0.0, 0.0, 1.0]])
    ident = np.stack([np.ones(g.N), np.zeros(g.N), np.ones(g.N)])
    lf = O.moments_lf(g, mom, _tetrad_field(g, e), ident)
    # transform back with the inverse tetrad: the tensor comes back
    back_in = lf[:6]
    back = O.moments_lf(g, back_in, _tetrad_field(g, np.linalg.inv(e)), ident)
    assert np.allclose(back[:6], mom, rtol=1e-13, atol=1e-13)
    # the boost mixes E and Fx: E' = gam^2 (E + 2 b Fx + b^2 Pxx)
    E, Fx, Pxx = mom[0], mom[1], mom[3]
    assert np.allclose(lf[0], gam * gam * (E + 2 * b * Fx + b * b * Pxx), rtol=1e-13)


def test_metric_norm_and_excision(O):
    g = _grid(O, 10)
    mom = _moments(g, 9)
    gamma = np.stack([np.full(g.N, 4.0), np.zeros(g.N), np.full(g.N, 9.0)])
    r2 = 0.3 ** 2
    lf = O.moments_lf(g, mom, _tetrad_field(g, np.eye(3)), gamma, r2)
    x = g.startx + (np.arange(g.N) % g.nx) * g.dx
    y = g.starty + (np.arange(g.N) // g.nx) * g.dy
    inside = x * x + y * y <= r2
    assert inside.any() and (~inside).any()
    assert np.all(lf[:, inside] == 0)
    want = np.sqrt(4 * mom[1] ** 2 + 9 * mom[2] ** 2)
    assert np.allclose(lf[6, ~inside], want[~inside], rtol=1e-14)
