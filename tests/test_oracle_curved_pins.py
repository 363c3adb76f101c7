"""Pins of the oracle's curved streams and of the moving-fluid pieces of its
collide by relations the mathematics fixes (CPU only).

Curved streaming (reference src/Radiation.cpp:454-512, and the adaptive
form :513-575) gathers (alpha/alpha_k)^3 IntensityAt(k) at the four corners
of the departure point, clamps each IntensityAt at 0, interpolates
bilinearly and multiplies by s^3.  Each relation below fails when one of
those factors is dropped, inverted or moved:

* a state alpha_k^3 f streams to alpha^3 (the stream of f with alpha = 1);
  swapping alpha and alpha_k would give alpha_k^6 / alpha^3 instead;
* a frequency-shift series scaled by c scales the output by exactly c^3;
* the clamp acts on each corner before the bilinear interpolation: with
  one corner's angular interpolant negative and the other positive, the
  output is half the positive corner (clamping after the interpolation, or
  not at all, would subtract half the negative lobe).

Collide with fluid velocity u (src/Radiation.cpp:664-698): the fluid-frame
energy density is (L T L^T)_00 of the inertial-frame moment tensor T under
the boost L(u); the reference's fluid-frame flux is gamma (L T L^T)_0i (its
expression carries one factor gamma more than the boosted flux, and the
oracle keeps it); the fluid-frame direction of a unit vector is the
aberrated unit vector and A = gamma (1 - u.n).
"""
import ctypes as C
import math

import numpy as np
import pytest

SQ2 = math.sqrt(2.0)


def grid_and_tables(O, nx=14, ny=12):
    g = O.make_grid(nx, ny, (-1, -1), (1, 1), 0.5)
    M = O.metric(g, 0)
    tables = O.geodesic_tables(g, M, -1.0)  # straight lines, s = 1
    return g, M, tables


def coords(g):
    i = np.arange(g.N) % g.nx
    j = np.arange(g.N) // g.nx
    return i, j, g.startx + i * g.dx, g.starty + j * g.dy


def stream(P, I, adaptive=False):
    if adaptive:
        return P.stream_adaptive(I)
    out = np.zeros(P.grid.N * P.nDir)
    P.L.or_stream_curved(C.byref(P.sim), np.ascontiguousarray(I, dtype=np.float64), out)
    return out


def smooth_positive(g, phi, seed):
    rng = np.random.default_rng(seed)
    _, _, x, y = coords(g)
    a, b, c = rng.uniform(0.5, 1.5, 3)
    return (1.5 + 0.4 * np.sin(a * x[:, None] + b * y[:, None] + c * phi[None, :])
            + 0.2 * np.cos(2 * phi[None, :] - x[:, None])).reshape(-1)


def interior(g, D):
    i, j, _, _ = coords(g)
    m = (i > 0) & (i < g.nx - 1) & (j > 0) & (j < g.ny - 1)
    return np.repeat(m, D)


def test_fourier_basis_orders(O):
    # the coefficient layouts below rely on B_0 = 1, B_1 = sqrt2 cos, B_2 = sqrt2 sin
    L = O.lib()
    for p in (0.3, 2.9, 5.1):
        assert L.or_fourier_basis(0, p) == 1.0
        assert abs(L.or_fourier_basis(1, p) - SQ2 * math.cos(p)) < 1e-15
        assert abs(L.or_fourier_basis(2, p) - SQ2 * math.sin(p)) < 1e-15


@pytest.mark.parametrize("adaptive", [False, True])
def test_lapse_ratio_factor(O, adaptive):
    g, M, tables = grid_and_tables(O)
    D = 16
    kind = 3 if adaptive else 2
    # each Problem reads the lapse from its own copy of the metric block
    Pa = O.Problem(g, D, kind, M.copy(), tables=tables)
    Pb = O.Problem(g, D, kind, M.copy(), tables=tables)
    _, _, x, y = coords(g)
    alpha = 1.0 + 0.3 * np.sin(2 * x) * np.cos(3 * y)
    Pa.alpha[:] = alpha  # alpha_c / alpha_k differ between neighbours
    if adaptive:
        th = np.full(g.N, 0.7)
        for P in (Pa, Pb):
            P.set_rotation(th, th)
    f = smooth_positive(g, Pa.phi, 3)
    a3 = np.repeat(alpha ** 3, D)
    out_a = stream(Pa, a3 * f, adaptive)
    out_b = stream(Pb, f, adaptive)
    m = interior(g, D)
    err = np.abs(out_a[m] - a3[m] * out_b[m]) / np.abs(out_b[m]).max()
    assert err.max() <= 1e-14, err.max()
    # and the relation is not trivially true: the lapse really varies
    assert np.abs(out_a[m] - out_b[m]).max() > 1e-2


@pytest.mark.parametrize("adaptive", [False, True])
def test_frequency_shift_enters_cubed(O, adaptive):
    g, M, tables = grid_and_tables(O)
    D = 16
    c = 1.1
    kind = 3 if adaptive else 2
    P1 = O.Problem(g, D, kind, M, tables=tables)
    scaled = [t.copy() for t in tables]
    scaled[0] *= c  # the S series: s(phi) -> c s(phi)
    Pc = O.Problem(g, D, kind, M, tables=scaled)
    if adaptive:
        th = np.full(g.N, 2.1)
        for P in (P1, Pc):
            P.set_rotation(th, th)
    f = smooth_positive(g, P1.phi, 5)
    o1 = stream(P1, f, adaptive)
    oc = stream(Pc, f, adaptive)
    m = interior(g, D)
    assert np.abs(oc[m] - c ** 3 * o1[m]).max() <= 1e-14 * np.abs(oc[m]).max()


@pytest.mark.parametrize("adaptive", [False, True])
def test_clamp_is_per_corner_before_bilinear(O, adaptive):
    """Departure half a cell to the left (corners i-1 and i, weight 1/2 each)
    and departure velocities turned by half a stencil step, so the angular
    interpolation falls between directions, where a spike has negative
    cubic lobes.  Even columns hold a spike, odd columns a constant 1."""
    D = 16
    g = O.make_grid(14, 12, (-1, -1), (1, 1), 0.5)
    M = O.metric(g, 0)
    i, j, x, y = coords(g)
    N, nF = g.N, 5
    delta = 0.5 * 2 * math.pi / D
    cS, cX, cY, cCx, cCy = (np.zeros(N * nF) for _ in range(5))
    cS[0::nF] = 1.0
    cX[0::nF] = x - 0.5 * g.dx
    cY[0::nF] = y
    cCx[1::nF] = math.cos(delta) / SQ2
    cCx[2::nF] = -math.sin(delta) / SQ2
    cCy[1::nF] = math.sin(delta) / SQ2
    cCy[2::nF] = math.cos(delta) / SQ2
    P = O.Problem(g, D, 3 if adaptive else 2, M, tables=[cS, cX, cY, cCx, cCy])
    if adaptive:
        z = np.zeros(N)
        P.set_rotation(z, z)
    d0 = 5
    spike = np.zeros((N, D))
    spike[:, d0] = 1.0
    even = (i % 2 == 0)[:, None]
    base = np.where(even, spike, 1.0)
    out1 = stream(P, base.reshape(-1), adaptive).reshape(N, D)
    # the unclamped interpolant u of the spike, from the same stream with the
    # whole field raised by 1 (the weights sum to 1, nothing is clamped then)
    out2 = stream(P, (base + 1.0).reshape(-1), adaptive).reshape(N, D)
    cells = (i % 2 == 1) & (i > 0) & (i < g.nx - 1) & (j > 0) & (j < g.ny - 1)
    u = 2.0 * (out2[cells] - 1.5)
    assert u.min() < -0.02, "no negative lobe reached"
    want = 0.5 * np.maximum(0.0, u) + 0.5
    assert np.abs(out1[cells] - want).max() <= 1e-12
    # clamping after the interpolation would differ where the lobe is negative
    assert np.abs(out1[cells] - np.maximum(0.0, 0.5 * u + 0.5)).max() > 1e-2


def boost(u):
    g = 1.0 / math.sqrt(1.0 - u @ u)
    L = np.zeros((3, 3))
    L[0, 0] = g
    L[0, 1:] = L[1:, 0] = -g * u
    L[1:, 1:] = np.eye(2) + (g - 1.0) * np.outer(u, u) / (u @ u)
    return L, g


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_fluid_frame_moments_and_directions(O, seed):
    L = O.lib()
    rng = np.random.default_rng(seed)
    u = rng.uniform(-0.6, 0.6, 2)
    D = 32
    phi = 2 * math.pi * (np.arange(D) + 0.5) / D
    n = np.stack([np.cos(phi), np.sin(phi)], 1)
    w = np.full(D, 1.0 / D)
    I = rng.uniform(0.2, 1.0, D)
    E, F = (w * I).sum(), (w * I) @ n
    P2 = (n.T * (w * I)) @ n
    m = np.array([E, F[0], F[1], P2[0, 0], P2[0, 1], P2[1, 1]])
    out = np.zeros(3)
    L.or_fluid_frame(u[0], u[1], m, out)
    T = np.array([[E, F[0], F[1]], [F[0], P2[0, 0], P2[0, 1]], [F[1], P2[1, 0], P2[1, 1]]])
    Lb, g = boost(u)
    TF = Lb @ T @ Lb.T
    assert abs(out[0] - TF[0, 0]) <= 1e-14 * abs(TF[0, 0])
    assert np.abs(out[1:] - g * TF[0, 1:]).max() <= 1e-13 * np.abs(TF[0, 1:]).max()
    # not the boosted flux itself: the reference's expression has gamma more
    assert np.abs(out[1:] - TF[0, 1:]).max() > 1e-3 * np.abs(TF[0, 1:]).max()
    # zero velocity: the inertial-frame values
    L.or_fluid_frame(0.0, 0.0, m, out)
    assert np.array_equal(out, m[:3])
    for k in range(0, D, 5):
        d = np.zeros(3)
        L.or_fluid_dir(u[0], u[1], n[k, 0], n[k, 1], d)
        un = u @ n[k]
        aberr = (n[k] - g * u + (g - 1.0) * un * u / (u @ u)) / (g * (1.0 - un))
        assert abs(d[2] - g * (1.0 - un)) <= 1e-15
        assert np.abs(d[:2] - aberr).max() <= 1e-14
        assert abs(np.hypot(d[0], d[1]) - 1.0) <= 1e-14
