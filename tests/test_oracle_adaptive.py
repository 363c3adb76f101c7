"""Pins of the oracle's adaptive streaming (rotated stencils with ghost
directions; src/Radiation.cpp:345-378, 415-453, 513-575; src/Stencil.cpp:177-332).

The reference has no assertions for this path either.  Each check below is
fixed by something other than the oracle's own formula: ghost layouts worked
out by hand from the construction rules, libm, exact identities of the
rotation at angle 0, relabelling of a uniform stencil under a rotation by
whole stencil steps, and flat-space geodesic tables."""
import math

import numpy as np
import pytest

from helpers import rel_l2

DEG = math.pi / 180.0


def circ(a, b):
    d = abs(a - b) % (2 * math.pi)
    return min(d, 2 * math.pi - d)


# ------------------------------------------------------------ ghost stencils
def test_myacos_error_bound(O):
    L = O.lib()
    xs = np.linspace(-1, 1, 4001)
    assert max(abs(L.or_myacos(v) - math.acos(np.float32(v))) for v in xs) < 1e-4
    assert L.or_myacos(1.0) == 0.0 and L.or_myacos(-1.0) == np.float32(math.pi)  # sqrt(1-x) vanishes at the ends


def test_ghost_stencil_closed_forms(O):
    # Stencil(4,2): no gap lies within pi/8 of pi, so the fallback bucket
    # [pi - pi/4, pi + pi/4] takes both ghosts at its quarter points, turned
    # by pi: 15pi/8 and 17pi/8 = pi/8.
    w, phi, cx, cy = O.stencil_ghost(4, 2)
    exp = sorted([math.pi / 4, 3 * math.pi / 4, 5 * math.pi / 4, 7 * math.pi / 4, math.pi / 8, 15 * math.pi / 8])
    assert np.allclose(phi, exp, atol=1e-14, rtol=0)
    assert list(w == 0) == [p in (math.pi / 8, 15 * math.pi / 8) for p in exp]
    # Stencil(16,4): one gap [15pi/16, 17pi/16] around pi holds all 4 ghosts
    # at 15pi/16 + (pi/8)(j+1/2)/4, turned by pi.
    w, phi, cx, cy = O.stencil_ghost(16, 4)
    g = np.sort(phi[w == 0])
    exp = sorted(((31 * math.pi / 16 + math.pi * (j + 0.5) / 32) % (2 * math.pi)) for j in range(4))
    assert np.allclose(g, exp, atol=1e-14, rtol=0)
    # Stencil(40,10): gaps between 157.5..202.5 deg (5 gaps of 9 deg; the
    # outer ends lie exactly pi/8 from pi, inside thanks to the +1e-8);
    # arccos samples 157.5 + 45 acos(1-2x)/pi at x = 0.05..0.95 fall
    # 1, 2, 4, 2, 1 into them.
    w, phi, cx, cy = O.stencil_ghost(40, 10)
    exp_deg = [162.0, 168.75, 173.25, 176.625, 178.875, 181.125, 183.375, 186.75, 191.25, 198.0]
    exp = sorted(((a + 180.0) % 360.0) * DEG for a in exp_deg)
    assert np.allclose(np.sort(phi[w == 0]), exp, atol=1e-12, rtol=0)


@pytest.mark.parametrize("n_order,n_ghost", [(16, 4), (40, 10), (80, 20), (160, 40), (32, 8)])
def test_ghost_stencil_invariants(O, n_order, n_ghost):
    L = O.lib()
    w, phi, cx, cy = O.stencil_ghost(n_order, n_ghost)
    assert len(w) == n_order + n_ghost
    assert np.all(np.diff(phi) > 0)                       # sorted, distinct
    assert np.sum(w == 0) == n_ghost
    assert abs(w.sum() - 1.0) < 1e-13
    fw, fphi, _, _ = O.stencil(n_order)
    assert np.array_equal(phi[w > 0], fphi)                # Fourier directions untouched
    g = phi[w == 0]
    assert max(circ(a, 0.0) for a in g) <= math.pi / 8 + 1e-8   # ghosts around the forward direction
    # the construction is mirror symmetric about pi, so the ghosts are about 0
    assert np.allclose(np.sort(g), np.sort((2 * math.pi - g) % (2 * math.pi)), atol=1e-12)
    for d in range(len(phi)):
        assert cx[d] == L.or_mycos(phi[d]) and cy[d] == L.or_mysin(phi[d])
    # cubic tables on the non-uniform stencil still reproduce cubics in angle
    Nc, Wc = O.interp_grid(phi, 10 * n_order)
    assert np.allclose(Wc.reshape(-1, 4).sum(1), 1.0, atol=1e-12)


def test_ghost_stencil_rejections(O):
    with pytest.raises(ValueError):
        O.stencil_ghost(15, 3)  # odd nOrder


# ------------------------------------------------------------ rotation update
def flat_problem(O, n=12, nDir=16, nGhost=0, streaming=1, fields=None, cfl=0.5):
    g = O.make_grid(n, n, (-1, -1), (1, 1), cfl)
    M = O.metric(g, 0)
    return O.Problem(g, nDir, streaming, M, fields=fields or {}, nGhost=nGhost), g


def test_update_rotation_closed_forms(O):
    P, g = flat_problem(O)
    N = g.N
    beta = 2.3
    mom = np.zeros((6, N))
    mom[0] = 1.0
    mom[1] = 0.5 * math.cos(beta)
    mom[2] = 0.5 * math.sin(beta)
    rot = np.full(N, 0.7)
    new = P.update_rotation(mom, rot)
    inner = np.zeros((g.ny, g.nx), bool)
    inner[1:-1, 1:-1] = True
    inner = inner.reshape(-1)
    assert np.all(np.abs(new[inner] - beta) < 2e-5)        # Hastings atan error
    assert np.all(new[~inner] == 0.0)                      # halo untouched
    # tiny energy or a flux below 1% of E keeps the old angle
    mom2 = mom.copy()
    mom2[0] = 1e-17
    assert np.all(P.update_rotation(mom2, rot)[inner] == 0.7)
    mom3 = mom.copy()
    mom3[1:3] *= 0.0199  # |F|/E = 0.00995
    assert np.all(P.update_rotation(mom3, rot)[inner] == 0.7)
    # a flux at one cell reaches exactly its 3x3 neighbourhood
    mom4 = np.zeros((6, N))
    mom4[0] = 1.0
    c = 5 + 6 * g.nx
    mom4[2, c] = 0.9                                        # pointing up: angle pi/2
    out = P.update_rotation(mom4, rot).reshape(g.ny, g.nx)
    moved = np.argwhere(out != 0.7)
    moved = [tuple(m) for m in moved if 0 < m[0] < g.ny - 1 and 0 < m[1] < g.nx - 1]
    assert sorted(moved) == sorted((6 + b, 5 + a) for b in (-1, 0, 1) for a in (-1, 0, 1))
    assert abs(out[6, 5] - math.pi / 2) < 1e-12


# ------------------------------------------------------------ adaptive stream
def random_state(g, nDir, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.1, 1.0, g.N * nDir)


def smooth_state(g, phi, seed):
    """Per cell a + b cos(phi - c) + e cos 2(phi - f) with random a > |b| + |e|:
    smooth in angle, so that angular interpolation adds no visible error."""
    rng = np.random.default_rng(seed)
    N = g.N
    b, e = rng.uniform(-0.3, 0.3, N), rng.uniform(-0.2, 0.2, N)
    a = 0.6 + rng.uniform(0, 0.4, N)
    c, f = rng.uniform(0, 2 * math.pi, N), rng.uniform(0, 2 * math.pi, N)
    I = a[:, None] + b[:, None] * np.cos(phi[None, :] - c[:, None]) + e[:, None] * np.cos(2 * (phi[None, :] - f[:, None]))
    return np.ascontiguousarray(I.reshape(-1))


def test_zero_rotation_is_fixed_streaming(O):
    # At angle 0 the adaptive stream reads each corner through IntensityAt
    # at the direction's own angle, which the fixed stream reads directly:
    # equal up to the Hastings atan error in the angle (~1e-5 rad).
    P, g = flat_problem(O, nDir=16)
    I = smooth_state(g, P.phi, 1)
    P.set_rotation(np.zeros(g.N), np.zeros(g.N))
    a = P.stream_adaptive(I)
    f = np.zeros_like(a)
    P.L.or_stream_flat(P.sim, I, f)
    assert np.max(np.abs(a - f)) < 1e-5
    # rotated moments at angle 0 are the fixed ones, bit for bit
    m_rot = P.moments(I)
    P.sim.rot = None
    assert np.array_equal(m_rot, P.moments(I))


def test_zero_rotation_simulation_is_fixed_simulation(O):
    fields = None
    P, g = flat_problem(O, nDir=16, fields=fields)
    rng = np.random.default_rng(5)
    N = g.N
    P2, _ = flat_problem(O, nDir=16, streaming=0)
    for arr, v in zip(P.fields, (rng.uniform(0, 1, N), rng.uniform(0, 2, N), rng.uniform(0, .3, N),
                                 np.zeros(N), rng.uniform(-.2, .2, N), rng.uniform(-.2, .2, N))):
        arr[:] = v
    for arr, src in zip(P2.fields, P.fields):
        arr[:] = src
    I0 = smooth_state(g, P.phi, 2)
    Ia, ma, ra, rna, mxa, sa, _ = P.run_adaptive(I0, 1)     # zero moments: the angles stay 0
    If, mf, mxf, sf, _ = P2.run(I0, 1)
    assert rel_l2(Ia, If) < 1e-5 and rel_l2(ma, mf) < 1e-5
    assert np.all(ra == 0) and np.all(rna == 0)


@pytest.mark.parametrize("k", [1, 3, -2])
def test_uniform_rotation_relabels_fixed_streaming(O, k):
    # Rotating every cell's stencil by k steps of 2pi/D turns direction d into
    # physical direction d+k: the adaptive stream of I equals the fixed stream
    # of I relabelled, up to the Hastings sin/cos error (~1e-7).
    D = 16
    P, g = flat_problem(O, nDir=D)
    I = smooth_state(g, P.phi, 3 + k).reshape(g.N, D)
    theta = (2 * math.pi * k / D) % (2 * math.pi)
    P.set_rotation(np.full(g.N, theta), np.full(g.N, theta))
    a = P.stream_adaptive(I.reshape(-1)).reshape(g.N, D)
    Irel = np.roll(I, k, axis=1)                          # Irel[:, d+k] = I[:, d]
    f = np.zeros(g.N * D)
    P.L.or_stream_flat(P.sim, np.ascontiguousarray(Irel.reshape(-1)), f)
    f = np.roll(f.reshape(g.N, D), -k, axis=1)            # back to the rotated labels
    inner = np.zeros((g.ny, g.nx), bool)
    inner[1:-1, 1:-1] = True
    inner = inner.reshape(-1)
    assert np.max(np.abs(a[inner] - f[inner])) < 2e-5
    # the wrong sense of rotation is far off
    f_wrong = np.zeros(g.N * D)
    P.L.or_stream_flat(P.sim, np.ascontiguousarray(np.roll(I, -k, axis=1).reshape(-1)), f_wrong)
    f_wrong = np.roll(f_wrong.reshape(g.N, D), k, axis=1)
    assert np.max(np.abs(a[inner] - f_wrong[inner])) > 1e-2


def test_rotated_moments_single_direction(O):
    D = 16
    P, g = flat_problem(O, nDir=D)
    rng = np.random.default_rng(9)
    theta = rng.uniform(0, 2 * math.pi, g.N)
    P.set_rotation(theta)
    d0 = 5
    I = np.zeros((g.N, D))
    I[:, d0] = 1.0
    m = P.moments(I.reshape(-1))
    ang = theta + P.phi[d0]
    w = 1.0 / D
    assert np.allclose(m[0], w, atol=1e-15)
    assert np.allclose(m[1], w * np.cos(ang), atol=1e-7 * w)
    assert np.allclose(m[2], w * np.sin(ang), atol=1e-7 * w)
    assert np.allclose(m[3], w * np.cos(ang) ** 2, atol=1e-7 * w)
    assert np.allclose(m[4], w * np.cos(ang) * np.sin(ang), atol=1e-7 * w)


def test_rotated_collide_kappa_zero(O):
    # kappa = 0: Inew = I + alpha dt eta / A^2 with the Doppler factor
    # A = gamma (1 - u.n) of the rotated direction n = R(theta) C(d).
    D = 16
    rng = np.random.default_rng(11)
    P, g = flat_problem(O, nDir=D)
    N = g.N
    eta = rng.uniform(0, 1, N)
    ux = rng.uniform(-0.3, 0.3, N)
    uy = rng.uniform(-0.3, 0.3, N)
    for arr, v in zip(P.fields, (eta, np.zeros(N), np.zeros(N), np.zeros(N), ux, uy)):
        arr[:] = v
    theta = rng.uniform(0, 2 * math.pi, N)
    P.set_rotation(theta)
    I = random_state(g, D, 12)
    mom = P.moments(I).reshape(-1).copy()
    out = I.copy()
    cnt = np.zeros(N, np.int32)
    P.L.or_collide(P.sim, I, out, mom, cnt.ctypes.data)
    inner = np.zeros((g.ny, g.nx), bool)
    inner[1:-1, 1:-1] = True
    inner = inner.reshape(-1)
    ang = theta[:, None] + P.phi[None, :]
    un = ux[:, None] * np.cos(ang) + uy[:, None] * np.sin(ang)
    gam = 1 / np.sqrt(1 - ux ** 2 - uy ** 2)
    A = gam[:, None] * (1 - un)
    exp = I.reshape(N, D) + g.dt * eta[:, None] / A ** 2
    assert np.allclose(out.reshape(N, D)[inner], exp[inner], rtol=1e-6, atol=0)


def test_curved_adaptive_flat_tables_match_flat_adaptive(O):
    # Minkowski geodesic tables are straight lines fitted exactly by the five
    # Fourier modes, so curved adaptive streaming evaluated at the rotated
    # angles must agree with flat adaptive streaming to the trig error.
    D = 16
    g = O.make_grid(14, 14, (-1, -1), (1, 1), 0.5)
    M = O.metric(g, 0)
    tables = O.geodesic_tables(g, M, -1.0)
    Pc = O.Problem(g, D + 4, 3, M, tables=tables, nGhost=4)
    Pf = O.Problem(g, D + 4, 1, M, nGhost=4)
    rng = np.random.default_rng(13)
    theta = rng.uniform(0, 2 * math.pi, g.N)
    theta_new = rng.uniform(0, 2 * math.pi, g.N)
    for P in (Pc, Pf):
        P.set_rotation(theta, theta_new)
    I = smooth_state(g, Pc.phi, 14)
    a = Pc.stream_adaptive(I)
    b = Pf.stream_adaptive(I)
    inner = np.zeros((g.ny, g.nx), bool)
    inner[1:-1, 1:-1] = True
    inner = np.repeat(inner.reshape(-1), D + 4)
    assert np.max(np.abs(a[inner] - b[inner])) < 1e-5
    # evaluating the series at the unrotated stencil angles would not agree
    Pc.set_rotation(theta, np.zeros(g.N))
    c = Pc.stream_adaptive(I)
    assert np.max(np.abs(c[inner] - b[inner])) > 1e-2


def test_adaptive_run_angle_buffers(O):
    # After one step rotationAngle holds the updated angles (halo: the old
    # rotationAngleNew), rotationAngleNew the old angles (src/Radiation.cpp:452).
    D = 20
    P, g = flat_problem(O, nDir=D, nGhost=4)
    N = g.N
    rng = np.random.default_rng(17)
    rot = rng.uniform(0, 2 * math.pi, N)
    rot_new = rng.uniform(0, 2 * math.pi, N)
    I = random_state(g, D, 18)
    P.set_rotation(rot, rot_new)
    mom = P.moments(I)
    expect = P.update_rotation(mom, rot)
    inner = np.zeros((g.ny, g.nx), bool)
    inner[1:-1, 1:-1] = True
    inner = inner.reshape(-1)
    _, _, r1, rn1, _, _, _ = P.run_adaptive(I, 1, mom=mom, rot=rot, rotNew=rot_new)
    assert np.array_equal(r1[inner], expect[inner])
    assert np.array_equal(r1[~inner], rot_new[~inner])
    assert np.array_equal(rn1, rot)
    # energy bookkeeping: pure streaming (no fields) keeps the intensity
    # positive and finite over several steps
    I5, m5, r5, rn5, _, _, _ = P.run_adaptive(I, 5, mom=mom, rot=rot, rotNew=rot_new)
    assert np.all(np.isfinite(I5)) and np.all(I5 >= 0)
