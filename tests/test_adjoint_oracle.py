"""CPU pins of the FP64 adjoint oracle: its forward against the reference
restatement, and its gradients against central finite differences of its own
forward (the independent check of a derivative)."""
import numpy as np
import pytest

from helpers import rel_l2


def ring(n, r):
    th = 2 * np.pi * (np.arange(n) + 0.5) / n
    return r * np.cos(th), r * np.sin(th)


def setup(O, kind, n=24, L=4.0, D=16, seed=7, k1=0.3):
    rng = np.random.default_rng(seed)
    g = O.make_grid(n, n, (-L, -L), (L, L), 0.5)
    N = g.N
    x = g.startx + (np.arange(N) % g.nx) * g.dx
    y = g.starty + (np.arange(N) // g.nx) * g.dy
    bh = 2.1 * 2.1 if kind else -1.0
    M = O.metric(g, kind)
    tables = O.geodesic_tables(g, M, bh, 1e-4) if kind else None
    fields = {"eta": rng.uniform(0.2, 1.0, N), "kappa0": rng.uniform(0, 1, N), "kappa1": k1 * rng.uniform(0, 1, N)}
    P = O.Problem(g, D, 2 if kind else 0, M, bh_r2=bh, tables=tables, fields=fields)
    I0 = rng.uniform(0.5, 1.5, N * D)  # anisotropic start: no moment vanishes by symmetry
    return P, g, I0


@pytest.mark.parametrize("kind", [0, 1])
def test_adjoint_forward_matches_reference_restatement(O, kind):
    P, g, I0 = setup(O, kind)
    states = P.adj_forward(I0, 12)
    Iref = P.run(I0, 12)[0]
    assert rel_l2(states[-1], Iref) <= 1e-12


@pytest.mark.parametrize("kind", [0, 1])
def test_adjoint_matches_central_differences(O, kind):
    P, g, I0 = setup(O, kind)
    nSteps = 12
    R = 2.5 if kind else 0.6 * (g.endx - g.dx)
    px, py = ring(32, R if kind else 0.7)
    target = np.zeros((32, P.nDir))
    states = P.adj_forward(I0, nSteps)
    L0, obs, lamT = P.observer_loss(states[-1], px, py, target, want_grad=True)
    assert L0 > 0
    ge, gk = P.adjoint(states, lamT)
    rng = np.random.default_rng(3)
    nx = g.nx
    cells = []
    while len(cells) < 8:
        i, j = rng.integers(2, nx - 2, 2)
        x, y = g.startx + i * g.dx, g.starty + j * g.dy
        if kind and x * x + y * y <= 2.1 * 2.1 + 0.5:
            continue
        cells.append(i + j * nx)

    def loss():
        st = P.adj_forward(I0, nSteps)
        return P.observer_loss(st[-1], px, py, target)[0]

    for f, grad in ((0, ge), (1, gk)):
        arr = P.fields[f]
        fd = []
        for c in cells:
            h = 1e-3
            old = arr[c]
            arr[c] = old + h
            lp = loss()
            arr[c] = old - h
            lm = loss()
            arr[c] = old
            fd.append((lp - lm) / (2 * h))
        fd = np.array(fd)
        ad = grad[cells]
        assert np.linalg.norm(ad - fd) <= 1e-5 * np.linalg.norm(fd) + 1e-12, (f, ad, fd)
