"""GPU adjoint of the fused step (gradients of an observer-ring loss with
respect to eta and kappa0) against the FP64 CPU adjoint oracle, central
differences on the GPU forward, and the torch.autograd wrapper."""
import importlib

import numpy as np
import pytest

from helpers import lib_context, rel_l2


def ring(n, r):
    th = 2 * np.pi * (np.arange(n) + 0.5) / n
    return r * np.cos(th), r * np.sin(th)


def build(rt, O, kind, n, L, D, precision, seed=11, k1=0.3):
    rng = np.random.default_rng(seed)
    gs = rt.GridSpec(n, n, (-L, -L), (L, L), 0.5)
    g = O.make_grid(n, n, (-L, -L), (L, L), 0.5)
    N = g.N
    bh = 2.1 * 2.1 if kind else -1.0
    M = O.metric(g, kind)
    tables = O.geodesic_tables(g, M, bh, 1e-4) if kind else None
    fields = {"eta": rng.uniform(0.2, 1.0, N), "kappa0": rng.uniform(0, 1, N), "kappa1": k1 * rng.uniform(0, 1, N)}
    P = O.Problem(g, D, 2 if kind else 0, M, bh_r2=bh, tables=tables, fields=fields)
    ctx, _ = lib_context(rt, gs, D, kind, fields, precision, bh_r2=bh, tables=tables)
    I0 = rng.uniform(0.5, 1.5, N * D)
    return P, ctx, g, fields, I0


CASES = [  # kind, n, L, D, steps, every, ring radius
    (0, 40, 1.0, 16, 20, 6, 0.7),
    (1, 40, 4.0, 16, 20, 6, 3.0),
    (1, 48, 4.0, 32, 15, 4, 3.5),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_adjoint_matches_oracle(rt, O, case):
    kind, n, L, D, steps, every, R = case
    print(f"[adjoint] kind={kind} n={n} D={D} steps={steps} every={every}", flush=True)
    px, py = ring(32, R)
    for prec, tol in ((rt.FP64, 1e-9), (rt.FP32_MIXED, 1e-4)):
        P, ctx, g, fields, I0 = build(rt, O, kind, n, L, D, prec)
        states = P.adj_forward(I0, steps)
        Lref, obs_ref, lamT = P.observer_loss(states[-1], px, py, np.zeros((32, D)), want_grad=True)
        ge_ref, gk_ref = P.adjoint(states, lamT)
        ctx.upload_I(I0)
        ctx.forward_checkpointed(steps, every)
        obs = ctx.sample_outgoing(px, py)
        assert rel_l2(obs, obs_ref) <= (1e-12 if prec == rt.FP64 else 1e-5)
        ge, gk = ctx.adjoint(px, py, obs)  # L = 1/2 sum obs^2
        assert rel_l2(ge, ge_ref) <= tol, (prec, rel_l2(ge, ge_ref))
        assert rel_l2(gk, gk_ref) <= tol, (prec, rel_l2(gk, gk_ref))
        # the state is left as the forward's final state
        assert rel_l2(ctx.sample_outgoing(px, py), obs) == 0.0
        ctx.close()


@pytest.mark.gpu
def test_adjoint_matches_central_differences_on_gpu(rt, O):
    """Schwarzschild 48x48, 32 directions, 30 steps: FD (h = 1e-3) of the FP64
    GPU forward on 8 random active cells against the GPU adjoint."""
    P, ctx, g, fields, I0 = build(rt, O, 1, 48, 4.0, 32, rt.FP64)
    px, py = ring(32, 3.5)
    steps, every = 30, 7

    def loss():
        ctx.upload_I(I0)
        ctx.forward_checkpointed(steps, every)
        o = ctx.sample_outgoing(px, py)
        return 0.5 * float(np.sum(o * o)), o

    L0, o = loss()
    ge, gk = ctx.adjoint(px, py, o)
    rng = np.random.default_rng(5)
    x = g.startx + (np.arange(g.N) % g.nx) * g.dx
    y = g.starty + (np.arange(g.N) // g.nx) * g.dy
    i = np.arange(g.N) % g.nx
    j = np.arange(g.N) // g.nx
    ok = (i > 1) & (i < g.nx - 2) & (j > 1) & (j < g.ny - 2) & (x * x + y * y > 2.1 * 2.1 + 0.5)
    cells = rng.choice(np.nonzero(ok)[0], 8, replace=False)
    for name, grad in (("eta", ge), ("kappa0", gk)):
        fd = []
        for c in cells:
            f = dict(fields)
            arr = f[name].copy()
            h = 1e-3
            arr[c] += h
            f[name] = arr
            ctx.set_fields(**f)
            lp = loss()[0]
            arr[c] -= 2 * h
            ctx.set_fields(**f)
            lm = loss()[0]
            fd.append((lp - lm) / (2 * h))
        ctx.set_fields(**fields)
        fd = np.array(fd)
        assert np.linalg.norm(grad[cells] - fd) <= 1e-5 * np.linalg.norm(fd), (name, grad[cells], fd)
    ctx.close()


@pytest.mark.gpu
def test_autograd_wrapper(rt, O):
    import torch
    ag = importlib.import_module("2dradiation_amd.autograd")
    P, ctx, g, fields, I0 = build(rt, O, 1, 40, 4.0, 16, rt.FP64)
    px, py = ring(32, 3.0)
    eta = torch.tensor(fields["eta"], requires_grad=True)
    k0 = torch.tensor(fields["kappa0"], requires_grad=True)
    extra = {"kappa1": fields["kappa1"]}
    out = ag.observer_ring(eta, k0, ctx, I0, 12, 5, px, py, extra)
    target = torch.zeros_like(out)
    loss = 0.5 * ((out - target) ** 2).sum()
    loss.backward()
    ge, gk = ctx.adjoint(px, py, out.detach().numpy())
    assert np.allclose(eta.grad.numpy(), ge, rtol=0, atol=1e-14 * np.abs(ge).max())
    assert np.allclose(k0.grad.numpy(), gk, rtol=0, atol=1e-14 * np.abs(gk).max())
    ctx.close()


@pytest.mark.gpu
def test_adjoint_rejects_bad_calls(rt, O):
    P, ctx, g, fields, I0 = build(rt, O, 0, 24, 1.0, 16, rt.FP64)
    px, py = ring(8, 0.5)
    with pytest.raises(rt.RtError) as e:
        ctx.adjoint(px, py, np.zeros((8, 16)))
    assert e.value.code == rt.RT_ERR_STATE
    ctx.forward_checkpointed(4, 2)
    with pytest.raises(rt.RtError):
        ctx.sample_outgoing(np.array([5.0]), np.array([0.0]))  # outside the grid
    f = dict(fields)
    f["ux"] = np.full(g.N, 0.1)
    ctx.set_fields(**f)
    with pytest.raises(rt.RtError) as e:
        ctx.forward_checkpointed(4, 2)
    assert e.value.code == rt.RT_ERR_UNSUPPORTED
    ctx.close()


@pytest.mark.gpu
def test_adjoint_after_download_keeps_checkpoint_convention(rt, O):
    """Curved FP64: downloading the state between the checkpointed forward and
    the adjoint turns the current state back into intensities; the adjoint
    still works in the convention its checkpoints were taken in."""
    kind, n, L, D, steps, every, R = CASES[1]
    px, py = ring(32, R)
    P, ctx, g, fields, I0 = build(rt, O, kind, n, L, D, rt.FP64)
    states = P.adj_forward(I0, steps)
    _, _, lamT = P.observer_loss(states[-1], px, py, np.zeros((32, D)), want_grad=True)
    ge_ref, gk_ref = P.adjoint(states, lamT)
    ctx.upload_I(I0)
    ctx.forward_checkpointed(steps, every)
    obs = ctx.sample_outgoing(px, py)
    I = ctx.download_I()
    assert rel_l2(I, states[-1]) <= 1e-12
    ge, gk = ctx.adjoint(px, py, obs)
    assert rel_l2(ge, ge_ref) <= 1e-9
    assert rel_l2(gk, gk_ref) <= 1e-9
    # and the current state is still the downloaded one
    assert rel_l2(ctx.download_I(), I) == 0.0
    ctx.close()


@pytest.mark.gpu
def test_autograd_backward_after_another_forward_raises(rt, O):
    import torch
    ag = importlib.import_module("2dradiation_amd.autograd")
    P, ctx, g, fields, I0 = build(rt, O, 1, 40, 4.0, 16, rt.FP64)
    px, py = ring(32, 3.0)
    extra = {"kappa1": fields["kappa1"]}
    eta = torch.tensor(fields["eta"], requires_grad=True)
    k0 = torch.tensor(fields["kappa0"], requires_grad=True)
    out1 = ag.observer_ring(eta, k0, ctx, I0, 6, 3, px, py, extra)
    out2 = ag.observer_ring(eta, k0, ctx, I0, 8, 3, px, py, extra)
    with pytest.raises(RuntimeError):
        (0.5 * (out1 ** 2).sum()).backward()
    (0.5 * (out2 ** 2).sum()).backward()
    ge, _ = ctx.adjoint(px, py, out2.detach().numpy())
    assert np.allclose(eta.grad.numpy(), ge, rtol=0, atol=1e-14 * np.abs(ge).max())
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("steps,every", [(13, 4), (12, 4), (5, 8), (2, 1)])
def test_adjoint_segment_layouts_and_repeat(rt, O, steps, every):
    """The checkpointed forward records its last segment (shorter than,
    equal to, or the only segment); the first adjoint starts on it, a second
    adjoint after the same forward recomputes it.  Both against the FP64
    adjoint oracle, and against each other."""
    print(f"[adjoint] layout steps={steps} every={every}", flush=True)
    P, ctx, g, fields, I0 = build(rt, O, 1, 40, 4.0, 16, rt.FP64)
    px, py = ring(32, 3.0)
    states = P.adj_forward(I0, steps)
    _, _, lamT = P.observer_loss(states[-1], px, py, np.zeros((32, 16)), want_grad=True)
    ge_ref, gk_ref = P.adjoint(states, lamT)
    ctx.upload_I(I0)
    ctx.forward_checkpointed(steps, every)
    obs = ctx.sample_outgoing(px, py)
    assert rel_l2(obs, P.observer_loss(states[-1], px, py, np.zeros((32, 16)))[1]) <= 1e-12
    ge1, gk1 = ctx.adjoint(px, py, obs)
    ge2, gk2 = ctx.adjoint(px, py, obs)
    for a, b, ref in ((ge1, ge2, ge_ref), (gk1, gk2, gk_ref)):
        assert rel_l2(a, ref) <= 1e-9
        assert rel_l2(b, ref) <= 1e-9
        assert rel_l2(a, b) <= 1e-12
    ctx.close()
