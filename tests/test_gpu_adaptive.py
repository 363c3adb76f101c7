"""Adaptive streaming on the GPU (StreamingType FlatAdaptive and
CurvedAdaptive: per-cell rotated stencils with ghost directions;
src/Radiation.cpp:345-378, 415-453, 513-575, 784-832) against the oracle.

FP64: rel-L2 <= 1e-12 of intensities and moments, rotation angles to 1e-12,
identical lambda-iteration statistics.  FP32-mixed (FP32 state, FP64
arithmetic): rel-L2 <= 1e-5 against FP64."""
import math

import numpy as np
import pytest

from helpers import rel_l2, grids

pytestmark = pytest.mark.gpu


def smooth_state(N, phi, seed):
    """Per cell a + b cos(phi - c) + e cos 2(phi - f), a > |b| + |e|: each cell
    has its own flux direction, so the rotation update turns the stencils."""
    rng = np.random.default_rng(seed)
    b, e = rng.uniform(-0.45, 0.45, N), rng.uniform(-0.2, 0.2, N)
    a = 0.7 + rng.uniform(0, 0.3, N)
    c, f = rng.uniform(0, 2 * math.pi, N), rng.uniform(0, 2 * math.pi, N)
    I = a[:, None] + b[:, None] * np.cos(phi[None, :] - c[:, None]) + e[:, None] * np.cos(2 * (phi[None, :] - f[:, None]))
    return np.ascontiguousarray(I.reshape(-1))


def fields_for(N, seed, u=True):
    rng = np.random.default_rng(seed)
    f = {"eta": rng.uniform(0, 1, N), "kappa0": rng.uniform(0, 1, N), "kappa1": rng.uniform(0, 0.3, N),
         "kappaA": rng.uniform(0, 0.2, N)}
    if u:
        f["ux"] = rng.uniform(-0.2, 0.2, N)
        f["uy"] = rng.uniform(-0.2, 0.2, N)
    return f


def setup(rt, O, kind, n_order, n_ghost, n, precision, fields, bh_r2=-1.0, span=1.0):
    """Oracle Problem and library Context of the same adaptive problem; the
    curved tables come from the oracle so both stream with the same series."""
    gs, og = grids(rt, O, n, n, (-span, -span), (span, span))
    D = n_order + n_ghost
    M = O.metric(og, kind, 1.0, 0.0)
    tables = O.geodesic_tables(og, M, bh_r2, 1e-4) if kind else None
    P = O.Problem(og, D, 3 if kind else 1, M, bh_r2=bh_r2, tables=tables, fields=fields, nGhost=n_ghost)
    ctx = rt.Context(gs, D, rt.CURVED_ADAPTIVE if kind else rt.FLAT_ADAPTIVE, precision, bh_r2)
    w, phi, cx, cy, Nc, Wc = rt.build_stencil_ghost(n_order, n_ghost)
    assert np.array_equal(phi, P.phi)
    ctx.set_stencil(w, phi, cx, cy, Nc, Wc)
    block, alpha, t11, t12, t21, t22 = rt.build_metric(kind, 1.0, 0.0, gs)
    ctx.set_geometry(alpha, t11, t12, t21, t22)
    if kind:
        ctx.set_streaming_stencil(*rt.build_stencil(5, tables=False))
        ctx.set_fourier_tables(*tables)
    ctx.set_fields(**fields)
    return P, ctx, gs, og


def initial(P, gs, seed):
    rng = np.random.default_rng(seed)
    I0 = smooth_state(gs.N, P.phi, seed)
    rot = rng.uniform(0, 2 * math.pi, gs.N)
    rot_new = rng.uniform(0, 2 * math.pi, gs.N)
    P.set_rotation(rot, rot_new)
    mom = P.moments(I0)
    return I0, rot, rot_new, mom


@pytest.mark.parametrize("kind,n_order,n_ghost,n,steps", [(0, 16, 4, 30, 8), (2, 16, 4, 30, 6), (0, 32, 8, 24, 5),
                                                          (2, 40, 10, 22, 4)])
def test_adaptive_steps_vs_oracle(rt, O, kind, n_order, n_ghost, n, steps):
    print(f"[adaptive] kind {kind}, Stencil({n_order},{n_ghost}), {n}^2, {steps} steps", flush=True)
    bh = 2.1 * 2.1 if kind else -1.0
    span = 4.0 if kind else 1.0
    N = (n + 2) ** 2
    fields = fields_for(N, 40 + kind, u=(kind == 0))
    P, ctx, gs, og = setup(rt, O, kind, n_order, n_ghost, n, rt.FP64, fields, bh, span)
    I0, rot, rot_new, mom = initial(P, gs, 7 + n_order)
    I, m, r, rn, mx, sm, _ = P.run_adaptive(I0, steps, mom=mom, rot=rot, rotNew=rot_new)
    assert not np.allclose(r, rot)                       # the stencils did turn
    ctx.upload_I(I0)
    ctx.set_rotation(rot, rot_new)
    ctx.set_moments(mom)
    ctx.reset_iteration_stats()
    ctx.step(steps)
    Ig = ctx.download_I()
    mg = ctx.moments()
    rg, rng_ = ctx.rotation()
    assert rel_l2(Ig, I) <= 1e-12, rel_l2(Ig, I)
    assert rel_l2(mg, m) <= 1e-12
    assert np.max(np.abs(rg - r)) <= 1e-12 and np.max(np.abs(rng_ - rn)) <= 1e-12
    assert ctx.iteration_stats() == (mx, sm)
    assert not ctx.uses_fused()
    ctx.close()


@pytest.mark.parametrize("kind,n_order,n_ghost,nx,ny", [(0, 8, 2, 37, 11), (2, 24, 6, 70, 5), (0, 120, 30, 9, 67)])
def test_adaptive_ragged_shapes(rt, O, kind, n_order, n_ghost, nx, ny):
    """Grids that are no multiple of the 64-cell workgroup strip and
    direction counts that are no multiple of 4 waves (Stencil(8,2) = 10,
    (24,6) = 30, (120,30) = 150 directions, the reference benchmark's
    stencils)."""
    print(f"[adaptive ragged] kind {kind}, Stencil({n_order},{n_ghost}), {nx}x{ny}", flush=True)
    bh = 2.1 * 2.1 if kind else -1.0
    span = 4.0 if kind else 1.0
    gs, og = grids(rt, O, nx, ny, (-span, -span), (span, span))
    D = n_order + n_ghost
    fields = fields_for(gs.N, 17, u=(kind == 0))
    M = O.metric(og, kind, 1.0, 0.0)
    tables = O.geodesic_tables(og, M, bh, 1e-4) if kind else None
    P = O.Problem(og, D, 3 if kind else 1, M, bh_r2=bh, tables=tables, fields=fields, nGhost=n_ghost)
    ctx = rt.Context(gs, D, rt.CURVED_ADAPTIVE if kind else rt.FLAT_ADAPTIVE, rt.FP64, bh)
    w, phi, cx, cy, Nc, Wc = rt.build_stencil_ghost(n_order, n_ghost)
    ctx.set_stencil(w, phi, cx, cy, Nc, Wc)
    block, alpha, t11, t12, t21, t22 = rt.build_metric(kind, 1.0, 0.0, gs)
    ctx.set_geometry(alpha, t11, t12, t21, t22)
    if kind:
        ctx.set_streaming_stencil(*rt.build_stencil(5, tables=False))
        ctx.set_fourier_tables(*tables)
    ctx.set_fields(**fields)
    I0, rot, rot_new, mom = initial(P, gs, 5 + D)
    I, m, r, rn, mx, sm, _ = P.run_adaptive(I0, 5, mom=mom, rot=rot, rotNew=rot_new)
    ctx.upload_I(I0)
    ctx.set_rotation(rot, rot_new)
    ctx.set_moments(mom)
    ctx.step(5)
    assert rel_l2(ctx.download_I(), I) <= 1e-12
    assert rel_l2(ctx.moments(), m) <= 1e-12
    assert np.max(np.abs(ctx.rotation()[0] - r)) <= 1e-12
    ctx.close()


def test_adaptive_single_methods_vs_oracle(rt, O):
    """UpdateRotationMatrizes, StreamCurvedAdaptive (with its swap of I and of
    the angle buffers), ComputeMomentsIF and Collide one by one."""
    n, n_order, n_ghost = 26, 16, 4
    N = (n + 2) ** 2
    fields = fields_for(N, 3, u=False)
    P, ctx, gs, og = setup(rt, O, 2, n_order, n_ghost, n, rt.FP64, fields, 2.1 * 2.1, 4.0)
    I0, rot, rot_new, mom = initial(P, gs, 11)
    ctx.upload_I(I0)
    ctx.set_rotation(rot, rot_new)
    ctx.set_moments(mom)
    ctx.update_rotation()
    r_cur, r_new = ctx.rotation()
    exp_new = P.update_rotation(mom, rot)
    inner = np.zeros((gs.ny, gs.nx), bool)
    inner[1:-1, 1:-1] = True
    inner = inner.reshape(-1)
    assert np.array_equal(r_cur, rot)
    assert np.max(np.abs(r_new[inner] - exp_new[inner])) <= 1e-12
    assert np.array_equal(r_new[~inner], rot_new[~inner])
    ctx.stream()
    P.set_rotation(rot, r_new)
    S = P.stream_adaptive(I0)
    Ig = ctx.download_I()
    D = n_order + n_ghost
    interior = np.repeat(inner, D)
    assert rel_l2(Ig[interior], S[interior]) <= 1e-12
    a, b = ctx.rotation()
    assert np.array_equal(a, r_new) and np.array_equal(b, rot)   # swapped
    ctx.compute_moments()
    P.set_rotation(r_new, rot)
    # the streamed buffer is the scratch buffer, zero outside the interior
    # after rt_upload_I (the reference's Inew)
    S_full = np.where(interior, S, 0.0)
    assert rel_l2(ctx.moments(), P.moments(S_full)) <= 1e-12
    ctx.collide()
    mom2 = P.moments(S_full).reshape(-1).copy()
    out = S_full.copy()
    out[:] = I0          # collide writes the interior active cells of the other buffer
    cnt = np.zeros(gs.N, np.int32)
    P.L.or_collide(P.sim, S_full, out, mom2, cnt.ctypes.data)
    assert rel_l2(ctx.download_I(), out) <= 1e-12
    assert np.array_equal(ctx.iteration_counts(), cnt)
    ctx.close()


@pytest.mark.parametrize("kind,motion", [(0, "rest"), (2, "rest"), (2, "moving"), (0, "mixed")])
def test_adaptive_fp32_mixed(rt, O, kind, motion):
    """FP32-mixed (moment-form collide for cells at rest, the reference cycle
    for moving ones, both in one launch for "mixed") against FP64 (the
    reference-order collide)."""
    n, n_order, n_ghost, steps = 28, 16, 4, 10
    N = (n + 2) ** 2
    bh = 2.1 * 2.1 if kind else -1.0
    span = 4.0 if kind else 1.0
    fields = fields_for(N, 9, u=motion != "rest")
    if motion == "mixed":
        still = np.random.default_rng(5).uniform(0, 1, N) < 0.5
        fields["ux"][still] = 0.0
        fields["uy"][still] = 0.0
    out = {}
    for prec in (rt.FP64, rt.FP32_MIXED):
        P, ctx, gs, og = setup(rt, O, kind, n_order, n_ghost, n, prec, fields, bh, span)
        I0, rot, rot_new, mom = initial(P, gs, 21)
        ctx.upload_I(I0)
        ctx.set_rotation(rot, rot_new)
        ctx.set_moments(mom)
        ctx.step(steps)
        out[prec] = (ctx.download_I(), ctx.moments())
        ctx.close()
    assert rel_l2(out[rt.FP32_MIXED][0], out[rt.FP64][0]) <= 1e-5
    assert rel_l2(out[rt.FP32_MIXED][1][:3], out[rt.FP64][1][:3]) <= 1e-5


def test_adaptive_wide_stencil(rt, O):
    """Stencil(160,40), the reference's widest adaptive stencil (200 directions)."""
    n = 16
    N = (n + 2) ** 2
    fields = fields_for(N, 5)
    P, ctx, gs, og = setup(rt, O, 0, 160, 40, n, rt.FP64, fields)
    I0, rot, rot_new, mom = initial(P, gs, 31)
    I, m, r, rn, mx, sm, _ = P.run_adaptive(I0, 3, mom=mom, rot=rot, rotNew=rot_new)
    ctx.upload_I(I0)
    ctx.set_rotation(rot, rot_new)
    ctx.set_moments(mom)
    ctx.step(3)
    assert rel_l2(ctx.download_I(), I) <= 1e-12
    assert rel_l2(ctx.moments(), m) <= 1e-12
    ctx.close()


def test_adaptive_sources_vs_oracle(rt, O):
    """keepSourceNodesActive: intensities and angles of the source cells are
    re-imposed before every step; a NaN angle leaves the cell's angle alone
    (LoadInitialData skips cells without initial flux)."""
    n, n_order, n_ghost, steps = 24, 16, 4, 6
    N = (n + 2) ** 2
    D = n_order + n_ghost
    fields = fields_for(N, 13, u=False)
    P, ctx, gs, og = setup(rt, O, 0, n_order, n_ghost, n, rt.FP64, fields)
    I0, rot, rot_new, mom = initial(P, gs, 41)
    rng = np.random.default_rng(3)
    cells = rng.choice(np.arange(gs.nx + 1, gs.N - gs.nx - 1), 8, replace=False)
    src = np.zeros(gs.N, np.uint8)
    src[cells] = 1
    src_rot = np.full(gs.N, np.nan)
    src_rot[cells[:5]] = rng.uniform(0, 2 * math.pi, 5)
    # oracle: one step at a time, sources put back first
    I, m, r, rn = I0.copy(), mom.copy(), rot.copy(), rot_new.copy()
    mask = np.repeat(src.astype(bool), D)
    for _ in range(steps):
        I[mask] = I0[mask]
        keep = ~np.isnan(src_rot)
        r[keep] = src_rot[keep]
        I, m, r, rn, _, _, _ = P.run_adaptive(I, 1, mom=m, rot=r, rotNew=rn)
    ctx.upload_I(I0)
    ctx.set_rotation(rot, rot_new)
    ctx.set_moments(mom)
    ctx.set_sources(src, I0)
    ctx.set_source_rotation(src_rot)
    ctx.step(steps)
    assert rel_l2(ctx.download_I(), I) <= 1e-12
    assert np.max(np.abs(ctx.rotation()[0] - r)) <= 1e-12
    ctx.close()


def test_adaptive_errors(rt, O):
    n = 12
    N = (n + 2) ** 2
    P, ctx, gs, og = setup(rt, O, 0, 16, 4, n, rt.FP64, fields_for(N, 1, u=False))
    with pytest.raises(rt.RtError, match="rotation angles"):
        ctx.set_rotation(np.full(gs.N, 9.0))
    with pytest.raises(rt.RtError):
        ctx.forward_checkpointed(2, 1)         # the adjoint covers the fused fixed step only
    with pytest.raises(rt.RtError):
        ctx.set_source_rotation(np.zeros(gs.N))  # no source cells named
    ctx.close()
    from helpers import lib_context
    fixed, _ = lib_context(rt, gs, 16, 0, {"eta": np.ones(gs.N)}, rt.FP64)
    with pytest.raises(rt.RtError, match="adaptive"):
        fixed.update_rotation()
    with pytest.raises(rt.RtError, match="adaptive"):
        fixed.set_rotation(np.zeros(gs.N))
    fixed.close()
