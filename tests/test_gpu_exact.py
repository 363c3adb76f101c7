"""The product step from I = 0 with scattering, against the oracle, on reduced
grids of the benchmark workloads (their fields, spacetimes and stencils).

FP64 (fused stream + exact collide): rel-L2 <= 1e-12 in I and the moments,
and the same lambda-cycle count in every cell at every step.  From an
isotropic start the reference's stopping test divides moments that are pure
rounding noise, so its cycle counts follow the exact rounding of the
d-ordered moment sums; matching them is what this file pins.

FP32-mixed on the same problems: rel-L2 <= 1e-5 against the FP64 oracle.
All calls go through the C ABI.
"""
import numpy as np
import pytest

from helpers import rel_l2, active_mask

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5

# (workload, reduced interior size, domain half-width, spacetime kind, nDir,
#  excision r^2, RK45 tol, steps)
CASES = {
    "schwarzschild_1024sq_d64": (128, 8.0, 1, 64, 2.1 ** 2, 1e-4, 24),
    "kerrschild_200sq_d32": (100, 4.0, 2, 32, 1.6 ** 2, 1e-10, 30),
    "schwarzschild_256sq_d32": (96, 4.0, 1, 32, 2.1 ** 2, 1e-4, 30),
}


def setup(rt, O, name, precision):
    n, L, kind, nDir, r2, tol, steps = CASES[name]
    gs = rt.GridSpec(n, n, (-L, -L), (L, L), 0.5)
    og = O.make_grid(n, n, (-L, -L), (L, L), 0.5)
    f = rt.workloads.fields(name, gs)
    M = O.metric(og, kind)
    tabs = O.geodesic_tables(og, M, r2, tol)
    P = O.Problem(og, nDir, 2, M, bh_r2=r2, tables=tabs, fields=f)
    ctx = rt.Context(gs, nDir, rt.CURVED_FIXED, precision, r2)
    w, phi, cx, cy, Nc, Wc = rt.build_stencil(nDir)
    ctx.set_stencil(w, phi, cx, cy, Nc, Wc)
    _, alpha, t11, t12, t21, t22 = rt.build_metric(kind, 1.0, 0.0, gs)
    ctx.set_geometry(alpha, t11, t12, t21, t22)
    ctx.set_streaming_stencil(*rt.build_stencil(5, tables=False))
    ctx.set_fourier_tables(*tabs)
    ctx.set_fields(**f)
    ctx.upload_I(np.zeros(gs.N * nDir))
    return gs, og, P, ctx, steps, nDir, r2


@pytest.mark.parametrize("name", list(CASES))
def test_fp64_step_from_zero_matches_oracle_cell_by_cell(rt, O, name):
    gs, og, P, ctx, steps, nDir, r2 = setup(rt, O, name, rt.FP64)
    print(f"[exact] {name} {gs.nx}x{gs.ny} D={nDir} steps={steps} fused={ctx.uses_fused()}", flush=True)
    Iref, mref, mx, sm, counts = P.run(np.zeros(og.N * nDir), steps, want_counts=True)
    counts = counts.reshape(steps, og.N)
    ctx.reset_iteration_stats()
    for s in range(steps):
        ctx.step(1)
        got = ctx.iteration_counts()
        bad = np.flatnonzero(got != counts[s])
        assert bad.size == 0, f"step {s}: {bad.size} cells differ in cycle count, first {bad[:5]}"
    I = ctx.download_I()
    assert rel_l2(I, Iref) <= TOL64
    act = active_mask(gs, r2)
    m = ctx.moments()
    for X in range(6):
        assert rel_l2(m[X][act], mref[X][act]) <= TOL64, X
    gmx, gsm = ctx.iteration_stats()
    assert (gmx, gsm) == (mx, sm)
    # the isotropic start is where the counts are decided by rounding noise
    assert counts[0].max() > 5


@pytest.mark.parametrize("name", list(CASES))
def test_fp32_mixed_step_from_zero_vs_oracle(rt, O, name):
    gs, og, P, ctx, steps, nDir, r2 = setup(rt, O, name, rt.FP32_MIXED)
    print(f"[exact] FP32-mixed {name} {gs.nx}x{gs.ny} D={nDir} steps={steps} fused={ctx.uses_fused()}", flush=True)
    Iref, mref, _, _, _ = P.run(np.zeros(og.N * nDir), steps)
    ctx.step(steps)
    assert rel_l2(ctx.download_I(), Iref) <= TOL32
    act = active_mask(gs, r2)
    m = ctx.moments()
    for X in range(3):
        assert rel_l2(m[X][act], mref[X][act]) <= TOL32, X


def test_fp32_mixed_d64_curved_kernel_runs_fused(rt, O):
    """The headline form: FP32-mixed, 64 directions, Schwarzschild, fused."""
    gs, og, P, ctx, steps, nDir, r2 = setup(rt, O, "schwarzschild_1024sq_d64", rt.FP32_MIXED)
    assert nDir == 64 and ctx.uses_fused()


def test_kerrschild_excision_inside_horizon_gpu_tables_vs_oracle(rt, O):
    """The Kerr-Schild workload's tables built on the GPU (RK45 tol 1e-10,
    cells inside r < 1.6 excised, so backward geodesics from 1.6 < r < 2
    start inside the horizon) against the oracle's, and the FP64 step on
    them from I = 0 against the oracle on its own tables."""
    n, L, kind, nDir, r2, tol, steps = CASES["kerrschild_200sq_d32"]
    gs = rt.GridSpec(n, n, (-L, -L), (L, L), 0.5)
    og = O.make_grid(n, n, (-L, -L), (L, L), 0.5)
    f = rt.workloads.fields("kerrschild_200sq_d32", gs)
    M = O.metric(og, kind)
    ref = O.geodesic_tables(og, M, r2, tol)
    ctx = rt.Context(gs, nDir, rt.CURVED_FIXED, rt.FP64, r2)
    w, phi, cx, cy, Nc, Wc = rt.build_stencil(nDir)
    ctx.set_stencil(w, phi, cx, cy, Nc, Wc)
    block, alpha, t11, t12, t21, t22 = rt.build_metric(kind, 1.0, 0.0, gs)
    ctx.set_geometry(alpha, t11, t12, t21, t22)
    ctx.set_streaming_stencil(*rt.build_stencil(5, tables=False))
    ctx.build_fourier_tables(block, tol)
    got = ctx.get_fourier_tables()
    for a, b in zip(got, ref):
        assert rel_l2(a, b) <= 1e-12
    ctx.set_fields(**f)
    ctx.upload_I(np.zeros(gs.N * nDir))
    print(f"[exact] Kerr-Schild GPU tables {n}x{n} fused={ctx.uses_fused()}", flush=True)
    P = O.Problem(og, nDir, 2, M, bh_r2=r2, tables=ref, fields=f)
    Iref, mref, mx, sm, _ = P.run(np.zeros(og.N * nDir), steps)
    ctx.reset_iteration_stats()
    ctx.step(steps)
    assert rel_l2(ctx.download_I(), Iref) <= TOL64
    assert ctx.iteration_stats() == (mx, sm)
