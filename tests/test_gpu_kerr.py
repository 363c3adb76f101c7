"""Kerr-Schild with spin (a != 0) and the per-step table
rebuild (rt_rebuild_fourier_tables, Config::updateFourierHarmonics,
src/Radiation.cpp:790-791).  The metric with spin follows
KerrSchild::MetricFunction (src/Spacetimes.cpp:87-111) in both the library's
C++ classes and the oracle; bars as for the other curved parity tests."""
import numpy as np
import pytest

from helpers import rel_l2, oracle_problem, lib_context, grids

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("spin", [0.6, -0.9])
def test_kerr_spin_tables_and_stream(rt, O, spin):
    print(f"[kerr] a = {spin}", flush=True)
    gs, og = grids(rt, O, 60, 44, (-4, -4), (4, 4))
    bh = 2.1 * 2.1
    x, y = gs.coords()
    r = np.sqrt(x * x + y * y)
    f = {"eta": ((r > 2.6) & (r < 3.4)).astype(float)}
    P, M, tabs = oracle_problem(O, og, 16, 2, f, bh_r2=bh, a=spin)
    ctx, _ = lib_context(rt, gs, 16, 2, f, rt.FP64, bh_r2=bh, a=spin)
    got = ctx.get_fourier_tables()
    for g_, t_ in zip(got, tabs):
        assert rel_l2(g_, t_) <= 1e-12
    # spin breaks the mirror symmetry y -> -y of the departure points
    if spin != 0:
        cX = tabs[1].reshape(og.N, 5)
        assert np.abs(cX[:, 2]).max() > 1e-6
    Iref = P.run(np.zeros(og.N * 16), 10)[0]
    assert ctx.uses_fused()
    ctx.step(10)
    assert rel_l2(ctx.download_I(), Iref) <= 1e-12
    # the per-step rebuild from the device copy of the metric gives the same tables
    ctx.rebuild_fourier_tables(1e-4)
    for a_, b_ in zip(ctx.get_fourier_tables(), got):
        assert np.array_equal(a_, b_)
    ctx.close()
