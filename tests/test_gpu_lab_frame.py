"""GPU lab-frame moments (rt_get_moments_lf, ComputeMomentsLF of
src/Radiation.cpp:270-301) against the oracle on the same inertial-frame
moments, tetrad and spatial metric.  Both evaluate the reference's loop order
in FP64 without contraction, so the bar is rel-L2 <= 1e-15 (round-off of the
final square root aside, which is correctly rounded on both sides)."""
import numpy as np
import pytest

from helpers import rel_l2, lib_context, grids

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_moments_lf_vs_oracle(rt, O, prec):
    print(f"[lab-frame] Schwarzschild 48x40, 16 dirs, {prec}", flush=True)
    gs, og = grids(rt, O, 48, 40, (-4, -4), (4, 4))
    bh = 2.1 * 2.1
    x, y = gs.coords()
    r = np.sqrt(x * x + y * y)
    fields = {"eta": ((r > 2.5) & (r < 3.5)).astype(float), "kappa0": np.full(gs.N, 0.3),
              "kappa1": np.full(gs.N, 0.2)}
    precision = rt.FP64 if prec == "fp64" else rt.FP32_MIXED
    ctx, _ = lib_context(rt, gs, 16, 2, fields, precision, bh_r2=bh)
    ctx.step(6, want_moments=False)
    ctx.compute_moments()
    mom = ctx.moments()
    rng = np.random.default_rng(11)
    tetrad = np.eye(3).reshape(9, 1) + 0.3 * rng.standard_normal((9, gs.N))
    gamma = np.stack([1 + rng.random(gs.N), 0.2 * rng.standard_normal(gs.N), 1 + rng.random(gs.N)])
    inside = x * x + y * y <= bh
    tetrad[:, inside] = np.nan  # never read inside the excision radius
    ctx.set_lab_frame(tetrad, gamma)
    lf = ctx.moments_lf()
    want = O.moments_lf(og, mom, tetrad, gamma, bh)
    assert np.all(lf[:, inside] == 0)
    assert np.all(np.isfinite(lf))
    for k in range(7):
        assert rel_l2(lf[k], want[k]) <= 1e-15, k
    ctx.close()


def test_moments_lf_requires_frame(rt):
    gs = rt.GridSpec(8, 8, (-1, -1), (1, 1), 0.5)
    ctx = rt.Context(gs, 16, rt.FLAT_FIXED, rt.FP64, -1.0)
    with pytest.raises(RuntimeError, match="rt_set_lab_frame"):
        ctx.moments_lf()
    ctx.close()
