"""rt_sample_outgoing_noalloc (staging in the library's fixed per-GPU array,
no device allocation) against rt_sample_outgoing on the same state: equal
bit for bit, on two calls in a row (the second with fewer points, so stale
staging from the first would show), on a flat FP64 and a curved FP32-mixed
(alpha^-3-scaled) context, and with two contexts on one GPU sharing the
array in turn.  Sizes beyond the array are rejected with an error."""
import numpy as np
import pytest

from helpers import lib_context

pytestmark = pytest.mark.gpu


def ring(n, r, phase=0.0):
    th = 2 * np.pi * (np.arange(n) + 0.5 + phase) / n
    return r * np.cos(th), r * np.sin(th)


def make(rt, O, kind, n, L, D, precision, seed):
    rng = np.random.default_rng(seed)
    gs = rt.GridSpec(n, n, (-L, -L), (L, L), 0.5)
    g = O.make_grid(n, n, (-L, -L), (L, L), 0.5)
    bh = 2.1 * 2.1 if kind else -1.0
    tables = O.geodesic_tables(g, O.metric(g, kind), bh, 1e-4) if kind else None
    fields = {"eta": rng.uniform(0.2, 1.0, g.N), "kappa0": rng.uniform(0, 1, g.N),
              "kappa1": 0.3 * rng.uniform(0, 1, g.N)}
    ctx, _ = lib_context(rt, gs, D, kind, fields, precision, bh_r2=bh, tables=tables)
    ctx.upload_I(rng.uniform(0.5, 1.5, g.N * D))
    ctx.step(7)
    return ctx


@pytest.mark.parametrize("kind,n,L,D,prec,R", [
    (0, 40, 1.0, 16, "FP64", 0.7),
    (1, 48, 4.0, 64, "FP32_MIXED", 3.5),
])
def test_noalloc_matches_sample_outgoing_twice(rt, O, kind, n, L, D, prec, R):
    print(f"[noalloc] kind={kind} n={n} D={D} {prec}", flush=True)
    ctx = make(rt, O, kind, n, L, D, getattr(rt, prec), 3)
    for npts, phase in ((96, 0.0), (17, 0.3)):
        px, py = ring(npts, R, phase)
        want = ctx.sample_outgoing(px, py)
        got = ctx.sample_outgoing_noalloc(px, py)
        assert got.shape == (npts, D)
        assert np.array_equal(got, want), (npts, np.abs(got - want).max())
        assert np.abs(want).max() > 0
    ctx.close()


def test_noalloc_shared_between_contexts_and_size_limit(rt, O):
    print("[noalloc] two contexts, size limit", flush=True)
    a = make(rt, O, 0, 40, 1.0, 16, rt.FP64, 5)
    b = make(rt, O, 1, 48, 4.0, 32, rt.FP64, 6)
    for ctx, R in ((a, 0.6), (b, 3.2), (a, 0.5)):
        px, py = ring(40, R)
        assert np.array_equal(ctx.sample_outgoing_noalloc(px, py), ctx.sample_outgoing(px, py))
    # the largest size the array holds at 32 directions, then one point more
    cap = rt.NOALLOC_SCRATCH_DOUBLES // (2 + 32)
    px, py = ring(cap, 3.0)
    assert np.array_equal(b.sample_outgoing_noalloc(px, py), b.sample_outgoing(px, py))
    px, py = ring(cap + 1, 3.0)
    with pytest.raises(rt.RtError, match="staging array"):
        b.sample_outgoing_noalloc(px, py)
    # the failed call left the context usable
    px, py = ring(8, 3.0)
    assert np.array_equal(b.sample_outgoing_noalloc(px, py), b.sample_outgoing(px, py))
    a.close()
    b.close()
