"""Batched stepping of independent problems (replayed from a HIP graph: the
FP32 members in one persistent launch draining a (problem, tile) queue, or
one grouped launch per class) against the same problems stepped one by
one."""
import numpy as np
import pytest


SPECS = [  # size, nDir, spacetime, precision
    (24, 16, "minkowski", "fp64"),
    (37, 20, "schwarzschild", "fp64"),
    (70, 32, "schwarzschild", "fp32"),
    (45, 16, "minkowski", "fp32"),
    (130, 24, "minkowski", "fp32"),
    (52, 32, "schwarzschild", "fp32"),
    (66, 48, "schwarzschild", "fp64"),
    (58, 48, "schwarzschild", "fp32"),
]
# FP32 members at 32 and 48 directions are small enough that stepping them
# alone takes the 8-wavefront form of the fused kernel, while the batch runs
# 4 wavefronts: the bitwise comparisons below cover that both agree.


def members(rt):
    out = []
    for k, (size, nd, st, prec) in enumerate(SPECS):
        L = 6.0 if st == "schwarzschild" else 1.0
        spec = dict(index=k, seed=77000 + k, size=size, nDir=nd, L=L, spacetime=st,
                    excision=max(2.1, 2.0 + 4.0 * L / (size - 1)))
        ctx, grid = rt.workloads.make_batch_member(spec, rt.FP64 if prec == "fp64" else rt.FP32_MIXED)
        rng = np.random.default_rng(k)
        ctx.upload_I(rng.uniform(0.0, 1.0, grid.N * nd))
        out.append(ctx)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("queue", [True, False])
def test_batch_matches_individual_steps(rt, queue):
    print(f"[batch] building members, queue={queue}", flush=True)
    A = members(rt)
    B = members(rt)
    b = rt.Batch(A)
    assert b.groups() == 7
    assert b.uses_queue()  # the default
    b.set_queue(queue)
    assert b.uses_queue() == queue
    for n in (7, 4):  # odd then even: both parities of the graph
        b.step(n)
        for c in B:
            c.step(n, want_moments=False)
    for a, c in zip(A, B):
        assert np.array_equal(a.download_I(), c.download_I())
    b.close()


@pytest.mark.gpu
def test_batch_autotune_keeps_results(rt):
    A = members(rt)
    B = members(rt)
    b = rt.Batch(A)
    chosen = b.autotune(1)
    assert len(chosen) == b.groups()
    b.step(5)
    for c in B:
        c.step(5, want_moments=False)
    for a, c in zip(A, B):
        assert np.array_equal(a.download_I(), c.download_I())
    # members stay usable on their own
    A[0].step(2, want_moments=False)
    B[0].step(2, want_moments=False)
    assert np.array_equal(A[0].download_I(), B[0].download_I())
    b.close()


@pytest.mark.gpu
def test_batch_rejects_uncovered_member(rt):
    spec = dict(index=0, seed=1, size=20, nDir=12, L=1.0, spacetime="minkowski")
    ctx, _ = rt.workloads.make_batch_member(spec, rt.FP64)  # nDir 12 is not a fused size
    with pytest.raises(rt.RtError) as e:
        rt.Batch([ctx])
    assert e.value.code == rt.RT_ERR_UNSUPPORTED


@pytest.mark.gpu
def test_full_batched_workload_matches_sequential(rt):
    """All 64 problems of the batched workload (64^2..512^2 cells, 16..64
    directions, Minkowski and Schwarzschild, FP32-mixed), stepped together
    and one by one: bit-identical states."""
    specs = rt.workloads.batch_specs(64)
    print(f"[batch] building {len(specs)} members", flush=True)
    A, B = [], []
    for sp in specs:
        A.append(rt.workloads.make_batch_member(sp, rt.FP32_MIXED)[0])
        B.append(rt.workloads.make_batch_member(sp, rt.FP32_MIXED)[0])
    assert 64 in {sp["nDir"] for sp in specs}
    b = rt.Batch(A)
    assert b.uses_queue()
    for n in (3, 4):
        b.step(n)
        for c in B:
            c.step(n, want_moments=False)
    for k, (a, c) in enumerate(zip(A, B)):
        assert np.array_equal(a.download_I(), c.download_I()), k
    # and one launch per class for the next steps
    b.set_queue(False)
    b.step(3)
    for c in B:
        c.step(3, want_moments=False)
    for k, (a, c) in enumerate(zip(A, B)):
        assert np.array_equal(a.download_I(), c.download_I()), k
    b.close()
    for c in A + B:
        c.close()
