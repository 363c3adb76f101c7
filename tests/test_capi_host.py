"""C ABI checks that need no GPU: symbol exports, host-side set-up helpers
(against the oracle, bit for bit) and error reporting."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(rt_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ("rt_create", "rt_destroy", "rt_last_error", "rt_set_stencil", "rt_set_fields",
                     "rt_upload_I", "rt_download_I", "rt_step", "rt_stream", "rt_collide",
                     "rt_compute_moments", "rt_build_fourier_tables", "rt_get_moments"):
        assert required in names


def test_library_exports_every_declared_symbol(rt):
    L = rt.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert L.rt_abi_version() == 1


@pytest.mark.parametrize("n", [5, 16, 20, 24, 32, 48, 64])
def test_stencil_builder_matches_oracle(rt, O, n):
    w, phi, cx, cy, Nc, Wc = rt.build_stencil(n)
    ow, ophi, ocx, ocy = O.stencil(n)
    oNc, oWc = O.interp_grid(ophi)
    for a, b in ((w, ow), (phi, ophi), (cx, ocx), (cy, ocy), (Nc, oNc), (Wc, oWc)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("n_order,n_ghost", [(4, 2), (16, 4), (32, 8), (40, 10), (80, 20), (160, 40)])
def test_ghost_stencil_builder_matches_oracle(rt, O, n_order, n_ghost):
    w, phi, cx, cy, Nc, Wc = rt.build_stencil_ghost(n_order, n_ghost)
    ow, ophi, ocx, ocy = O.stencil_ghost(n_order, n_ghost)
    oNc, oWc = O.interp_grid(ophi, 10 * n_order)
    for a, b in ((w, ow), (phi, ophi), (cx, ocx), (cy, ocy), (Nc, oNc), (Wc, oWc)):
        assert np.array_equal(a, b)


def test_ghost_stencil_builder_rejections(rt):
    with pytest.raises(rt.RtError):
        rt.build_stencil_ghost(15, 3)   # odd nOrder
    with pytest.raises(rt.RtError):
        rt.build_stencil_ghost(16, -1)


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_metric_builder_matches_oracle(rt, O, kind):
    gs = rt.GridSpec(37, 29, (-4, -3), (4, 3), 0.5)
    og = O.make_grid(37, 29, (-4, -3), (4, 3), 0.5)
    block, alpha, t11, t12, t21, t22 = rt.build_metric(kind, 1.0, 0.0, gs)
    M = O.metric(og, kind)
    B = block.reshape(-1, rt.NMETRIC)
    order = [O.ALPHA, O.B1U, O.B2U, O.DA1, O.DA2, O.D1B1U, O.D1B2U, O.D2B1U, O.D2B2U, O.K11, O.K12, O.K22,
             O.GA11U, O.GA12U, O.GA22U, O.D1GA11L, O.D1GA12L, O.D1GA22L, O.D2GA11L, O.D2GA12L, O.D2GA22L,
             O.T00 + 4, O.T00 + 5, O.T00 + 7, O.T00 + 8, O.TI00 + 4, O.TI00 + 5, O.TI00 + 7, O.TI00 + 8]
    for q, oq in enumerate(order):
        assert np.array_equal(B[:, q], M[oq], equal_nan=True), q  # NaN only at r=0 for Kerr-Schild
    assert np.array_equal(alpha, M[O.ALPHA], equal_nan=True)
    assert np.array_equal(t21, M[O.T00 + 7], equal_nan=True)


def test_grid_spec_matches_reference_grid(rt, O):
    for args in ((100, 100, (-1, -1), (1, 1), 0.5), (1024, 1024, (-8, -8), (8, 8), 0.5),
                 (7, 3, (0.5, -2), (3, 4), 0.9)):
        gs = rt.GridSpec(*args[:4], cfl=args[4])
        og = O.make_grid(*args)
        assert (gs.nx, gs.ny, gs.dx, gs.dy, gs.dt, gs.startx, gs.starty) == \
               (og.nx, og.ny, og.dx, og.dy, og.dt, og.startx, og.starty)


def test_errors_are_reported_not_fatal(rt):
    L = rt.lib()
    h = C.c_void_p()
    # invalid descriptions are rejected before any device work
    bad = rt.Desc(2, 5, 0.1, 0.1, 0.05, 0.0, 0.0, 16, 5, rt.FP64, rt.FLAT_FIXED, -1.0, 0)
    assert L.rt_create(C.byref(bad), C.byref(h)) == rt.RT_ERR_ARG
    assert b"at least 2" in L.rt_last_error(None)
    bad = rt.Desc(10, 10, 0.1, 0.1, 0.05, 0.0, 0.0, 16, 5, rt.FP64, rt.GEODESIC_FIXED, -1.0, 0)
    assert L.rt_create(C.byref(bad), C.byref(h)) == rt.RT_ERR_UNSUPPORTED
    bad = rt.Desc(10, 10, -0.1, 0.1, 0.05, 0.0, 0.0, 16, 5, rt.FP64, rt.FLAT_FIXED, -1.0, 0)
    assert L.rt_create(C.byref(bad), C.byref(h)) == rt.RT_ERR_ARG
    bad = rt.Desc(10, 10, 0.1, 0.1, 0.05, 0.0, 0.0, 16, 5, 7, rt.FLAT_FIXED, -1.0, 0)
    assert L.rt_create(C.byref(bad), C.byref(h)) == rt.RT_ERR_ARG
    # null context calls fail cleanly
    assert L.rt_step(None, 1, 0) == rt.RT_ERR_ARG
    assert L.rt_synchronize(None) == rt.RT_ERR_ARG
    with pytest.raises(rt.RtError):
        rt.build_stencil(0)
