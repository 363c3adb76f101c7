"""Shared helpers for the parity tests: build the same problem for the oracle
and for the library, and compare results."""
import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def oracle_problem(O, grid, nDir, kind, fields, bh_r2=-1.0, tol=1e-4, tables=None, fast=False, a=0.0):
    """Oracle Problem for a grid given as an oracle Grid (spin a for Kerr-Schild)."""
    M = O.metric(grid, kind, 1.0, a)
    streaming = 0 if kind == 0 else 2
    if streaming == 2 and tables is None:
        tables = O.geodesic_tables(grid, M, bh_r2, tol)
    return O.Problem(grid, nDir, streaming, M, bh_r2=bh_r2, tables=tables, fields=fields, fast=fast), M, tables


def lib_context(rt, gspec, nDir, kind, fields, precision, bh_r2=-1.0, tol=1e-4, tables=None, a=0.0):
    """Library Context built through its own C++ stencil/metric classes; tables
    either built on the GPU (tables=None) or uploaded."""
    curved = kind != 0
    ctx = rt.Context(gspec, nDir, rt.CURVED_FIXED if curved else rt.FLAT_FIXED, precision, bh_r2)
    w, phi, cx, cy, Nc, Wc = rt.build_stencil(nDir)
    ctx.set_stencil(w, phi, cx, cy, Nc, Wc)
    block, alpha, t11, t12, t21, t22 = rt.build_metric(kind, 1.0, a, gspec)
    ctx.set_geometry(alpha, t11, t12, t21, t22)
    if curved:
        ctx.set_streaming_stencil(*rt.build_stencil(5, tables=False))
        if tables is None:
            ctx.build_fourier_tables(block, tol)
        else:
            ctx.set_fourier_tables(*tables)
    ctx.set_fields(**fields)
    ctx.upload_I(np.zeros(gspec.N * nDir))
    return ctx, block


def grids(rt, O, nx_, ny_, start, end, cfl=0.5):
    return rt.GridSpec(nx_, ny_, start, end, cfl), O.make_grid(nx_, ny_, start, end, cfl)


def active_mask(gspec, bh_r2):
    x, y = gspec.coords()
    i = np.arange(gspec.N) % gspec.nx
    j = np.arange(gspec.N) // gspec.nx
    interior = (i > 0) & (i < gspec.nx - 1) & (j > 0) & (j < gspec.ny - 1)
    inside = (bh_r2 >= 0) & (x * x + y * y <= bh_r2)
    return interior & ~inside
