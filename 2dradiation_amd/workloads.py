"""Seeded synthetic workloads of the stream-collide benchmark and tests.

Every random field comes from a counter-based SplitMix64 generator (Steele,
Lea, Flood 2014) so CPU and GPU runs see bit-identical inputs; normals use the
Box-Muller transform; Gaussian random fields are white noise convolved (FFT,
periodic) with a Gaussian of the stated correlation length and scaled to unit
variance.

Workloads (name: grid, stencil, spacetime, fields, steps):
  minkowski_100sq_d20      100x100 on [-1,1]^2, 20 dirs, flat, CFL 0.5, Gaussian eta (sigma 0.1), kappa 0, 100 steps
  kerrschild_200sq_d32     200x200 on [-4,4]^2, 32 dirs, Kerr-Schild M=1 a=0, excision r<1.6, RK45 tol 1e-10,
                           eta=1 on 3<=r<=3.5, kappa0=exp(-1+0.5 z), 500 steps
  schwarzschild_1024sq_d64 1024x1024 on [-8,8]^2, 64 dirs, Schwarzschild M=1 (excision r<=2.1), eta=max(0,GRF(0.5)),
                           kappa0~U[0,2], kappa1=0.3*U[0,1], 2000 steps
  schwarzschild_256sq_d32  256x256 on [-4,4]^2, 32 dirs, Schwarzschild M=1, kappa0~U[0,1], eta*=max(0,GRF(0.5)), 500 steps
CFL 0.5 everywhere; seed 1000 + workload number.
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed, n, stream=0):
    """n pseudo-random uint64 values, value k = mix(seed * G + stream * 2^40 + k + 1)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64) + np.uint64(stream) * np.uint64(1 << 40)
        z = np.uint64(seed) * _GOLDEN + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n, stream=0):
    """U[0,1) doubles with 53 random bits."""
    return (splitmix64(seed, n, stream) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, n, stream=0):
    """Standard normals by Box-Muller from two uniform streams."""
    u1 = uniform(seed, n, 2 * stream + 1)
    u2 = uniform(seed, n, 2 * stream + 2)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def gaussian_random_field(seed, nx, ny, sigma_x, sigma_y, stream=0):
    """Unit-variance periodic Gaussian random field on an ny x nx array."""
    z = normal(seed, nx * ny, stream).reshape(ny, nx)
    kx = np.fft.fftfreq(nx)
    ky = np.fft.fftfreq(ny)
    gx = np.exp(-0.5 * (2 * np.pi * kx * sigma_x) ** 2)
    gy = np.exp(-0.5 * (2 * np.pi * ky * sigma_y) ** 2)
    f = np.real(np.fft.ifft2(np.fft.fft2(z) * np.outer(gy, gx)))
    f -= f.mean()
    s = f.std()
    return (f / s if s > 0 else f).reshape(-1)


def _grid(spec):
    from . import GridSpec
    return GridSpec(*spec["grid"], cfl=spec["cfl"])


WORKLOADS = {
    "minkowski_100sq_d20": dict(number=1, grid=(100, 100, (-1.0, -1.0), (1.0, 1.0)), cfl=0.5, nDir=20,
                                spacetime="minkowski", steps=100),
    "kerrschild_200sq_d32": dict(number=2, grid=(200, 200, (-4.0, -4.0), (4.0, 4.0)), cfl=0.5, nDir=32,
                                 spacetime="kerrschild", excision=1.6, tol=1e-10, steps=500),
    "schwarzschild_1024sq_d64": dict(number=3, grid=(1024, 1024, (-8.0, -8.0), (8.0, 8.0)), cfl=0.5, nDir=64,
                                     spacetime="schwarzschild", excision=None, tol=1e-4, steps=2000),
    "schwarzschild_256sq_d32": dict(number=4, grid=(256, 256, (-4.0, -4.0), (4.0, 4.0)), cfl=0.5, nDir=32,
                                    spacetime="schwarzschild", excision=None, tol=1e-4, steps=500),
    # the 1024^2 Schwarzschild problem streamed adaptively: Stencil(48,16),
    # 64 directions (48 Fourier + 16 ghost), same fields as the fixed one
    "schwarzschild_1024sq_adaptive_48_16": dict(number=6, fields_like=3, grid=(1024, 1024, (-8.0, -8.0), (8.0, 8.0)),
                                                cfl=0.5, nDir=64, nOrder=48, nGhost=16, streaming="adaptive",
                                                spacetime="schwarzschild", excision=None, tol=1e-4, steps=2000),
}


def fields(name, grid, seed=None):
    """Fluid fields of a workload as a dict of [N] float64 arrays."""
    spec = WORKLOADS[name]
    number = spec.get("fields_like", spec["number"])
    seed = 1000 + number if seed is None else seed
    x, y = grid.coords()
    r = np.sqrt(x * x + y * y)
    N = grid.N
    f = {k: np.zeros(N) for k in ("eta", "kappa0", "kappa1", "kappaA", "ux", "uy")}
    if number == 1:
        f["eta"] = np.exp(-(x * x + y * y) / (2 * 0.1 ** 2))
    elif number == 2:
        f["eta"] = ((r >= 3.0) & (r <= 3.5)).astype(np.float64)
        f["kappa0"] = np.exp(-1.0 + 0.5 * normal(seed, N, 0))
    elif number == 3:
        g = gaussian_random_field(seed, grid.nx, grid.ny, 0.5 / grid.dx, 0.5 / grid.dy, 0)
        f["eta"] = np.maximum(0.0, g)
        f["kappa0"] = 2.0 * uniform(seed, N, 10)
        f["kappa1"] = 0.3 * uniform(seed, N, 11)
    elif number == 4:
        # the point where gradients are taken: a strictly positive emissivity,
        # so no intensity sits on the IntensityAt clamp at 0 (a kink where
        # central differences and the adjoint legitimately disagree)
        g = gaussian_random_field(seed, grid.nx, grid.ny, 0.5 / grid.dx, 0.5 / grid.dy, 0)
        f["eta"] = np.maximum(0.05, 0.5 + 0.25 * g)
        f["kappa0"] = uniform(seed, N, 10)
    return f


def target_eta(name, grid, seed=None):
    """Emissivity eta* of the observer-ring target of the gradient workload:
    a clipped unit-variance GRF (correlation length 0.5) with its own seed."""
    spec = WORKLOADS[name]
    seed = 2000 + spec["number"] if seed is None else seed
    g = gaussian_random_field(seed, grid.nx, grid.ny, 0.5 / grid.dx, 0.5 / grid.dy, 0)
    return np.maximum(0.0, g).reshape(-1)


def spacetime_kind(spec):
    from . import MINKOWSKI, SCHWARZSCHILD, KERRSCHILD
    return {"minkowski": MINKOWSKI, "schwarzschild": SCHWARZSCHILD, "kerrschild": KERRSCHILD}[spec["spacetime"]]


def excision_r2(spec, m=1.0):
    """InsideBH threshold: the reference's 2.1 m unless the workload sets its own radius."""
    if spec["spacetime"] == "minkowski":
        return -1.0
    if spec.get("excision") is None:
        return 2.1 * 2.1 * m * m
    return spec["excision"] ** 2


def make_context(name, precision, overrides=None, seed=None, grid_override=None):
    """Builds a ready-to-step Context for a workload through the library:
    stencil and metric from its C++ classes, geodesic tables on the GPU."""
    from . import (Context, build_stencil, build_stencil_ghost, build_metric, FLAT_FIXED, CURVED_FIXED,
                   FLAT_ADAPTIVE, CURVED_ADAPTIVE)
    spec = dict(WORKLOADS[name])
    if overrides:
        spec.update(overrides)
    grid = _grid(spec) if grid_override is None else grid_override
    kind = spacetime_kind(spec)
    curved = kind != 0
    r2 = excision_r2(spec)
    adaptive = spec.get("streaming") == "adaptive"
    if adaptive:
        streaming = CURVED_ADAPTIVE if curved else FLAT_ADAPTIVE
    else:
        streaming = CURVED_FIXED if curved else FLAT_FIXED
    ctx = Context(grid, spec["nDir"], streaming, precision, r2)
    if adaptive:
        w, phi, cx, cy, Nc, Wc = build_stencil_ghost(spec["nOrder"], spec["nGhost"])
    else:
        w, phi, cx, cy, Nc, Wc = build_stencil(spec["nDir"])
    ctx.set_stencil(w, phi, cx, cy, Nc, Wc)
    block, alpha, t11, t12, t21, t22 = build_metric(kind, 1.0, 0.0, grid)
    ctx.set_geometry(alpha, t11, t12, t21, t22)
    if curved:
        ctx.set_streaming_stencil(*build_stencil(5, tables=False))
        ctx.build_fourier_tables(block, spec["tol"])
    f = fields(name, grid, seed)
    ctx.set_fields(**f)
    ctx.upload_I(np.zeros(grid.N * spec["nDir"]))
    return ctx, dict(spec=spec, grid=grid, fields=f, metric_block=block, alpha=alpha,
                     tetrad=(t11, t12, t21, t22), excision_r2=r2, kind=kind)


def batch_specs(n=64, seed=1005, lo=64, hi=512, dirs=(16, 24, 32, 48, 64)):
    """The batched workload: n independent problems, nx = ny = round(exp(U[ln lo,
    ln hi])), nDir uniform from `dirs`, Minkowski on [-1,1]^2 or Schwarzschild
    (M = 1) on [-6,6]^2 by a fair coin, CFL 0.5.  The Schwarzschild excision
    radius is max(2.1, 2 + 2 dx): on coarse grids the backward geodesics of
    cells just outside r = 2.1 reach the horizon's stiff region and leave the
    cell's neighbourhood, so those cells are excised too."""
    u = uniform(seed, 3 * n, 20).reshape(3, n)
    specs = []
    for k in range(n):
        size = int(round(np.exp(np.log(lo) + u[0, k] * (np.log(hi) - np.log(lo)))))
        nd = int(dirs[min(len(dirs) - 1, int(u[1, k] * len(dirs)))])
        curved = bool(u[2, k] >= 0.5)
        L = 6.0 if curved else 1.0
        dx = 2 * L / (size - 1)
        specs.append(dict(index=k, seed=seed * 1000 + k, size=size, nDir=nd, L=L,
                          spacetime="schwarzschild" if curved else "minkowski",
                          excision=max(2.1, 2.0 + 2.0 * dx) if curved else None))
    return specs


def batch_fields(spec, grid):
    """eta = max(0, GRF) with correlation length L/8, kappa0 ~ lognormal(0, 0.5),
    kappa1 ~ U[0, 0.5] (seeded per problem)."""
    N = grid.N
    corr = spec["L"] / 8.0
    g = gaussian_random_field(spec["seed"], grid.nx, grid.ny, corr / grid.dx, corr / grid.dy, 0)
    f = {k: np.zeros(N) for k in ("eta", "kappa0", "kappa1", "kappaA", "ux", "uy")}
    f["eta"] = np.maximum(0.0, g).reshape(-1)
    f["kappa0"] = np.exp(0.5 * normal(spec["seed"], N, 1))
    f["kappa1"] = 0.5 * uniform(spec["seed"], N, 2)
    return f


def make_batch_member(spec, precision):
    """A ready-to-step Context for one problem of batch_specs."""
    from . import (Context, GridSpec, build_stencil, build_metric, CURVED_FIXED, FLAT_FIXED,
                   MINKOWSKI, SCHWARZSCHILD)
    n, L = spec["size"], spec["L"]
    grid = GridSpec(n, n, (-L, -L), (L, L), cfl=0.5)
    curved = spec["spacetime"] == "schwarzschild"
    r2 = spec.get("excision", 2.1) ** 2 if curved else -1.0
    ctx = Context(grid, spec["nDir"], CURVED_FIXED if curved else FLAT_FIXED, precision, r2)
    w, phi, cx, cy, Nc, Wc = build_stencil(spec["nDir"])
    ctx.set_stencil(w, phi, cx, cy, Nc, Wc)
    block, alpha, t11, t12, t21, t22 = build_metric(SCHWARZSCHILD if curved else MINKOWSKI, 1.0, 0.0, grid)
    ctx.set_geometry(alpha, t11, t12, t21, t22)
    if curved:
        ctx.set_streaming_stencil(*build_stencil(5, tables=False))
        ctx.build_fourier_tables(block, 1e-4)
    ctx.set_fields(**batch_fields(spec, grid))
    ctx.upload_I(np.zeros(grid.N * spec["nDir"]))
    return ctx, grid
