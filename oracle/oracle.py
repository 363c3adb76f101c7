"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, the smoke check and the CPU-baseline leg of bench.py import this
module, and only as the checker or the timed CPU baseline.  The product
library never imports it.  The C sources restate the reference algorithm with
citations (rt_oracle.c).
"""
import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

NM = 56  # number of metric arrays (OR_NM)
# indices into the metric block (rt_oracle.h enum)
(G00L, G01L, G02L, G11L, G12L, G22L, G00U, G01U, G02U, G11U, G12U, G22U,
 ALPHA, B1U, B2U, B1L, B2L, GA11L, GA12L, GA22L, GA11U, GA12U, GA22U,
 DA1, DA2, D1B1U, D1B2U, D2B1U, D2B2U,
 D1GA11L, D1GA12L, D1GA22L, D2GA11L, D2GA12L, D2GA22L, K11, K12, K22) = range(38)
T00 = 38
TI00 = 47

dptr = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
iptr = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("dx", C.c_double), ("dy", C.c_double),
                ("dt", C.c_double), ("startx", C.c_double), ("starty", C.c_double),
                ("endx", C.c_double), ("endy", C.c_double)]

    @property
    def N(self):
        return self.nx * self.ny


class Sim(C.Structure):
    _fields_ = [("grid", Grid), ("nDir", C.c_int32),
                ("w", C.c_void_p), ("phi", C.c_void_p), ("cx", C.c_void_p), ("cy", C.c_void_p),
                ("nGrid", C.c_int32), ("Nc", C.c_void_p), ("Wc", C.c_void_p),
                ("streaming", C.c_int32), ("nCoeff", C.c_int32),
                ("cS", C.c_void_p), ("cX", C.c_void_p), ("cY", C.c_void_p),
                ("cCx", C.c_void_p), ("cCy", C.c_void_p),
                ("alpha", C.c_void_p), ("t11", C.c_void_p), ("t12", C.c_void_p),
                ("t21", C.c_void_p), ("t22", C.c_void_p), ("bh_r2", C.c_double),
                ("eta", C.c_void_p), ("kappa0", C.c_void_p), ("kappa1", C.c_void_p),
                ("kappaA", C.c_void_p), ("ux", C.c_void_p), ("uy", C.c_void_p),
                ("rot", C.c_void_p), ("rotNew", C.c_void_p)]


def build():
    """Compile both oracle libraries with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name in _LIBS:
        return _LIBS[name]
    path = os.path.join(_HERE, "_build", name)
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    d, i32 = C.c_double, C.c_int32
    for f in ("or_mysin", "or_mycos", "or_myatan"):
        getattr(L, f).restype = d
        getattr(L, f).argtypes = [d]
    L.or_myatan2.restype = d
    L.or_myatan2.argtypes = [d, d]
    L.or_phi.restype = d
    L.or_phi.argtypes = [d, d]
    L.or_fourier_basis.restype = d
    L.or_fourier_basis.argtypes = [i32, d]
    L.or_grid_make.argtypes = [i32, i32, d, d, d, d, d, C.POINTER(Grid)]
    L.or_stencil.argtypes = [i32, dptr, dptr, dptr, dptr]
    L.or_interp_grid.argtypes = [i32, dptr, i32, iptr, dptr]
    L.or_metric_function.argtypes = [i32, d, d, d, d, dptr]
    L.or_invert3.argtypes = [dptr, dptr]
    L.or_metric_build.argtypes = [C.POINTER(Grid), i32, d, d, dptr]
    L.or_rk45.restype = d
    L.or_rk45.argtypes = [C.POINTER(Grid), dptr, d, d, d, C.POINTER(d), C.POINTER(d),
                          C.POINTER(d), C.POINTER(d)]
    L.or_geodesic_tables.argtypes = [C.POINTER(Grid), dptr, i32, d, d, dptr, dptr, dptr, dptr, dptr]
    L.or_simulate.restype = C.c_int
    L.or_simulate.argtypes = [C.POINTER(Sim), i32, dptr, C.c_void_p, C.POINTER(i32),
                              C.POINTER(d), C.c_void_p]
    for f in ("or_stream_flat", "or_stream_curved"):
        getattr(L, f).argtypes = [C.POINTER(Sim), dptr, dptr]
    L.or_moments.argtypes = [C.POINTER(Sim), dptr, dptr]
    L.or_moments_lf.argtypes = [C.POINTER(Grid), dptr, dptr, dptr, d, dptr]
    L.or_collide.argtypes = [C.POINTER(Sim), dptr, dptr, dptr, C.c_void_p]
    L.or_fluid_frame.argtypes = [d, d, dptr, dptr]
    L.or_fluid_dir.argtypes = [d, d, d, d, dptr]
    L.or_myacos.restype = d
    L.or_myacos.argtypes = [d]
    L.or_stencil_ghost.restype = C.c_int
    L.or_stencil_ghost.argtypes = [i32, i32, dptr, dptr, dptr, dptr]
    L.or_update_rotation.argtypes = [C.POINTER(Sim), dptr, dptr, dptr]
    for f in ("or_stream_flat_adaptive", "or_stream_curved_adaptive"):
        getattr(L, f).argtypes = [C.POINTER(Sim), dptr, dptr]
    L.or_simulate_adaptive.restype = C.c_int
    L.or_simulate_adaptive.argtypes = [C.POINTER(Sim), i32, dptr, dptr, dptr, dptr, C.POINTER(i32),
                                       C.POINTER(d), C.c_void_p]
    L.or_num_threads.restype = i32
    L.or_set_threads.argtypes = [i32]
    _bind_adjoint(L)
    _LIBS[name] = L
    return L


def _bind_adjoint(L):
    i32 = C.c_int32
    L.or_adj_forward.restype = C.c_int
    L.or_adj_forward.argtypes = [C.POINTER(Sim), i32, dptr, dptr]
    L.or_observer_loss.restype = C.c_double
    L.or_observer_loss.argtypes = [C.POINTER(Sim), dptr, i32, dptr, dptr, dptr, C.c_void_p, C.c_void_p]
    L.or_adjoint_run.restype = C.c_int
    L.or_adjoint_run.argtypes = [C.POINTER(Sim), i32, dptr, dptr, dptr, dptr]


def make_grid(nx_, ny_, start, end, cfl):
    g = Grid()
    lib().or_grid_make(nx_, ny_, start[0], start[1], end[0], end[1], cfl, C.byref(g))
    return g


def stencil(n):
    w, phi, cx, cy = (np.zeros(n) for _ in range(4))
    lib().or_stencil(n, w, phi, cx, cy)
    return w, phi, cx, cy


def stencil_ghost(nOrder, nGhost):
    """Stencil(nOrder, nGhost) sorted by angle: (w, phi, cx, cy), nOrder+nGhost each."""
    n = nOrder + nGhost
    w, phi, cx, cy = (np.zeros(n) for _ in range(4))
    rc = lib().or_stencil_ghost(nOrder, nGhost, w, phi, cx, cy)
    if rc != 0:
        raise ValueError(f"Stencil({nOrder},{nGhost}) rejected ({rc})")
    return w, phi, cx, cy


def interp_grid(phi, nGrid=None):
    n = len(phi)
    nGrid = 10 * n if nGrid is None else nGrid
    Nc = np.zeros(4 * nGrid, np.int32)
    Wc = np.zeros(4 * nGrid)
    lib().or_interp_grid(n, np.ascontiguousarray(phi, dtype=np.float64), nGrid, Nc, Wc)
    return Nc, Wc


def metric(g, kind, m=1.0, a=0.0):
    out = np.zeros(NM * g.N)
    lib().or_metric_build(C.byref(g), kind, m, a, out)
    return out.reshape(NM, g.N)


def moments_lf(g, mom, tetrad, gamma, r2=-1.0):
    """ComputeMomentsLF (src/Radiation.cpp:270-301): mom [6][N], tetrad
    [9][N] (e[a][A] row-major), gamma [3][N] (11, 12, 22) -> [7][N]
    (E, Fx, Fy, Pxx, Pxy, Pyy, F) in the lab frame."""
    out = np.zeros(7 * g.N)
    lib().or_moments_lf(C.byref(g), np.ascontiguousarray(mom, dtype=float).reshape(-1),
                        np.ascontiguousarray(tetrad, dtype=float).reshape(-1),
                        np.ascontiguousarray(gamma, dtype=float).reshape(-1), r2, out)
    return out.reshape(7, g.N)


def geodesic_tables(g, M, bh_r2, tol=1e-4, nF=5):
    arrs = [np.zeros(g.N * nF) for _ in range(5)]
    lib().or_geodesic_tables(C.byref(g), np.ascontiguousarray(M).reshape(-1), nF, bh_r2, tol, *arrs)
    return arrs


class Problem:
    """Holds every array an oracle run needs (keeps them alive for ctypes)."""

    def __init__(self, grid, nDir, streaming, M, bh_r2=-1.0, tables=None, fields=None, fast=False, nGhost=0):
        self.L = lib(fast)
        self.grid = grid
        self.nDir = nDir
        self.nOrder = nDir - nGhost
        if nGhost:
            self.w, self.phi, self.cx, self.cy = stencil_ghost(self.nOrder, nGhost)
        else:
            self.w, self.phi, self.cx, self.cy = stencil(nDir)
        self.nGrid = 10 * self.nOrder
        self.Nc, self.Wc = interp_grid(self.phi, self.nGrid)
        self.M = np.ascontiguousarray(M)
        N = grid.N
        self.alpha = np.ascontiguousarray(self.M[ALPHA])
        self.t = [np.ascontiguousarray(self.M[T00 + q]) for q in (4, 5, 7, 8)]
        self.tables = tables if tables is not None else [np.zeros(N * 5) for _ in range(5)]
        f = fields or {}
        self.fields = [np.ascontiguousarray(f.get(k, np.zeros(N)), dtype=np.float64)
                       for k in ("eta", "kappa0", "kappa1", "kappaA", "ux", "uy")]
        s = Sim()
        s.grid = grid
        s.nDir = nDir
        s.w, s.phi, s.cx, s.cy = (a.ctypes.data for a in (self.w, self.phi, self.cx, self.cy))
        s.nGrid = self.nGrid
        s.Nc, s.Wc = self.Nc.ctypes.data, self.Wc.ctypes.data
        s.streaming = streaming
        s.nCoeff = 5
        s.cS, s.cX, s.cY, s.cCx, s.cCy = (a.ctypes.data for a in self.tables)
        s.alpha = self.alpha.ctypes.data
        s.t11, s.t12, s.t21, s.t22 = (a.ctypes.data for a in self.t)
        s.bh_r2 = bh_r2
        s.eta, s.kappa0, s.kappa1, s.kappaA, s.ux, s.uy = (a.ctypes.data for a in self.fields)
        self.sim = s

    def adj_forward(self, I0, nSteps):
        """Moment-space forward (u = 0) storing every state: [(nSteps+1), N*nDir]."""
        ND = self.grid.N * self.nDir
        states = np.zeros((nSteps + 1) * ND)
        rc = self.L.or_adj_forward(C.byref(self.sim), nSteps, np.ascontiguousarray(I0, dtype=np.float64), states)
        assert rc == 0, "adjoint oracle needs zero fluid velocity"
        return states.reshape(nSteps + 1, ND)

    def observer_loss(self, I, px, py, target, want_grad=False):
        """Ring loss 1/2 sum (obs - target)^2 over outward directions; returns
        (L, obs [npts, nDir], dL/dI or None)."""
        npts = len(px)
        obs = np.zeros(npts * self.nDir)
        g = np.zeros(self.grid.N * self.nDir) if want_grad else None
        L = self.L.or_observer_loss(C.byref(self.sim), np.ascontiguousarray(I, dtype=np.float64), npts,
                                    np.ascontiguousarray(px, dtype=np.float64),
                                    np.ascontiguousarray(py, dtype=np.float64),
                                    np.ascontiguousarray(target, dtype=np.float64).reshape(-1),
                                    obs.ctypes.data, g.ctypes.data if want_grad else None)
        return L, obs.reshape(npts, self.nDir), g

    def adjoint(self, states, lamT):
        """dL/d eta and dL/d kappa0 ([N] each) from stored states and dL/dI_T."""
        nSteps = states.shape[0] - 1
        ge = np.zeros(self.grid.N)
        gk = np.zeros(self.grid.N)
        rc = self.L.or_adjoint_run(C.byref(self.sim), nSteps, np.ascontiguousarray(states.reshape(-1)),
                                   np.ascontiguousarray(lamT, dtype=np.float64), ge, gk)
        assert rc == 0
        return ge, gk

    def moments(self, I):
        """ComputeMomentsIF of a state: [6][N] (E, Fx, Fy, Pxx, Pxy, Pyy)."""
        mom = np.zeros(6 * self.grid.N)
        self.L.or_moments(C.byref(self.sim), np.ascontiguousarray(I, dtype=np.float64), mom)
        return mom.reshape(6, self.grid.N)

    def set_rotation(self, rot, rotNew=None):
        """Rotation angles used by moments(), stream_adaptive() and collide."""
        N = self.grid.N
        self.rot = np.ascontiguousarray(rot, dtype=np.float64).copy()
        self.rotNew = (np.zeros(N) if rotNew is None else np.ascontiguousarray(rotNew, dtype=np.float64).copy())
        self.sim.rot = self.rot.ctypes.data
        self.sim.rotNew = self.rotNew.ctypes.data

    def update_rotation(self, mom, rot):
        """UpdateRotationMatrizes: returns rotationAngleNew ([N]; halo = 0)."""
        out = np.zeros(self.grid.N)
        self.L.or_update_rotation(C.byref(self.sim), np.ascontiguousarray(mom, dtype=np.float64).reshape(-1),
                                  np.ascontiguousarray(rot, dtype=np.float64), out)
        return out

    def stream_adaptive(self, I):
        """One adaptive stream with the angles of set_rotation (no swap)."""
        out = np.zeros(self.grid.N * self.nDir)
        f = self.L.or_stream_curved_adaptive if self.sim.streaming == 3 else self.L.or_stream_flat_adaptive
        f(C.byref(self.sim), np.ascontiguousarray(I, dtype=np.float64), out)
        return out

    def run_adaptive(self, I, nSteps, mom=None, rot=None, rotNew=None, want_counts=False):
        """nSteps adaptive steps; returns (I, moments, rot, rotNew, max, sum, counts)."""
        N = self.grid.N
        I = np.ascontiguousarray(I, dtype=np.float64).copy()
        mom = np.zeros(6 * N) if mom is None else np.ascontiguousarray(mom, dtype=np.float64).reshape(-1).copy()
        rot = np.zeros(N) if rot is None else np.ascontiguousarray(rot, dtype=np.float64).copy()
        rotNew = np.zeros(N) if rotNew is None else np.ascontiguousarray(rotNew, dtype=np.float64).copy()
        mx = C.c_int32(0)
        sm = C.c_double(0)
        counts = np.zeros(nSteps * N, np.int32) if want_counts else None
        rc = self.L.or_simulate_adaptive(C.byref(self.sim), nSteps, I, mom, rot, rotNew, C.byref(mx), C.byref(sm),
                                         counts.ctypes.data if want_counts else None)
        assert rc == 0
        return I, mom.reshape(6, N), rot, rotNew, mx.value, sm.value, counts

    def run(self, I, nSteps, want_counts=False):
        """Advance I (AoS [N*nDir], float64) by nSteps; returns (I, moments, max, sum, counts)."""
        I = np.ascontiguousarray(I, dtype=np.float64).copy()
        N = self.grid.N
        mom = np.zeros(6 * N)
        mx = C.c_int32(0)
        sm = C.c_double(0)
        counts = np.zeros(nSteps * N, np.int32) if want_counts else None
        rc = self.L.or_simulate(C.byref(self.sim), nSteps, I, mom.ctypes.data, C.byref(mx),
                                C.byref(sm), counts.ctypes.data if want_counts else None)
        assert rc == 0
        return I, mom.reshape(6, N), mx.value, sm.value, counts
