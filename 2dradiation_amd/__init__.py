"""MI355X (gfx950) stream-collide library for 2dRadiation-style GR radiation transport.

Python access to the C ABI declared in include/rt2d.h (ctypes).  The compute
path is the HIP library lib/librt2d.so; importing this package fails loudly
when it has not been built, and nothing here falls back to a CPU path.

The package directory name starts with a digit, so import it with
``importlib.import_module("2dradiation_amd")``.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "librt2d.so")

RT_OK = 0
RT_ERR_ARG, RT_ERR_STATE, RT_ERR_HIP, RT_ERR_OOM, RT_ERR_UNSUPPORTED, RT_ERR_NONFINITE = -1, -2, -3, -4, -5, -6
FP64, FP32_MIXED = 0, 1
FLAT_FIXED, FLAT_ADAPTIVE, CURVED_FIXED, CURVED_ADAPTIVE, GEODESIC_FIXED = 0, 1, 2, 3, 4
MINKOWSKI, SCHWARZSCHILD, KERRSCHILD = 0, 1, 2
NMETRIC = 29
NOALLOC_SCRATCH_DOUBLES = 1 << 17  # RT_NOALLOC_SCRATCH_DOUBLES


class RtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rt2d error {code}: {msg}")
        self.code = code


class Desc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("dx", C.c_double), ("dy", C.c_double),
                ("dt", C.c_double), ("startx", C.c_double), ("starty", C.c_double),
                ("nDir", C.c_int32), ("nCoeff", C.c_int32), ("precision", C.c_int32),
                ("streaming", C.c_int32), ("excision_r2", C.c_double), ("device", C.c_int32)]


_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    """Compile lib/librt2d.so for gfx950 with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", os.path.join(_HERE, "csrc"), f"-j{min(8, os.cpu_count() or 1)}"],
                   check=True)


def lib():
    """The loaded HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(LIB_PATH)
    vp, i32, d = C.c_void_p, C.c_int32, C.c_double
    L.rt_abi_version.restype = C.c_int
    L.rt_last_error.restype = C.c_char_p
    L.rt_last_error.argtypes = [vp]
    L.rt_create.argtypes = [C.POINTER(Desc), C.POINTER(vp)]
    L.rt_destroy.argtypes = [vp]
    L.rt_destroy.restype = None
    L.rt_synchronize.argtypes = [vp]
    L.rt_set_stencil.argtypes = [vp, _dp, _dp, _dp, _dp, i32, vp, vp]
    L.rt_set_streaming_stencil.argtypes = [vp, _dp, _dp, _dp, _dp]
    L.rt_set_geometry.argtypes = [vp, _dp, _dp, _dp, _dp, _dp]
    L.rt_set_fourier_tables.argtypes = [vp, _dp, _dp, _dp, _dp, _dp]
    L.rt_build_fourier_tables.argtypes = [vp, _dp, d]
    L.rt_get_fourier_tables.argtypes = [vp, _dp, _dp, _dp, _dp, _dp]
    L.rt_set_fields.argtypes = [vp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.rt_upload_I.argtypes = [vp, _dp]
    L.rt_download_I.argtypes = [vp, _dp]
    L.rt_get_moments.argtypes = [vp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.rt_build_flux_tables.argtypes = [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.POINTER(C.c_int32)]
    L.rt_set_sources.argtypes = [vp, C.c_void_p, C.c_void_p]
    L.rt_rebuild_fourier_tables.argtypes = [vp, C.c_double]
    L.rt_set_lab_frame.argtypes = [vp, C.POINTER(C.POINTER(C.c_double)), _dp, _dp, _dp]
    L.rt_get_moments_lf.argtypes = [vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.rt_iteration_stats.argtypes = [vp, C.POINTER(i32), C.POINTER(d)]
    L.rt_reset_iteration_stats.argtypes = [vp]
    L.rt_step.argtypes = [vp, i32, i32]
    L.rt_stream.argtypes = [vp]
    L.rt_compute_moments.argtypes = [vp]
    L.rt_collide.argtypes = [vp]
    L.rt_uses_fused_path.argtypes = [vp, C.POINTER(i32)]
    L.rt_time_steps.argtypes = [vp, i32, C.POINTER(C.c_float)]
    L.rt_get_iteration_counts.argtypes = [vp, _ip]
    L.rt_forward_checkpointed.argtypes = [vp, i32, i32]
    L.rt_sample_outgoing.argtypes = [vp, i32, _dp, _dp, _dp]
    L.rt_sample_outgoing_noalloc.argtypes = [vp, i32, _dp, _dp, _dp]
    L.rt_adjoint.argtypes = [vp, i32, _dp, _dp, _dp, _dp, _dp]
    L.rt_batch_create.argtypes = [C.POINTER(vp), i32, C.POINTER(vp)]
    L.rt_batch_destroy.argtypes = [vp]
    L.rt_batch_last_error.restype = C.c_char_p
    L.rt_batch_last_error.argtypes = [vp]
    L.rt_batch_step.argtypes = [vp, i32]
    L.rt_batch_time_steps.argtypes = [vp, i32, C.POINTER(C.c_float)]
    L.rt_batch_autotune.argtypes = [vp, i32, vp]
    L.rt_batch_groups.argtypes = [vp, C.POINTER(i32)]
    L.rt_batch_uses_queue.argtypes = [vp, C.POINTER(i32)]
    L.rt_batch_set_queue.argtypes = [vp, i32]
    L.rt_build_stencil.argtypes = [i32, _dp, _dp, _dp, _dp, vp, vp]
    L.rt_build_stencil_ghost.argtypes = [i32, i32, _dp, _dp, _dp, _dp, vp, vp]
    L.rt_set_rotation.argtypes = [vp, _dp, vp]
    L.rt_get_rotation.argtypes = [vp, vp, vp]
    L.rt_update_rotation.argtypes = [vp]
    L.rt_set_source_rotation.argtypes = [vp, _dp]
    L.rt_set_moments.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.rt_build_metric.argtypes = [i32, d, d, i32, i32, d, d, d, d, vp, vp, vp, vp, vp, vp]
    _lib = L
    return L


def _check(rc, ctx=None):
    if rc != RT_OK:
        msg = lib().rt_last_error(ctx)
        raise RtError(rc, msg.decode() if msg else "")


def _ptr(a):
    return None if a is None else a.ctypes.data


def build_stencil(n, tables=True):
    """Stencil(n, 0): (w, phi, cx, cy[, Nc, Wc]) from the library's C++ Stencil class."""
    w, phi, cx, cy = (np.zeros(n) for _ in range(4))
    Nc = np.zeros(40 * n, np.int32) if tables else None
    Wc = np.zeros(40 * n) if tables else None
    _check(lib().rt_build_stencil(n, w, phi, cx, cy, _ptr(Nc), _ptr(Wc)))
    return (w, phi, cx, cy, Nc, Wc) if tables else (w, phi, cx, cy)


def build_stencil_ghost(n_order, n_ghost):
    """Stencil(n_order, n_ghost) with ghost directions, sorted by angle:
    (w, phi, cx, cy, Nc, Wc); the cubic tables have 10*n_order fine angles."""
    n = n_order + n_ghost
    w, phi, cx, cy = (np.zeros(n) for _ in range(4))
    Nc = np.zeros(40 * n_order, np.int32)
    Wc = np.zeros(40 * n_order)
    _check(lib().rt_build_stencil_ghost(n_order, n_ghost, w, phi, cx, cy, Nc.ctypes.data, Wc.ctypes.data))
    return w, phi, cx, cy, Nc, Wc


def build_flux_tables(n):
    """Stencil(n).fluxToSigmaTable / fluxToNormalizationTable: (relative flux,
    sigma, log-normalisation) arrays from the library's C++ Stencil class."""
    count = C.c_int32()
    _check(lib().rt_build_flux_tables(n, 0, None, None, None, C.byref(count)))
    f, s_, z = (np.zeros(count.value) for _ in range(3))
    _check(lib().rt_build_flux_tables(n, count.value, f.ctypes.data, s_.ctypes.data, z.ctypes.data, C.byref(count)))
    return f, s_, z


def build_metric(kind, m, a, grid):
    """Metric arrays from the library's C++ spacetimes: (block[N,29], alpha, t11, t12, t21, t22)."""
    N = grid.nx * grid.ny
    block = np.zeros(N * NMETRIC)
    alpha, t11, t12, t21, t22 = (np.zeros(N) for _ in range(5))
    _check(lib().rt_build_metric(kind, m, a, grid.nx, grid.ny, grid.dx, grid.dy, grid.startx, grid.starty,
                                 block.ctypes.data, alpha.ctypes.data, t11.ctypes.data, t12.ctypes.data,
                                 t21.ctypes.data, t22.ctypes.data))
    return block, alpha, t11, t12, t21, t22


class GridSpec:
    """Grid(nx_, ny_, start, end) + SetCFL(cfl) of the reference (src/Grid.cpp:4-37)."""

    def __init__(self, nx_, ny_, start, end, cfl=1.0):
        self.nx, self.ny = nx_ + 2, ny_ + 2
        self.dx = (end[0] - start[0]) / (self.nx - 1.0 - 2.0)
        self.dy = (end[1] - start[1]) / (self.ny - 1.0 - 2.0)
        self.dt = cfl * min(self.dx, self.dy)
        self.startx, self.starty = start[0] - self.dx, start[1] - self.dy
        self.endx, self.endy = end[0] + self.dx, end[1] + self.dy

    @property
    def N(self):
        return self.nx * self.ny

    def coords(self):
        i = np.arange(self.N) % self.nx
        j = np.arange(self.N) // self.nx
        return self.startx + i * self.dx, self.starty + j * self.dy


class Context:
    """One problem on one GPU (wraps rt_ctx)."""

    def __init__(self, grid, nDir, streaming=FLAT_FIXED, precision=FP64, excision_r2=-1.0, nCoeff=5, device=0):
        L = lib()
        self.grid, self.nDir, self.nCoeff = grid, nDir, nCoeff
        d = Desc(grid.nx, grid.ny, grid.dx, grid.dy, grid.dt, grid.startx, grid.starty, nDir, nCoeff,
                 precision, streaming, excision_r2, device)
        h = C.c_void_p()
        _check(L.rt_create(C.byref(d), C.byref(h)))
        self.h = h

    def _c(self, rc):
        _check(rc, self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().rt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stencil(self, w, phi, cx, cy, Nc=None, Wc=None):
        nGrid = 0 if Nc is None else len(Nc) // 4
        self._c(lib().rt_set_stencil(self.h, w, phi, cx, cy, nGrid, _ptr(Nc), _ptr(Wc)))

    def set_streaming_stencil(self, w, phi, cx, cy):
        self._c(lib().rt_set_streaming_stencil(self.h, w, phi, cx, cy))

    def set_geometry(self, alpha, t11, t12, t21, t22):
        self._c(lib().rt_set_geometry(self.h, *(np.ascontiguousarray(a, dtype=np.float64)
                                                for a in (alpha, t11, t12, t21, t22))))

    def set_fourier_tables(self, cS, cX, cY, cCx, cCy):
        self._c(lib().rt_set_fourier_tables(self.h, *(np.ascontiguousarray(a, dtype=np.float64)
                                                      for a in (cS, cX, cY, cCx, cCy))))

    def build_fourier_tables(self, block, tol=1e-4):
        self._c(lib().rt_build_fourier_tables(self.h, np.ascontiguousarray(block, dtype=np.float64), tol))

    def get_fourier_tables(self):
        n = self.grid.N * self.nCoeff
        arrs = [np.zeros(n) for _ in range(5)]
        self._c(lib().rt_get_fourier_tables(self.h, *arrs))
        return arrs

    def set_fields(self, eta=None, kappa0=None, kappa1=None, kappaA=None, ux=None, uy=None):
        N = self.grid.N
        f = [np.ascontiguousarray(np.zeros(N) if a is None else a, dtype=np.float64)
             for a in (eta, kappa0, kappa1, kappaA, ux, uy)]
        self._c(lib().rt_set_fields(self.h, *f))

    def upload_I(self, I):
        self._c(lib().rt_upload_I(self.h, np.ascontiguousarray(I, dtype=np.float64)))

    def download_I(self):
        I = np.zeros(self.grid.N * self.nDir)
        self._c(lib().rt_download_I(self.h, I))
        return I

    def moments(self):
        arrs = [np.zeros(self.grid.N) for _ in range(6)]
        self._c(lib().rt_get_moments(self.h, *arrs))
        return np.stack(arrs)

    def rebuild_fourier_tables(self, tol=1e-4):
        """RK45 tables again from the metric block kept on the device."""
        self._c(lib().rt_rebuild_fourier_tables(self.h, tol))

    def set_sources(self, is_source, I):
        """Source cells re-imposed before every step (keepSourceNodesActive):
        is_source [N] (nonzero marks a cell), I host AoS [N*nDir]; None
        removes them."""
        if is_source is None:
            self._c(lib().rt_set_sources(self.h, None, None))
            return
        m = np.ascontiguousarray(is_source, dtype=np.uint8)
        v = np.ascontiguousarray(I, dtype=np.float64)
        self._c(lib().rt_set_sources(self.h, m.ctypes.data, v.ctypes.data))

    def set_rotation(self, angle, angle_new=None):
        """rotationAngle and rotationAngleNew ([N] each; None: 0) of an
        adaptive context."""
        a = np.ascontiguousarray(angle, dtype=np.float64)
        b = None if angle_new is None else np.ascontiguousarray(angle_new, dtype=np.float64)
        self._c(lib().rt_set_rotation(self.h, a, _ptr(b)))

    def rotation(self):
        """(rotationAngle, rotationAngleNew) as held on the device."""
        a, b = np.zeros(self.grid.N), np.zeros(self.grid.N)
        self._c(lib().rt_get_rotation(self.h, a.ctypes.data, b.ctypes.data))
        return a, b

    def update_rotation(self):
        self._c(lib().rt_update_rotation(self.h))

    def set_source_rotation(self, angle):
        """Angle re-imposed on each source cell before every step ([N], NaN: none)."""
        self._c(lib().rt_set_source_rotation(self.h, np.ascontiguousarray(angle, dtype=np.float64)))

    def set_moments(self, mom):
        """E, Fx, Fy, Pxx, Pxy, Pyy ([6][N]; a None row is left as it is)."""
        rows = [None if m is None else np.ascontiguousarray(m, dtype=np.float64) for m in mom]
        self._c(lib().rt_set_moments(self.h, *[_ptr(r) for r in rows]))

    def set_lab_frame(self, tetrad, gamma):
        """Full tetrad e[a][A] as [9][N] (00, 01, ..., 22) and gamma_ll as
        [3][N] (11, 12, 22) for moments_lf (Metric::tetradAB_ul,
        gammaIJ_ll)."""
        t = [np.ascontiguousarray(tetrad[k], dtype=np.float64) for k in range(9)]
        g = [np.ascontiguousarray(gamma[k], dtype=np.float64) for k in range(3)]
        self._frame = (t, g)  # the library copies; kept for the caller's inspection only
        arr = (C.POINTER(C.c_double) * 9)(*[x.ctypes.data_as(C.POINTER(C.c_double)) for x in t])
        self._c(lib().rt_set_lab_frame(self.h, arr, *g))

    def moments_lf(self):
        """Lab-frame E, Fx, Fy, Pxx, Pxy, Pyy, F of the moments last computed
        (Radiation::ComputeMomentsLF, src/Radiation.cpp:270-301), [7][N]."""
        arrs = [np.zeros(self.grid.N) for _ in range(7)]
        self._c(lib().rt_get_moments_lf(self.h, *arrs))
        return np.stack(arrs)

    def iteration_stats(self):
        mx, sm = C.c_int32(), C.c_double()
        self._c(lib().rt_iteration_stats(self.h, C.byref(mx), C.byref(sm)))
        return mx.value, sm.value

    def reset_iteration_stats(self):
        self._c(lib().rt_reset_iteration_stats(self.h))

    def step(self, n=1, want_moments=True):
        self._c(lib().rt_step(self.h, n, 1 if want_moments else 0))

    def stream(self):
        self._c(lib().rt_stream(self.h))

    def compute_moments(self):
        self._c(lib().rt_compute_moments(self.h))

    def collide(self):
        self._c(lib().rt_collide(self.h))

    def synchronize(self):
        self._c(lib().rt_synchronize(self.h))

    def uses_fused(self):
        f = C.c_int32()
        self._c(lib().rt_uses_fused_path(self.h, C.byref(f)))
        return bool(f.value)

    def iteration_counts(self):
        """Per-cell lambda-cycle counts of the last collide that records them:
        the reference-order collide, and every step of an FP64 context with a
        fixed stencil (its exact collide)."""
        cnt = np.zeros(self.grid.N, dtype=np.int32)
        self._c(lib().rt_get_iteration_counts(self.h, cnt))
        return cnt

    def forward_checkpointed(self, n, every):
        """n steps keeping a checkpoint every `every` steps for adjoint().
        Each call replaces the checkpoints and bumps checkpoint_generation."""
        self.checkpoint_generation = getattr(self, "checkpoint_generation", 0) + 1
        self._c(lib().rt_forward_checkpointed(self.h, n, every))

    def sample_outgoing(self, x, y):
        """Bilinear samples [npts, nDir] of the current state at (x, y), outward directions only."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros(len(x) * self.nDir)
        self._c(lib().rt_sample_outgoing(self.h, len(x), x, y, out))
        return out.reshape(len(x), self.nDir)

    def sample_outgoing_noalloc(self, x, y):
        """sample_outgoing staged in the library's fixed per-GPU array (no
        device allocation); 2*npts + npts*nDir <= NOALLOC_SCRATCH_DOUBLES."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros(len(x) * self.nDir)
        self._c(lib().rt_sample_outgoing_noalloc(self.h, len(x), x, y, out))
        return out.reshape(len(x), self.nDir)

    def adjoint(self, x, y, dL_dout):
        """dL/d eta and dL/d kappa0 ([N] each) for dL/d(sample_outgoing(x, y)) after
        the last forward_checkpointed run."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        g = np.ascontiguousarray(dL_dout, dtype=np.float64).reshape(-1)
        ge = np.zeros(self.grid.N)
        gk = np.zeros(self.grid.N)
        self._c(lib().rt_adjoint(self.h, len(x), x, y, g, ge, gk))
        return ge, gk

    def time_steps(self, n):
        ms = C.c_float()
        self._c(lib().rt_time_steps(self.h, n, C.byref(ms)))
        return ms.value


class Batch:
    """Many Contexts advanced together (include/rt2d.h: rt_batch_*): one
    grouped launch per (precision, nDir, streaming) class per step, two steps
    captured in a HIP graph.  The Contexts must stay alive while it exists."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        arr = (C.c_void_p * len(self.contexts))(*[c.h for c in self.contexts])
        self.h = C.c_void_p()
        rc = lib().rt_batch_create(arr, len(self.contexts), C.byref(self.h))
        if rc != RT_OK:
            msg = lib().rt_last_error(self.contexts[0].h)
            raise RtError(rc, msg.decode() if msg else "")

    def _c(self, rc):
        if rc != RT_OK:
            raise RtError(rc, lib().rt_batch_last_error(self.h).decode())

    def groups(self):
        n = C.c_int32()
        self._c(lib().rt_batch_groups(self.h, C.byref(n)))
        return n.value

    def step(self, n):
        self._c(lib().rt_batch_step(self.h, n))

    def time_steps(self, n):
        ms = C.c_float()
        self._c(lib().rt_batch_time_steps(self.h, n, C.byref(ms)))
        return ms.value

    def uses_queue(self):
        """True when the FP32 members step in one persistent queue launch."""
        q = C.c_int32()
        self._c(lib().rt_batch_uses_queue(self.h, C.byref(q)))
        return bool(q.value)

    def set_queue(self, on):
        """Step the FP32 members in one queue launch (True) or one launch per class."""
        self._c(lib().rt_batch_set_queue(self.h, 1 if on else 0))

    def autotune(self, trial_steps=2):
        """Rows per workgroup chosen per class (0 = default rule)."""
        chosen = np.zeros(self.groups(), dtype=np.int32)
        self._c(lib().rt_batch_autotune(self.h, trial_steps, chosen.ctypes.data))
        return chosen

    def close(self):
        if getattr(self, "h", None):
            lib().rt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


from . import workloads  # noqa: E402,F401
