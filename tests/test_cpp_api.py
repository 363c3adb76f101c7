"""The C++ classes with the reference's API (Grid, Stencil, Metric,
Radiation) driven by a reference-style user program (tests/cpp), checked
against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from helpers import rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "_build", "radiation_driver")


def test_driver_is_built():
    assert os.path.exists(DRIVER), "tests/cpp/_build/radiation_driver missing: run __graft_entry__.build()"


def run_driver(mode, tmp_path):
    out = str(tmp_path / f"{mode}.bin")
    print(f"[cpp-driver] {mode}", flush=True)
    r = subprocess.run([DRIVER, mode, out], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return np.fromfile(out, dtype=np.float64), r.stdout


@pytest.mark.gpu
def test_radiation_flat_minkowski_100sq_matches_oracle(O, tmp_path):
    """Minkowski 100x100 on [-1,1]^2, 20 directions, CFL 0.5, 100 steps,
    Gaussian emission (sigma 0.1), no scattering: FP64 <= 1e-12."""
    data, out = run_driver("flat", tmp_path)
    g = O.make_grid(100, 100, (-1, -1), (1, 1), 0.5)
    N, D = g.N, 20
    x = g.startx + (np.arange(N) % g.nx) * g.dx
    y = g.starty + (np.arange(N) // g.nx) * g.dy
    P = O.Problem(g, D, 0, O.metric(g, 0), fields={"eta": np.exp(-(x * x + y * y) / (2 * 0.1 ** 2))})
    Iref, mom, mx = P.run(np.zeros(N * D), 100)[:3]
    I = data[:N * D]
    E, Fx, Fy = data[N * D:N * D + N], data[N * D + N:N * D + 2 * N], data[N * D + 2 * N:]
    assert rel_l2(I, Iref) <= 1e-12
    for q, v in enumerate((E, Fx, Fy)):
        assert rel_l2(v, mom[q]) <= 1e-12, q
    assert f"maxItterationCount {mx}" in out


@pytest.mark.gpu
def test_radiation_curved_schwarzschild_matches_oracle(O, tmp_path):
    """Schwarzschild 40x40 on [-4,4]^2, 16 directions, 10 steps, ring
    emission, GPU-built geodesic tables (tol 1e-4) against the oracle's."""
    data, _ = run_driver("curved", tmp_path)
    g = O.make_grid(40, 40, (-4, -4), (4, 4), 0.5)
    N, D = g.N, 16
    x = g.startx + (np.arange(N) % g.nx) * g.dx
    y = g.starty + (np.arange(N) // g.nx) * g.dy
    r = np.sqrt(x * x + y * y)
    bh = 2.1 * 2.1
    M = O.metric(g, 1)
    P = O.Problem(g, D, 2, M, bh_r2=bh, tables=O.geodesic_tables(g, M, bh, 1e-4),
                  fields={"eta": ((r > 3) & (r < 3.5)).astype(float)})
    Iref = P.run(np.zeros(N * D), 10)[0]
    assert rel_l2(data[:N * D], Iref) <= 1e-12
