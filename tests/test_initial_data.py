"""The von Mises flux-to-sigma tables of the C++ Stencil
(Stencil::PopulateLookUpTable, src/Stencil.cpp:378-455), pinned by what the
tables mean rather than by retracing the walk:

* each entry's log-normalisation makes the unit-energy distribution
  exp(sigma C_x - z) integrate to 1 under the stencil weights;
* each entry's relative flux is the flux of that distribution;
* flux and sigma increase together, sigma starts at 0;
* every accepted sigma keeps the mean relative error of the cubic angular
  interpolation within 1% near the flux direction, and a step of 1e-7 past
  the last one does not (the walk stops at refinement 8).
"""
import numpy as np
import pytest


def _myacos(x):
    """Hastings arccos of the reference (float argument and result)."""
    x = np.float32(x)
    a = abs(float(x))
    p = 1.57078786 + a * (-0.21412453 + a * (0.08466649 + a * (-0.03575663 + a * 0.00864884)))
    asin = np.pi / 2 - np.sqrt(1.0 - a) * p
    return float(np.float32(np.pi / 2 - asin if x >= 0 else np.pi / 2 + asin))


def _interp_error(O, sigma, w, cx, Nc, Wc, nGrid):
    """Mean relative error of the cubic interpolation of exp(sigma C_x - z)
    at the fine angles near angle 0; the fine-angle cosines come from the
    oracle's MyCos (float-rounded, as the library's interpolation grid)."""
    z = np.log(np.sum(w * np.exp(sigma * cx)))
    I = np.exp(sigma * cx - z)
    err, cnt = 0.0, 0
    for g in range(nGrid):
        cg = O.lib().or_mycos(2.0 * np.pi * g / nGrid)
        if _myacos(cg) > (np.pi / 8 + 1e-8) / 2:
            continue
        exact = np.exp(sigma * cg - z)
        cubic = np.dot(Wc[4 * g:4 * g + 4], I[Nc[4 * g:4 * g + 4]])
        err += abs((exact - cubic) / exact)
        cnt += 1
    return err / cnt


@pytest.mark.parametrize("n", [16, 32])
def test_flux_tables(rt, O, n):
    f, s, z = rt.build_flux_tables(n)
    w, phi, cx, cy, Nc, Wc = rt.build_stencil(n)
    assert s[0] == 0.0 and np.all(np.diff(f) > 0) and np.all(np.diff(s) > 0)
    for k in range(len(f)):
        e = np.exp(s[k] * cx - z[k])
        assert abs(np.sum(w * e) - 1.0) < 1e-13
        assert abs(np.hypot(np.sum(w * e * cx), np.sum(w * e * cy)) - f[k]) < 1e-13
    nGrid = 10 * n
    for k in np.linspace(0, len(f) - 1, 12).astype(int):
        assert _interp_error(O, s[k], w, cx, Nc, Wc, nGrid) <= 0.01
    assert _interp_error(O, s[-1] + 1e-7, w, cx, Nc, Wc, nGrid) > 0.01
