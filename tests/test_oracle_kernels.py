"""Smoothing kernels of the oracle (MPSSPHKernel.cuh:9-84) pinned by the mathematics.

- W (poly6 / cubic spline) integrates to 1 over its support: pins CONST_W, CONST_CUBICW and
  the powers of h and q.
- G is the radial derivative divided by r of a normalised kernel (spiky for the default
  family, the cubic spline itself for the cubic family): -(4 pi / 3) int r^4 G dr = 1 pins
  CONST_G / CONST_CUBICG, and for the cubic spline G r = dW/dr pins both branches.
- STG r = d STW / dr on (h/2, h) and STW is flat below h/2: pins CONST_STW against CONST_STG.
"""
import numpy as np
import pytest

from oracle.binding import kernel_eval

H = 0.04


def _radial(n=200001):
    r = np.linspace(0.0, H, n)
    return r


@pytest.mark.parametrize("kernel", [0, 1])
def test_w_normalised(kernel):
    r = _radial()
    w, _, _, _ = kernel_eval(64, kernel, r, 1.0 / H)
    integral = np.trapezoid(4.0 * np.pi * r ** 2 * w, r)
    assert abs(integral - 1.0) < 1e-6, integral


@pytest.mark.parametrize("kernel", [0, 1])
def test_gradient_normalised(kernel):
    # int_0^h 4 pi r^2 W_g dr = -(4 pi/3) int_0^h r^3 W_g'(r) dr and W_g' = G r
    r = _radial()
    _, g, _, _ = kernel_eval(64, kernel, r, 1.0 / H)
    integral = -(4.0 * np.pi / 3.0) * np.trapezoid(r ** 4 * g, r)
    assert abs(integral - 1.0) < 1e-6, integral


def test_cubic_gradient_is_derivative():
    r = np.linspace(0.05 * H, 0.95 * H, 37)
    eps = 1e-7 * H
    wp, _, _, _ = kernel_eval(64, 1, r + eps, 1.0 / H)
    wm, _, _, _ = kernel_eval(64, 1, r - eps, 1.0 / H)
    _, g, _, _ = kernel_eval(64, 1, r, 1.0 / H)
    dwdr = (wp - wm) / (2 * eps)
    np.testing.assert_allclose(g * r, dwdr, rtol=1e-6, atol=1e-3 * np.abs(dwdr).max() * 1e-3)


def test_spiky_gradient_shape():
    # spiky: G r = -(45/pi) (h - r)^2 / h^6, i.e. proportional to (1-q)^2 with no other q dependence
    r = np.linspace(0.1 * H, 0.9 * H, 9)
    _, g, _, _ = kernel_eval(64, 0, r, 1.0 / H)
    ratio = (g * r) / (1.0 - r / H) ** 2
    np.testing.assert_allclose(ratio, ratio[0], rtol=1e-12)


def test_surface_tension_kernels_consistent():
    r = np.linspace(0.52 * H, 0.98 * H, 24)
    eps = 1e-7 * H
    _, _, sp, _ = kernel_eval(64, 0, r + eps, 1.0 / H)
    _, _, sm, _ = kernel_eval(64, 0, r - eps, 1.0 / H)
    _, _, _, stg = kernel_eval(64, 0, r, 1.0 / H)
    np.testing.assert_allclose(stg * r, (sp - sm) / (2 * eps), rtol=1e-6)
    # plateau of STW and zero STG below h/2
    rl = np.linspace(0.0, 0.5 * H, 11)
    _, _, stw, stg = kernel_eval(64, 0, rl, 1.0 / H)
    np.testing.assert_allclose(stw, stw[0], rtol=1e-14)
    assert np.all(stg == 0.0)


@pytest.mark.parametrize("kernel", [0, 1])
def test_support_and_origin(kernel):
    r = np.array([H, 1.0001 * H, 2 * H])
    w, g, stw, stg = kernel_eval(64, kernel, r, 1.0 / H)
    assert np.all(w == 0) and np.all(g == 0) and np.all(stw == 0) and np.all(stg == 0)
    _, g0, _, _ = kernel_eval(64, kernel, np.array([0.0]), 1.0 / H)
    assert g0[0] == 0.0  # ratio < 1e-20 guard


@pytest.mark.parametrize("kernel", [0, 1])
def test_fp32_tracks_fp64(kernel):
    # sample points avoid q = 0.5 exactly, where STG (and the cubic branches) switch
    r = np.linspace(0.013 * H, 0.987 * H, 97)
    a = kernel_eval(64, kernel, r, 1.0 / H)
    b = kernel_eval(32, kernel, r, 1.0 / H)
    for x, y in zip(a, b):
        np.testing.assert_allclose(y, x, rtol=1e-4, atol=1e-5 * np.abs(x).max())
