"""Oracle pin: smoothing kernels (include/kernel.hpp:35-100).

Checked against textbook forms written independently of the oracle's factorised ones:
  M4 cubic    W = 1 - 3/2 q^2 + 3/4 q^3 (q < 1),  1/4 (2 - q)^3 (1 <= q < 2), 0 beyond
  M6 quintic  W = (3-q)_+^5 - 6 (2-q)_+^5 + 15 (1-q)_+^5
plus the 3-D normalisation  4 pi cnormk int_0^R W q^2 dq = 1, the derivative by central
differences, continuity at the piece boundaries, and compact support.
"""
import numpy as np
from scipy import integrate

from oracle import oracle as O


def textbook_cubic(q):
    q = np.asarray(q, dtype=float)
    return np.where(q < 1, 1 - 1.5 * q**2 + 0.75 * q**3, np.where(q < 2, 0.25 * (2 - q) ** 3, 0.0))


def textbook_quintic(q):
    q = np.asarray(q, dtype=float)
    p = lambda a: np.maximum(a, 0.0) ** 5
    return p(3 - q) - 6 * p(2 - q) + 15 * p(1 - q)


def test_cubic_closed_form_and_derivative():
    q = np.linspace(0, 2.5, 2001)
    w, dw = O.kernel(q, quintic=False)
    np.testing.assert_allclose(w, textbook_cubic(q), rtol=0, atol=1e-14)
    eps = 1e-6
    qi = q[(q > 0.01) & (np.abs(q - 1) > 1e-3) & (q < 1.99)]
    fd = (textbook_cubic(qi + eps) - textbook_cubic(qi - eps)) / (2 * eps)
    np.testing.assert_allclose(O.kernel(qi)[1], fd, rtol=0, atol=1e-8)


def test_quintic_closed_form_and_derivative():
    q = np.linspace(0, 3.5, 3501)
    w, dw = O.kernel(q, quintic=True)
    np.testing.assert_allclose(w, textbook_quintic(q), rtol=0, atol=1e-11)
    eps = 1e-6
    qi = q[(q > 0.01) & (q < 2.99)]
    fd = (textbook_quintic(qi + eps) - textbook_quintic(qi - eps)) / (2 * eps)
    np.testing.assert_allclose(O.kernel(qi, quintic=True)[1], fd, rtol=0, atol=1e-6)


def test_normalisation_3d():
    for quintic in (False, True):
        k = O.kernel_consts(quintic)
        f = lambda q: O.kernel(np.array([q]), quintic)[0][0] * q * q
        pieces = [0, 1, 2] + ([3] if quintic else [])
        total = sum(integrate.quad(f, a, b, epsabs=1e-14, epsrel=1e-14)[0] for a, b in zip(pieces[:-1], pieces[1:]))
        assert abs(4 * np.pi * k["cnormk"] * total - 1.0) < 1e-12


def test_constants_and_support():
    c = O.kernel_consts(False)
    assert (c["rad"], c["rad2"], c["hfact"]) == (2.0, 4.0, 1.2)
    assert abs(c["cnormk"] - 1 / np.pi) < 1e-16
    q = O.kernel_consts(True)
    assert (q["rad"], q["rad2"], q["hfact"]) == (3.0, 9.0, 1.0)
    assert abs(q["cnormk"] - 1 / (120 * np.pi)) < 1e-18
    for quintic, r in ((False, 2.0), (True, 3.0)):
        w, dw = O.kernel(np.array([r, r + 0.1, 10.0]), quintic)
        assert np.all(w == 0) and np.all(dw == 0)


def test_continuity_at_piece_boundaries():
    for quintic, edges in ((False, [1.0, 2.0]), (True, [1.0, 2.0, 3.0])):
        for e in edges:
            wl, dl = O.kernel(np.array([e - 1e-9]), quintic)
            wr, dr = O.kernel(np.array([e + 1e-9]), quintic)
            assert abs(wl[0] - wr[0]) < 1e-6 and abs(dl[0] - dr[0]) < 1e-6
