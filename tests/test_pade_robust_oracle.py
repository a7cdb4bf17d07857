"""Pins of the robust Pade oracle (algo/analytic/src/pade_robust.f90) against
closed forms: rational functions and polynomials sampled on a circle are
recovered with their true degrees and coefficients, the zero function gives
(0, 0), full-rank data satisfy the Pade conditions of the (m, n) block, and
complex Taylor coefficients give the conjugate denominator that the
reference's QR of the plain transpose (C D)^T produces."""
import numpy as np
import pytest

from oracle import pade_robust_oracle as pr


def _rat(a, b):
    a, b = np.asarray(a, np.complex128), np.asarray(b, np.complex128)
    return lambda x: np.polyval(a[::-1], x) / np.polyval(b[::-1], x)


@pytest.mark.parametrize("a,b,npts,radius", [
    ([1.0, 0.3, 0.1], [1.0, -0.2, 0.05], 32, 1.0),
    ([0.5, -0.4], [1.0, -0.25, -5 / 72, 1 / 72], 40, 1.0),     # poles at 3, -4, 6
    ([2.0, 0.1, 0.0, 0.03], [1.0, 0.3], 24, 1.0),
    ([1.0, 0.3], [1.0, -0.1 + 1 / 12, -1 / 120], 24, 2.0),   # poles at 10, -12
])
def test_real_rational_recovered(a, b, npts, radius):
    f = _rat(a, b)
    z = pr.circle(radius, npts)
    deg = npts // 2 - 2
    ga, gb = pr.pade_robust(radius, f(z), deg, deg)
    assert (ga.size - 1, gb.size - 1) == (len(a) - 1, len(b) - 1)
    assert np.abs(ga - np.asarray(a)).max() <= 1e-12
    assert np.abs(gb - np.asarray(b)).max() <= 1e-12
    # real Taylor coefficients are made exactly real (pade_derivative)
    assert np.all(ga.imag == 0) and np.all(gb.imag == 0)


def test_polynomial_and_constant():
    z = pr.circle(1.0, 20)
    p = np.array([1.0, 2.0, 0.0, -0.5])
    ga, gb = pr.pade_robust(1.0, np.polyval(p[::-1], z), 8, 8)
    assert ga.size == 4 and gb.size == 1
    assert np.abs(ga - p).max() <= 1e-13 and gb[0] == 1
    ga, gb = pr.pade_robust(1.0, np.full(20, 0.7 - 0.2j), 8, 8)
    assert ga.size == 1 and gb.size == 1
    assert abs(ga[0] - (0.7 - 0.2j)) <= 1e-15


def test_zero_function():
    ga, gb = pr.pade_robust(1.0, np.zeros(16, np.complex128), 6, 6)
    assert ga.tolist() == [0] and gb.tolist() == [1]


def test_full_rank_data_meet_pade_conditions():
    """For generic values the block keeps (m, n) and (c * b)_i = a_i for
    i <= m, (c * b)_i = 0 for m < i <= m + n, c the Taylor coefficients."""
    rng = np.random.default_rng(5)
    npts, radius = 16, 1.5
    f = rng.standard_normal(npts) + 1j * rng.standard_normal(npts)
    ga, gb = pr.pade_robust(radius, f, 6, 6)
    assert (ga.size, gb.size) == (7, 7)
    c = np.fft.fft(f) / npts / radius ** np.arange(npts)
    # a = Z b; for complex c the null vector of C is conj(b) (the reference's
    # QR of the plain transpose, see the complex test below)
    assert np.abs(np.convolve(c[:13], gb)[:7] - ga).max() <= 1e-12 * np.abs(ga).max()
    null = np.convolve(c[:13], np.conj(gb))[7:13]
    assert np.abs(null).max() <= 1e-12 * np.abs(c).max() * np.abs(gb).max()


def test_complex_coefficients_give_conjugate_denominator():
    """(C D)^T = QR, q_{n+1} orthogonal to the columns: (C D) conj(q) = 0,
    so b = D q is the conjugate of C's null vector, and a = Z b."""
    at = np.array([1.0, 0.3 + 0.2j, 0.1])
    bt = np.array([1.0, -(0.2 - 0.1j), 0.05j])
    z = pr.circle(1.0, 32)
    ga, gb = pr.pade_robust(1.0, _rat(at, bt)(z), 14, 14)
    assert (ga.size, gb.size) == (3, 3)
    assert np.abs(gb - np.conj(bt)).max() <= 1e-13
    c = np.fft.fft(_rat(at, bt)(z)) / 32
    za = np.array([sum(c[i - j] * gb[j] for j in range(i + 1)) for i in range(3)])
    assert np.abs(ga - za).max() <= 1e-13


def test_coeff_layout_eval_and_checks():
    a, b = [1.0, 0.3, 0.1], [1.0, -0.2, 0.05]
    f = _rat(a, b)
    z = pr.circle(1.0, 20)
    func = np.stack([f(z), np.zeros(20), 3.0 * f(z)])
    coeff = pr.pade_coeff_robust(z, func)
    assert coeff.shape == (3, 20)
    assert coeff[0, :2].tolist() == [2, 2]
    assert np.abs(coeff[0, 2:8] - np.array(a + b)).max() <= 1e-12
    assert np.all(coeff[0, 8:] == 0)
    assert coeff[1, :4].tolist() == [0, 0, 0, 1]
    for x in (0.3 + 0.1j, -0.5j, 0.9):
        assert abs(pr.pade_eval_robust(coeff[0], x) - f(x)) <= 1e-12
        assert abs(pr.pade_eval_robust(coeff[2], x) - 3 * f(x)) <= 1e-12
    with pytest.raises(ValueError, match="at least 10"):
        pr.pade_coeff_robust(pr.circle(1.0, 8), np.ones((1, 8)))
    bad = z.copy()
    bad[3] *= 1.01
    with pytest.raises(ValueError, match="circle"):
        pr.pade_coeff_robust(bad, np.ones((1, 20)))
