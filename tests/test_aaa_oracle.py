"""Pins of the AAA oracle (vendor/analytic/src/aaa.f90 and its uses in
algo/analytic/src/analytic.f90) against the reference's own unit tests
(vendor/analytic/test/testAAA.pf, numbers in tests/golden/aaa_reference.json)
and closed forms."""
import json
import os

import numpy as np
import pytest

from oracle import aaa_oracle as aa

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "aaa_reference.json")))


def _c(pairs):
    return np.array([complex(a, b) for a, b in pairs])


def test_realistic_example_support_points_and_weights():
    """test_generate_realistic_example (testAAA.pf:24-45): 35 points of a GW
    screened interaction at 1e-10: the 14 support points, their values, and
    the weights up to a phase (1e-6)."""
    r = GOLD["realistic"]
    z = 1j * np.array(r["z_imag"])
    f = _c(r["f"])
    pos, val, wgt = aa.aaa_generate(r["threshold"], -1, z, f)
    sel = np.array(r["selection_1based"]) - 1
    assert np.array_equal(np.sort(np.searchsorted(z.imag, pos.imag)), sel)
    assert np.abs(np.sort_complex(pos) - np.sort_complex(z[sel])).max() <= 1e-14
    order = np.argsort(pos.imag)
    ref_w = _c(r["weight"])
    w = wgt[order]
    phase = w[0] / ref_w[0]
    assert np.abs(ref_w * phase - w).max() <= 1e-6


def test_realistic_example_poles_and_residues():
    """test_pole_residual_realistic_example (:47-73): from the reference's
    support points and weights, 13 poles (sorted by 1/|pole|) to 1e-6 and
    residues to 2e-6.  The residues are means of r(p + s) s over shifts of
    1e-6, where the barycentric denominator's slope is ~1e-3, so rounding
    alone moves them by ~1e-6: against 40-digit residues of the same weights
    the oracle is off by 1.2e-6, the reference's numbers by 6.7e-8 (its test
    asserts 1e-6)."""
    r = GOLD["realistic"]
    z = 1j * np.array(r["z_imag"])
    f = _c(r["f"])
    sel = np.array(r["selection_1based"]) - 1
    pos, val, wgt = z[sel], f[sel], _c(r["weight"])
    p = aa.poles(pos, wgt)
    res = aa.residues(pos, val, wgt, p)
    order = np.argsort(1.0 / np.abs(p), kind="stable")
    assert np.abs(p[order] - _c(r["pole"])).max() <= 1e-6
    assert np.abs(res[order] - _c(r["residual"])).max() <= 2e-6


def test_tangent_support_points_values_weights():
    """test_tangent (:222-269): tan(pi z / 2) on 11 imaginary points at
    1e-14: 6 support points and values (1e-14), weights up to a phase (1e-10)."""
    t = GOLD["tangent"]
    z = 1j * np.array(t["z_imag"])
    f = np.tan(np.pi / 2 * z)
    pos, val, wgt = aa.aaa_generate(1e-14, -1, z, f)
    order = np.argsort(pos.imag)
    assert pos.size == 6
    assert np.abs(pos[order] - 1j * np.array(t["position_imag"])).max() <= 1e-14
    assert np.abs(val[order] - 1j * np.array(t["value_imag"])).max() <= 1e-14
    ref_w = np.array(t["weight"], np.complex128)
    w = wgt[order]
    assert np.abs(ref_w * (w[0] / ref_w[0]) - w).max() <= 1e-10


def test_tangent_evaluation_and_poles():
    """test_evaluate_tangent (:271-315) at 1e-12; test_pole_residual
    (:371-403): the 5 poles at 1e-10, residues at 1e-8 (the far pole's at
    1e-3).  ZGGEV's order of the +-conj pairs differs between LAPACK builds,
    so each reference pole is matched to the nearest one."""
    t = GOLD["tangent"]
    pos = 1j * np.array(t["position_imag"])
    val = 1j * np.array(t["value_imag"])
    wgt = np.array(t["weight"], np.complex128)
    zz = np.array(t["eval_z_real_plus_i"]) + 1j
    assert np.abs(aa.aaa_evaluate(pos, val, wgt, zz) - _c(t["eval_f"])).max() <= 1e-12
    p = aa.poles(pos, wgt)
    ref_p = _c(t["pole"])
    assert p.size == ref_p.size
    near = np.array([int(np.argmin(np.abs(p - q))) for q in ref_p])
    assert sorted(near.tolist()) == list(range(p.size))
    assert np.abs(p[near] - ref_p).max() <= 1e-10
    res = aa.residues(pos, val, wgt, p)[near]
    ref = _c(t["residual"])
    assert abs(res[0] - ref[0]) <= 1e-3
    assert np.abs(res[1:] - ref[1:]).max() <= 1e-8


def test_threshold_steps():
    """test_threshold (:317-348): log(2 - z^4) / (1 - 16 z^4) on 49 points of
    the circle |z| = sqrt(1/2): support-point counts per threshold."""
    g = GOLD["threshold"]
    n = g["num_point"]
    z = np.exp(2j * np.pi * np.arange(n) / n) * np.sqrt(0.5)
    f = np.log(2 - z ** 4) / (1 - 16 * z ** 4)
    for thres, steps in zip(g["thres"], g["steps"]):
        pos, _, _ = aa.aaa_generate(thres, -1, z, f)
        assert pos.size == steps, (thres, pos.size)


def test_constant_rational_restriction():
    """test_constant_function, test_rational_function, test_restriction,
    test_error_raised (:164-220, :350-369, :405-420)."""
    pos, val, wgt = aa.aaa_generate(1e-14, -1, np.array([1.0 + 0j]), np.array([2.0 + 0j]))
    assert pos.tolist() == [1] and val.tolist() == [2] and abs(abs(wgt[0]) - 1) <= 1e-14
    z = np.array([1.0, 2.0], np.complex128)
    f = 6.0 / (z + 1.0)
    pos, val, wgt = aa.aaa_generate(1e-14, -1, z, f)
    assert pos.size == 2
    assert np.abs(aa.aaa_evaluate(pos, val, wgt, z) - f).max() <= 1e-14
    rng = np.random.default_rng(1)
    z = rng.random(100) + 1j * rng.random(100)
    f = rng.random(100) + 1j * rng.random(100)
    for mp in (1, 7, 30):
        assert aa.aaa_generate(1e-8, mp, z, f)[0].size <= mp
    with pytest.raises(ValueError):
        aa.aaa_generate(1e-14, -1, np.ones(1), np.ones(2))


def test_rational_function_recovered_and_pole_form():
    """Closed form: r(z) = sum_k c_k / (z - p_k) with 3 poles off the mesh is
    recovered by AAA with 4 support points; pole_correction gives its poles
    and residues, and both coefficient forms evaluate to r."""
    p = np.array([0.5 + 2.0j, -1.0 - 1.5j, 2.5 + 0.3j])
    c = np.array([1.0 - 0.5j, 0.3j, -0.7 + 0.2j])
    z = 1j * np.linspace(0.0, 4.0, 24)
    r = lambda x: (c[None, :] / (np.atleast_1d(x)[:, None] - p[None, :])).sum(1)
    f = r(z)
    pos, val, wgt = aa.aaa_generate(1e-12, -1, z, f)
    assert pos.size == 4
    pp = aa.poles(pos, wgt)
    assert pp.size == 3
    order = [int(np.argmin(np.abs(pp - q))) for q in p]
    assert np.abs(pp[order] - p).max() <= 1e-9
    res = aa.residues(pos, val, wgt, pp)
    assert np.abs(res[order] - c).max() <= 1e-6
    c4 = aa.aaa_coeff(4, 1e-12, z, f)
    c5 = aa.aaa_coeff(5, 1e-12, z, f)
    assert c4.size == 24 and c5.size == 24
    for x in (0.3 + 0.1j, 1.0 + 1.0j, -2.0j):
        want = r(x)[0]
        assert abs(aa.aaa_eval(4, c4, x) - want) <= 1e-9 * abs(want)
        assert abs(aa.aaa_eval(5, c5, x) - want) <= 1e-5 * abs(want)
