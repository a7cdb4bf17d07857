"""Collect the numbers of the reference's AAA unit tests into a JSON fixture.

Source: ``vendor/analytic/test/testAAA.pf`` of SternheimerGW.  The fixture
holds data only (points, values, expected support points, weights, poles,
residues, step counts), keyed by the test that asserts them:

* ``realistic``: make_realistic_example (:76-162), asserted by
  test_generate_realistic_example (:24-45) and
  test_pole_residual_realistic_example (:47-73);
* ``tangent``: test_tangent (:222-269), test_evaluate_tangent (:271-315),
  test_pole_residual (:371-403);
* ``threshold``: test_threshold (:317-348).

Run in a container where the reference tree is mounted; the JSON is committed.
"""
import json
import os
import re

SRC = "/root/reference/vendor/analytic/test/testAAA.pf"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "aaa_reference.json")

NUM = r"[-+]?\d+\.\d*(?:[eE][-+]?\d+)?"


def _reals(text):
    return [float(x) for x in re.findall(NUM + r"(?=_dp)", text)]


def _cmplx(text):
    """CMPLX(re, im, KIND=dp) pairs in order."""
    out = []
    for m in re.finditer(r"CMPLX\(\s*(" + NUM + r")(?:_dp)?\s*,\s*(" + NUM + r")(?:_dp)?", text):
        out.append([float(m.group(1)), float(m.group(2))])
    return out


def _between(src, start, end):
    i = src.index(start)
    return src[i: src.index(end, i)]


def main():
    src = open(SRC).read()
    real = _between(src, "FUNCTION make_realistic_example", "END FUNCTION make_realistic_example")
    zz = _reals(_between(real, "example%zz = imag *", "example%ff ="))
    ff = _cmplx(_between(real, "example%ff =", "ALLOCATE(example%selection"))
    sel = [int(x) for x in re.findall(r"\d+", _between(real, "example%selection = [", "]")[21:])]
    wgt = _cmplx(_between(real, "example%ref_weight =", "example%ref_pole ="))
    pole = _cmplx(_between(real, "example%ref_pole =", "example%ref_residual ="))
    res = _cmplx(real[real.index("example%ref_residual ="):])
    thr = re.search(r"example%threshold = (eps\d+)", real).group(1)
    realistic = {"threshold": float("1e-" + thr[3:]), "z_imag": zz, "f": ff,
                 "selection_1based": sel, "weight": wgt, "pole": pole, "residual": res}

    tan = _between(src, "SUBROUTINE test_tangent", "END SUBROUTINE test_tangent")
    tz = _reals(_between(tan, "zz(11) = [", "* imag"))
    pos = _reals(_between(tan, "pos_ref(6) = [", "* imag"))
    val = _reals(_between(tan, "val_ref(6) = [", "* imag"))
    tw = _reals(_between(tan, "weight_ref(6) = [", "]"))
    ev = _between(src, "SUBROUTINE test_evaluate_tangent", "END SUBROUTINE test_evaluate_tangent")
    ez = _reals(_between(ev, "zz(11) = [", "+ imag"))
    eref = _cmplx(_between(ev, "ff_ref(11) = [", "COMPLEX(dp), ALLOCATABLE"))
    pr = _between(src, "SUBROUTINE test_pole_residual\n", "END SUBROUTINE test_pole_residual\n")
    p1 = _reals(re.search(r"pole1 = ([^\n]*)", pr).group(1))
    p2 = _reals(re.search(r"pole2 = ([^\n]*)", pr).group(1))
    p0 = _reals(re.search(r"ref_pole\(5\) = \[([^,]*)", pr).group(1))
    r1 = _reals(re.search(r"res1 = ([^\n]*)", pr).group(1))
    r2 = _reals(re.search(r"res2 = ([^\n]*)", pr).group(1))
    r0 = _reals(re.search(r"ref_res\(5\) = \[([^+]*)", pr).group(1))
    # ref_pole = [p0 i, pole1, -conj(pole1), pole2, -conj(pole2)],
    # ref_res = [r0, res1, conj(res1), res2, conj(res2)]
    tangent = {"threshold": 1e-14, "z_imag": tz, "position_imag": pos, "value_imag": val,
               "weight": tw, "eval_z_real_plus_i": ez, "eval_f": eref,
               "pole": [[0.0, p0[0]], p1, [-p1[0], p1[1]], p2, [-p2[0], p2[1]]],
               "residual": [[r0[0], 0.0], r1, [r1[0], -r1[1]], r2, [r2[0], -r2[1]]]}

    th = _between(src, "SUBROUTINE test_threshold", "END SUBROUTINE test_threshold")
    thres = [float(x) for x in re.findall(r"(\d+e-\d+)_dp", _between(th, "thres_list(6) = [", "]"))]
    steps = [int(x) for x in re.findall(r"\d+", _between(th, "step_list(6) = [", "]")[16:])]
    threshold = {"num_point": 49, "thres": thres, "steps": steps}

    with open(OUT, "w") as fh:
        json.dump({"source": "vendor/analytic/test/testAAA.pf", "realistic": realistic,
                   "tangent": tangent, "threshold": threshold}, fh, indent=1)
    print(OUT)


if __name__ == "__main__":
    main()
