"""Extracts the golden vectors the reference's own tests hold into
tests/golden/reference_vectors.json and copies the translation-coefficient CSV
fixtures next to it.  Run once where /root/reference is mounted:

    python tests/golden/extract_reference_vectors.py

Only numbers (inputs and expected outputs) are written; no reference source text.
"""
import json
import os
import re
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

NUM = r"[-+]?\d*\.?\d+(?:[dDeE][-+]?\d+)?(?:_8)?"


def _f(tok):
    return float(tok.replace("_8", "").replace("d", "e").replace("D", "e"))


def data_block(path, first_line):
    """Numbers of a Fortran `data name / ... /` statement starting at first_line (1-based)."""
    lines = open(path).read().split("\n")[first_line - 1:]
    txt = "\n".join(lines)
    body = txt[txt.index("/") + 1:]
    body = body[:body.index("/")]
    return [_f(t) for t in re.findall(NUM, body)]


def main():
    out = {}
    bessel = os.path.join(REF, "libmath/src/math/lib_math_bessel.f90")
    # test_lib_math_bessel_spherical_first_kind_real: x = 5, n = 0..40, rel 1e-13
    out["sph_j_x5"] = {"x": 5.0, "values": data_block(bessel, 2459), "rtol": 1e-13,
                       "cite": "libmath/src/math/lib_math_bessel.f90:2433-2515"}
    # test_lib_math_bessel_spherical_second_kind_real: x = 5, n = 0..40, rel 1e-13
    out["sph_y_x5"] = {"x": 5.0, "values": data_block(bessel, 2627), "rtol": 1e-13,
                       "cite": "libmath/src/math/lib_math_bessel.f90:2601-2683"}
    # test_lib_math_bessel_spherical_third_kind_1_real: x = 2, h1_n assignments (n-1 indexed)
    src = open(bessel).read().split("\n")
    h1 = {}
    for ln in src[2783:2874]:
        m = re.match(r"\s*ground_truth_y\((\d+)\)\s*=\s*cmplx\((" + NUM + r"),\s*([-+])\s*(" + NUM + r")", ln)
        if m:
            idx = int(m.group(1)) - 1
            im = _f(m.group(4)) * (-1.0 if m.group(3) == "-" else 1.0)
            h1[idx] = [_f(m.group(2)), im]
    out["sph_h1_x2"] = {"x": 2.0, "values": {str(k): v for k, v in sorted(h1.items())}, "log_tol": 1e-13,
                        "cite": "libmath/src/math/lib_math_bessel.f90:2784-2874"}
    # Wigner 3j (Mathematica), lib_math_wigner.f90:127-164
    wig = open(os.path.join(REF, "libmath/src/math/lib_math_wigner.f90")).read().split("\n")
    vals = []
    for ln in wig[150:165]:
        m = re.match(r"\s*ground_truth\((\d)\)\s*=\s*(" + NUM + ")", ln)
        if m:
            vals.append(_f(m.group(2)))
    out["wigner3j"] = {"j1": 98, "j2": 115, "j3": [213, 212, 211, 210, 209], "m1": -69, "m2": -100, "m3": 169,
                       "values": vals, "atol": 1e-14, "cite": "libmath/src/math/lib_math_wigner.f90:98-187"}
    # Xu-Cruzan a, b (lib_math_vector_spherical_harmonics.f90:2799-2874)
    vsh = open(os.path.join(REF, "libmath/src/math/lib_math_vector_spherical_harmonics.f90")).read().split("\n")
    ga, gb = {}, {}
    for ln in vsh[2820:2845]:
        m = re.match(r"\s*ground_truth_([ab])\((\d)\)\s*=\s*(" + NUM + ")", ln)
        if m:
            (ga if m.group(1) == "a" else gb)[int(m.group(2))] = _f(m.group(3))
    out["xu_ab"] = {"m": 15, "n": 16, "mu": -15, "nu": 20, "q": [1, 2, 3, 4, 5],
                    "a": [ga[q] for q in range(1, 6)], "b": [gb[q] for q in range(1, 6)], "atol": 1e-13,
                    "cite": "libmath/src/math/lib_math_vector_spherical_harmonics.f90:2799-2874"}
    # Mie coefficients, Barber & Hill program S2 values (lib_mie_ss_helper_functions.f90:1639-1766)
    mie = open(os.path.join(REF, "FoggySim/src/lib/mie_theory/lib_mie_ss_helper_functions.f90")).read().split("\n")

    def grab_ab(lo, hi):
        a, b = {}, {}
        for ln in mie[lo:hi]:
            m = re.match(r"\s*ground_truth_([ab])_n\((\d)\)\s*=\s*cmplx\((" + NUM + r"),\s*(" + NUM + ")", ln)
            if m:
                (a if m.group(1) == "a" else b)[int(m.group(2))] = [_f(m.group(3)), _f(m.group(4))]
        return [a[i] for i in range(1, 5)], [b[i] for i in range(1, 5)]
    a, b = grab_ab(1655, 1680)
    out["mie_s2_real"] = {"x": 10.0, "m": [1.5, 0.0], "a": a, "b": b, "rtol": 1e-6,
                          "cite": "FoggySim/src/lib/mie_theory/lib_mie_ss_helper_functions.f90:1639-1702"}
    a, b = grab_ab(1718, 1745)
    out["mie_s2_cmplx"] = {"x": 10.0, "m": [1.28, -1.37], "a": a, "b": b, "atol": 1e-6,
                           "cite": "FoggySim/src/lib/mie_theory/lib_mie_ss_helper_functions.f90:1704-1766"}
    # libtree fixtures (lib_tree_helper_functions.f90:2166-2579, lib_tree.f90:2425-2681)
    out["tree"] = {
        "uindex": {"point": [0.75, 0.5, 2.0 ** -9 + 2.0 ** -8], "level": 1, "n": 6},
        "uindex_lmax": {"point": [1.0 - 2.0 ** -21] * 3, "level": 21, "n": 0x7FFFFFFFFFFFFFFF},
        "parent": {"n": 6, "parent": 0},
        "children": {"n": 1, "children": [8, 9, 10, 11, 12, 13, 14, 15]},
        "centre": {"n": 1, "level": 1, "centre": [0.25, 0.25, 0.75]},
        "neighbours_35": {"k": 1, "n": 35, "level": 2,
                          "out": [4, 5, 12, 6, 7, 14, 20, 21, 28, 32, 33, 40, 34, 42, 48, 49, 56, 36, 37, 44, 38,
                                  39, 46, 52, 53, 60]},
        "neighbours_5": {"k": 1, "n": 5, "level": 1, "out": [0, 1, 2, 3, 4, 6, 7] + [-1] * 19},
        "i2_sum": {"n": 0, "level": 2, "sum": 28},
        "level_min": {"k": 1, "l_min": 2},
        "level_max": {"s": [1, 2, 3], "l_max": [3, 2, 1], "sorted": [5, 4, 3, 2, 1],
                      "points": [[1.0 - (0.999 * i) / 5.0] * 3 for i in range(1, 6)], "l_th": 16},
        "cite": "libtree/src/lib/lib_tree_helper_functions.f90:2166-2579; libtree/src/lib/lib_tree.f90:2425-2681",
    }
    with open(os.path.join(HERE, "reference_vectors.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    gt = os.path.join(REF, "FoggySim/ground_truth/lib_math_vector_spherical_harmonics")
    for fn in ("ground_truth_AB_mnmunu.csv", "ground_truth_AB_mnmunu_minus_phi.csv"):
        shutil.copy(os.path.join(gt, fn), os.path.join(HERE, fn))
    print("wrote", os.path.join(HERE, "reference_vectors.json"))


if __name__ == "__main__":
    main()
