"""Writes the golden fixtures of this directory (run from the repository
root: python tests/golden/make_golden.py).  Nothing here runs the
reference; every expected value comes from mathematics or from numbers the
reference publishes.

  readme_1e9.json      the 10^9-digit run published in the reference README
                       (README.md:72-74): RET length, RET hash, RET hex
                       preview, first and last 64 decimal digits, and the
                       digit count the sizing rounds up to
                       (pi_bs_gpu.cu:924-944).
  convert_cases.json   the four convert_decimal correctness cases of
                       convert_decimal_test.cu:301-304 (477 * 2^k digits, k =
                       0..3): input words from the test's generator
                       (:58-80) and the exact expansion floor(X 10^d / 2^bits)
                       (:82-107), computed with Python integers.
  pi_digits.txt        "3" and the first 20000 decimal digits of pi after the
                       point, from mpmath (an implementation independent of
                       this repository and of the reference).
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import refcases  # noqa: E402

PI_DIGITS = 20000


def main() -> None:
    with open(os.path.join(HERE, "readme_1e9.json"), "w") as f:
        json.dump(refcases.README_1E9, f, indent=1, sort_keys=True)
        f.write("\n")
    cases = [{"digits": t, "bits": b, "words": w, "expected": e} for t, b, w, e in refcases.convert_cases()]
    with open(os.path.join(HERE, "convert_cases.json"), "w") as f:
        json.dump(cases, f)
        f.write("\n")
    import mpmath
    mpmath.mp.dps = PI_DIGITS + 40
    s = mpmath.nstr(mpmath.pi, PI_DIGITS + 20, strip_zeros=False).replace(".", "")[:PI_DIGITS + 1]
    with open(os.path.join(HERE, "pi_digits.txt"), "w") as f:
        f.write(s + "\n")


if __name__ == "__main__":
    main()
