"""CPU: pin the oracle's pipeline stages to mathematics and to the reference's
own known answers.

- leaf terms: rational identities with the unreduced Chudnovsky terms;
- the tree: R / (Q 2^(15*17N)) is exactly the partial Chudnovsky sum;
- sqrt(10005): Pell norm 1 and the size of (4001 + 40 sqrt 10005)^(2^k);
- the inverse: distance to floor(2^L / x);
- digits against mpmath's pi, RET against pi * 2^(32 prec);
- convert_decimal: the reference's four known-answer cases;
- sizing formulas: the reference's published lengths."""
import ctypes
import os
import math
from fractions import Fraction

import mpmath
import numpy as np
import pytest

import oracle_ffi as of
import refcases
from helpers import rand_rows, to_int

A_CONST, B_CONST = 13591409, 545140134
C3_OVER_24 = 640320 ** 3 // 24  # = 10005^2 * 3335 * 2^15


def term_index(is_head, t):
    return 17 * t + 16 if is_head else t + (t >> 4)


@pytest.mark.parametrize("is_head", [0, 1])
def test_leaf_terms_rational_identities(is_head):
    count = 4000
    P = np.zeros(count * 3, np.uint32)
    Q = np.zeros(count * 4, np.uint32)
    R = np.zeros(count * 6, np.uint32)
    of.lib().o_leaf17(is_head, P.ctypes.data, Q.ctypes.data, R.ctypes.data, count)
    reduced = 0
    for t in range(count):
        i = term_index(is_head, t)
        p = (6 * i + 1) * (2 * i + 1) * (6 * i + 5)
        q = (i + 1) ** 3 * C3_OVER_24 >> 15
        a = A_CONST + B_CONST * i
        Pi, Qi, Ri = to_int(P[3 * t:3 * t + 3]), to_int(Q[4 * t:4 * t + 4]), to_int(R[6 * t:6 * t + 6])
        assert Pi * q == p * Qi          # p/q preserved by the cancellation
        assert 3335 * Ri == a * Qi       # R/Q = a/3335
        assert q % Qi == 0 and p % Pi == 0
        reduced += Qi < q
    assert reduced > count // 2          # common factors were actually cancelled


def chudnovsky_partial(n):
    """sum_{k<n} (-1)^k a_k/(3335 2^15) prod_{j<k} p_j/(q_j 2^15) (exact)."""
    s = Fraction(0)
    ratio = Fraction(1)
    for k in range(n):
        a = A_CONST + B_CONST * k
        s += (-1) ** k * ratio * Fraction(a, 3335 << 15)
        p = (6 * k + 1) * (2 * k + 1) * (6 * k + 5)
        ratio *= Fraction(p, (k + 1) ** 3 * C3_OVER_24)
    return s


@pytest.mark.parametrize("groups", [1, 2, 4, 8])
def test_binary_split_is_the_chudnovsky_sum(groups):
    oq, orr = of.OArray(), of.OArray()
    assert of.lib().o_binary_split(groups, ctypes.byref(oq), ctypes.byref(orr)) == 0
    Q, R = to_int(of.take_array(oq)), to_int(of.take_array(orr))
    assert Fraction(R, Q << (15 * 17 * groups)) == chudnovsky_partial(17 * groups)


def group_sum(k0):
    """(P/Q, R/(Q 2^255)) of the 17 terms k0..k0+16, each group's series
    starting with a + sign."""
    s, ratio = Fraction(0), Fraction(1)
    for j in range(17):
        k = k0 + j
        s += (-1) ** j * ratio * Fraction(A_CONST + B_CONST * k, 3335 << 15)
        ratio *= Fraction((6 * k + 1) * (2 * k + 1) * (6 * k + 5), (k + 1) ** 3 * C3_OVER_24)
    return ratio * (1 << (15 * 17)), s


@pytest.mark.parametrize("groups", [1, 4])
def test_leaf_stage_is_each_groups_sum(groups):
    # the 17-term group stage (leaf terms, 4 pairwise levels, head merge,
    # exact division): row g holds group g's P/Q and series value
    op, oq, orr = of.OArray(), of.OArray(), of.OArray()
    assert of.lib().o_leaf_stage(groups, ctypes.byref(op), ctypes.byref(oq), ctypes.byref(orr)) == 0
    P, Q, R = of.take_rows(op), of.take_rows(oq), of.take_rows(orr)
    assert P.shape[0] == Q.shape[0] == R.shape[0] == groups
    for g in range(groups):
        pq, s = group_sum(17 * g)
        Pg, Qg, Rg = to_int(P[g]), to_int(Q[g]), to_int(R[g])
        assert Fraction(Pg, Qg) == pq
        assert Fraction(Rg, Qg << (15 * 17)) == s
    if groups == 1:  # one group: the whole tree
        oq2, or2 = of.OArray(), of.OArray()
        assert of.lib().o_binary_split(1, ctypes.byref(oq2), ctypes.byref(or2)) == 0
        assert to_int(of.take_array(oq2)) == to_int(Q[0]) and to_int(of.take_array(or2)) == to_int(R[0])


@pytest.mark.parametrize("digits", [5, 8, 16, 100, 5000, 100000])
def test_sqrt_10005_pell(digits):
    op, oq = of.OArray(), of.OArray()
    assert of.lib().o_sqrt_10005(digits, ctypes.byref(op), ctypes.byref(oq)) == 0
    P, Q = to_int(of.take_array(op)), to_int(of.take_array(oq))
    assert P * P - 10005 * Q * Q == 1
    k = 0
    while 7.5 * 2 ** k < digits:
        k += 1
    # P + Q sqrt(10005) = (4001 + 40 sqrt(10005))^(2^k) ~ 2P; neighbouring
    # powers differ by ~13 bits, so the size fixes the exponent
    want_bits = 2 ** k * math.log2(4001 + 40 * math.sqrt(10005)) - 1
    assert P > 0 and abs(math.log2(P) - want_bits) < 0.5


@pytest.mark.parametrize("limbs", [3, 100, 1025, 3000])
def test_recursive_inverse_close_to_floor(limbs):
    rng = np.random.default_rng(limbs)
    x = rand_rows(rng, 1, limbs)
    x[0, -1] |= 0x80000000
    L = 2 * 32 * limbs - 1
    h = of.HostArray(x)
    oy = of.OArray()
    assert of.lib().o_recursive_inverse(h.oarray(), L, ctypes.byref(oy)) == 0
    y = to_int(of.take_array(oy))
    exact = (1 << L) // to_int(x[0])
    assert abs(y - exact) <= 2


@pytest.mark.parametrize("digits", [1, 50, 1000, 5000, 20000])
def test_pi_digits_and_ret_vs_mpmath(digits):
    ret, dec, cd = of.pi_digits(digits)
    mpmath.mp.dps = digits + 50
    pistr = mpmath.nstr(mpmath.pi, digits + 30, strip_zeros=False).replace(".", "")
    assert "3" + dec == pistr[:digits + 1]
    prec = of.lib().o_prec_limbs_for_digits(digits)
    assert len(ret) == prec + 1
    mpmath.mp.prec = 32 * prec + 128
    scaled = int(mpmath.floor(mpmath.pi * mpmath.mpf(2) ** (32 * prec)))
    # RET / 2^(32 prec) is pi to well past the requested digits; above a few
    # dozen digits the difference is within the last two limbs (the
    # truncations and the approximate inverse decide those)
    err = abs(to_int(ret) - scaled)
    assert err * 10 ** digits < (1 << (32 * prec))
    if digits >= 50:
        assert err < (1 << 64)


@pytest.mark.parametrize("total,bits,words,expected", refcases.convert_cases(),
                         ids=[str(c[0]) for c in refcases.convert_cases()])
def test_convert_reference_cases(total, bits, words, expected):
    leaf = 477
    levels = (total // leaf).bit_length() - 1
    powers = (of.OArray * max(levels, 1))()
    assert of.lib().o_convert_precompute_powers(leaf, levels, powers) == 0
    h = of.HostArray(np.array([words], dtype=np.uint32))
    out = ctypes.create_string_buffer(total)
    assert of.lib().o_convert_equal_split(h.oarray(), bits, total, leaf, powers, levels, out) == 0
    assert out.raw[:total].decode() == expected
    for i in range(levels):
        of.lib().o_free(powers[i].data)


def test_sizing_matches_reference_published_lengths():
    L = of.lib()
    leaf = ctypes.c_uint32(0)
    assert L.o_round_up_decimal_digits(10 ** 9, ctypes.byref(leaf)) == refcases.README_1E9["convert_digits"]
    assert leaf.value == 477
    # README.md:72 (rounded convention) and convert_decimal_test.cu:30 (requested)
    assert L.o_prec_limbs_for_digits(refcases.README_1E9["convert_digits"]) + 1 == refcases.README_1E9["ret_len"]
    assert L.o_prec_limbs_for_digits(10 ** 9) + 1 == 103812502


def test_threaded_oracle_gives_the_same_results():
    # the CPU baseline splits the rows of every batched product over
    # threads; results must not depend on it
    ret1, dec1, _ = of.pi_digits(30000)
    of.set_threads(4)
    try:
        ret4, dec4, _ = of.pi_digits(30000)
        rng = np.random.default_rng(5)
        A = rng.integers(0, 1 << 32, size=(37, 300), dtype=np.uint64).astype(np.uint32)
        B = rng.integers(0, 1 << 32, size=(37, 200), dtype=np.uint64).astype(np.uint32)
        hc = of.HostArray.empty(37, 500)
        assert of.lib().o_mul(of.HostArray(A).oarray(), of.HostArray(B).oarray(), hc.oarray()) == 0
        got = hc.rows()
    finally:
        of.set_threads(1)
    assert np.array_equal(ret1, ret4) and dec1 == dec4
    for i in range(37):
        assert int.from_bytes(got[i].tobytes(), "little") == (int.from_bytes(A[i].tobytes(), "little") *
                                                              int.from_bytes(B[i].tobytes(), "little"))


def test_classic_cpu_baseline_digits_match_the_oracle():
    # oracle/pi_oracle --classic (the textbook GMP Chudnovsky of the CPU
    # baseline) prints the same digits as the restated pipeline
    import re
    import subprocess
    cli = os.path.join(os.path.dirname(of.__file__), "pi_oracle")
    out = subprocess.run([cli, "20000", "--classic", "--threads", "3"], capture_output=True, text=True,
                         check=True).stdout
    _, dec, _ = of.pi_digits(20000)
    assert f"DEC_hash = {of.lib().o_decimal_hash(dec.encode(), 20000)}" in out
    assert re.search(r"DEC = 31415926535897932384\.\.\.", out)
