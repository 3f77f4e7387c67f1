"""CPU: pin the oracle's batched ops to Python big-integer closed forms.

Shapes follow the reference's own op tests: fixed and odd strides, long carry
chains (runs of 0xffffffff), zero lengths, broadcast rows (stride 0), and the
exact-division divisors of batch_exactdiv_small_test.cu:279-305."""
import numpy as np
import pytest

import oracle_ffi as of
from helpers import closed_form, rand_rows, to_int

PAD = 0xDEADBEEF

CASES = [
    # (N, La, Lb, Lc, mode)
    (5, 0, 0, 1, "random"), (7, 1, 1, 1, "random"), (7, 1, 1, 2, "ones"), (9, 3, 2, 5, "random"),
    (4, 8, 8, 9, "ones"), (6, 17, 5, 18, "top"), (3, 64, 64, 65, "ones"), (3, 65, 1, 66, "top"),
    (2, 200, 150, 201, "random"), (2, 200, 150, 120, "random"), (3, 40, 70, 72, "sparse"),
]


def run_op(kind, A, B, scalar, shift, lc, sa=None, sb=None, sc=None):
    L = of.lib()
    n = A.shape[0]
    ha = of.HostArray(A, sa)
    hb = of.HostArray(B, sb) if B is not None else None
    hc = of.HostArray.empty(n, lc, sc, pad=PAD)
    if kind in ("add", "sub"):
        st = getattr(L, "o_" + kind)(ha.oarray(), hb.oarray(), hc.oarray())
    elif kind in ("add_small", "sub_small", "mul_small"):
        st = getattr(L, "o_" + kind)(ha.oarray(), scalar, hc.oarray())
    else:
        st = getattr(L, "o_" + kind)(ha.oarray(), hb.oarray(), shift, hc.oarray())
    assert st == 0
    stride = hc.stride
    for i in range(n):  # padding guard (batch_add_test.cu:33-35)
        assert np.all(hc.buf[i * stride + lc:(i + 1) * stride] == PAD)
    return hc.rows()


@pytest.mark.parametrize("kind", ["add", "sub", "add_small", "sub_small", "mul_small", "shift_add", "shift_sub"])
@pytest.mark.parametrize("case", CASES)
def test_addsub_family_matches_closed_form(kind, case):
    n, la, lb, lc, mode = case
    rng = np.random.default_rng(123456789)
    A = rand_rows(rng, n, la, mode)
    B = rand_rows(rng, n, lb, mode)
    for shift in ([0] if "shift" not in kind else [0, 1, 31, 32, 33, 255, 1000]):
        for scalar in ([0] if "small" not in kind else [0, 1, 10005, 0xFFFFFFFF]):
            for sa, sb, sc in [(None, None, None), (la + 3, lb + 1, lc + 2)]:
                got = run_op(kind, A, B, scalar, shift, lc, sa, sb, sc)
                for i in range(n):
                    want = closed_form(kind, to_int(A[i]), to_int(B[i]), scalar, shift, lc)
                    assert to_int(got[i]) == want, (kind, case, shift, scalar, i)


def test_broadcast_and_inplace():
    L = of.lib()
    rng = np.random.default_rng(7)
    A = rand_rows(rng, 4, 9, "random")
    row = rand_rows(rng, 1, 6, "ones")
    ha = of.HostArray(A)
    hb = of.HostArray(row)
    # broadcast B (stride 0) and in-place C == A
    assert L.o_add(ha.oarray(), hb.oarray(batch=4, stride=0), ha.oarray()) == 0
    for i in range(4):
        assert to_int(ha.rows()[i]) == (to_int(A[i]) + to_int(row[0])) % (1 << (32 * 9))
    # shift ops reject aliasing (batch_arith.cu:286)
    assert L.o_shift_add(ha.oarray(), hb.oarray(batch=4, stride=0), 3, ha.oarray()) == -1


@pytest.mark.parametrize("la,lb", [(1, 1), (3, 3), (4, 6), (17, 5), (64, 64), (100, 37), (513, 512), (1, 700)])
def test_mul(la, lb):
    rng = np.random.default_rng(100)
    n = 3
    A = rand_rows(rng, n, la)
    B = rand_rows(rng, n, lb)
    got = of.HostArray.empty(n, la + lb, la + lb + 2, pad=PAD)
    assert of.lib().o_mul(of.HostArray(A, la + 1).oarray(), of.HostArray(B).oarray(), got.oarray()) == 0
    for i in range(n):
        assert to_int(got.rows()[i]) == to_int(A[i]) * to_int(B[i])
    # C.length must be La + Lb (batch_arith.cu:177)
    bad = of.HostArray.empty(n, la + lb + 1)
    assert of.lib().o_mul(of.HostArray(A).oarray(), of.HostArray(B).oarray(), bad.oarray()) == -1


@pytest.mark.parametrize("divisor", [1, 3, 5, 0xFFFFFFFF, 0xFFFFFFFD, 0x13579BDF, 830297, 156590819])
@pytest.mark.parametrize("length", [1, 2, 7, 8, 33, 128])
def test_exactdiv(divisor, length):
    rng = np.random.default_rng(0x6FDDC0FFEE123456 % (1 << 63))
    n = 5
    rows = []
    quots = []
    for _ in range(n):
        q = int(rng.integers(1, 1 << 62)) << (32 * length - 64) if length > 2 else int(rng.integers(1, 1 << 30))
        q %= (1 << (32 * length)) // divisor
        quots.append(q)
        rows.append([(q * divisor >> (32 * k)) & 0xFFFFFFFF for k in range(length)])
    h = of.HostArray(np.array(rows, dtype=np.uint32), length + 1, pad=PAD)
    assert of.lib().o_exactdiv_small(h.oarray(), divisor) == 0
    for i in range(n):
        assert to_int(h.rows()[i]) == quots[i]
    # not exact -> flagged; even divisor / too long -> invalid (batch_arith.cu:229-231)
    if divisor > 1:
        h2 = of.HostArray(np.array([[1] + [0] * (length - 1)], dtype=np.uint32))
        assert of.lib().o_exactdiv_small(h2.oarray(), divisor) == -2
    assert of.lib().o_exactdiv_small(h.oarray(), 2) == -1


@pytest.mark.parametrize("shift", [0, 1, 5, 31, 32, 33, 64, 100, -1, -31, -32, -33, -100])
def test_shift_bits(shift):
    rng = np.random.default_rng(11)
    A = rand_rows(rng, 3, 7)
    for lout in (1, 7, 12):
        out = of.HostArray.empty(3, lout, pad=PAD)
        assert of.lib().o_shift_bits(of.HostArray(A).oarray(), out.oarray(), shift) == 0
        for i in range(3):
            a = to_int(A[i])
            want = (a << shift) if shift >= 0 else (a >> -shift)
            assert to_int(out.rows()[i]) == want % (1 << (32 * lout))


def test_lengths():
    rng = np.random.default_rng(5)
    A = rand_rows(rng, 6, 40, "sparse")
    A[:, 30:] = 0
    A[2, 29] = 1
    h = of.HostArray(A, 41)
    assert of.lib().o_limblength_max(h.oarray()) == 30
    assert of.lib().o_bitlength_max(h.oarray()) == max(to_int(r).bit_length() for r in A)
    assert of.lib().o_limblength_max(of.HostArray.empty(3, 5).oarray()) == 0


@pytest.mark.parametrize("L_limb,L_dec", [(1, 9), (1, 10), (2, 18), (51, 477), (52, 489), (3, 40)])
def test_decimal_naive(L_limb, L_dec):
    rng = np.random.default_rng(3)
    A = rand_rows(rng, 4, L_limb)
    A[1] = np.where(np.arange(L_limb) % 2 == 0, 0, 0xFFFFFFFF)  # batch_decimal_naive_test.cu:151
    out = np.zeros(4 * L_dec, dtype=np.uint8)
    of.lib().o_decimal_naive(A.ctypes.data, out.ctypes.data, 4, L_limb, L_limb, L_dec, L_dec)
    for i in range(4):
        x = to_int(A[i])
        want = str((x * 10 ** L_dec) >> (32 * L_limb)).zfill(L_dec)
        got = out[i * L_dec:(i + 1) * L_dec].tobytes().decode()[::-1]  # least significant first
        assert got == want
