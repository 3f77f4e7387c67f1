"""GPU: every batched add/sub/shift/mul_small/exactdiv/shift_bits/length op
through the C ABI, bit-exact against the oracle (and the closed form), with
padding guards around every output row (batch_add_test.cu:33-35, 373-381)."""
import numpy as np
import pytest

import gpupi_amd as g
import oracle_ffi as of
from helpers import closed_form, rand_rows, to_int
from test_oracle_ops import run_op as oracle_op

pytestmark = pytest.mark.gpu
PAD = 0xDEADBEEF


def dev_op(ctx, kind, A, B, scalar, shift, lc, sa=None, sb=None, sc=None):
    n = A.shape[0]
    ba, Aa = g.upload_rows(A, sa)
    bb, Ba = g.upload_rows(B, sb) if B is not None else (None, None)
    sc = lc if sc is None else sc
    bc, Ca = g.upload_rows(np.full((n, sc), PAD, dtype=np.uint32), sc)
    Ca.length = lc
    if kind in ("add", "sub"):
        err = getattr(ctx, kind)(Aa, Ba, Ca)
    elif kind in ("add_small", "sub_small", "mul_small"):
        err = getattr(ctx, kind)(Aa, scalar, Ca)
    else:
        err = getattr(ctx, kind)(Aa, Ba, shift, Ca)
    g.check(err, kind)
    g.synchronize()
    full = bc.download(np.uint32, count=n * sc).reshape(n, sc)
    assert np.all(full[:, lc:] == PAD), "write past the row length"
    return full[:, :lc]


CASES = [
    (5, 0, 0, 1, "random"), (7, 1, 1, 2, "ones"), (9, 3, 2, 5, "random"), (300, 6, 4, 7, "random"),
    (4, 8, 8, 9, "ones"), (6, 17, 5, 18, "top"), (3, 64, 64, 65, "ones"), (3, 65, 1, 66, "top"),
    (2, 200, 150, 201, "random"), (2, 200, 150, 120, "random"), (5, 1030, 900, 1031, "top"),
    (3, 5000, 4000, 5001, "ones"), (2, 70000, 3, 70001, "top"), (1, 300000, 299999, 300001, "random"),
    (1000, 40, 33, 41, "sparse"),
    # rows of 129..512 limbs: one wave of four or eight 64-limb segments per row
    (3, 150, 120, 151, "ones"), (2, 300, 280, 301, "random"), (2, 380, 10, 381, "top"),
    (2, 440, 400, 441, "ones"), (2, 500, 500, 512, "random"),
    # rows a multiple of 4 limbs: 16-byte aligned rows, four limbs per lane
    (3, 1030, 900, 1032, "top"), (4, 5000, 4000, 5004, "ones"), (3, 2048, 1500, 2052, "random"),
]


@pytest.mark.parametrize("kind", ["add", "sub", "add_small", "sub_small", "mul_small", "shift_add", "shift_sub"])
@pytest.mark.parametrize("case", CASES, ids=[f"N{c[0]}_La{c[1]}_Lb{c[2]}_Lc{c[3]}_{c[4]}" for c in CASES])
def test_addsub_family(ctx, kind, case):
    n, la, lb, lc, mode = case
    rng = np.random.default_rng(123456789)
    A = rand_rows(rng, n, la, mode)
    B = rand_rows(rng, n, lb, mode)
    shifts = [0] if "shift" not in kind else [0, 15, 32, 255, 1000]
    scalars = [0] if "small" not in kind else [1, 10005, 0xFFFFFFFF]
    for shift in shifts:
        for scalar in scalars:
            for strides in [(None, None, None), (la + 3, lb + 1, lc + 2)]:
                got = dev_op(ctx, kind, A, B, scalar, shift, lc, *strides)
                want = oracle_op(kind, A, B, scalar, shift, lc, *strides)
                assert np.array_equal(got, want), (kind, case, shift, scalar, strides)
    if n * lc <= 4000:
        for i in range(n):
            assert to_int(got[i]) == closed_form(kind, to_int(A[i]), to_int(B[i]), scalar, shift, lc)


@pytest.mark.parametrize("length", [3, 70, 2050, 100000])
def test_long_carry_chain_across_chunks(ctx, length):
    # 0xffff...ff + 1 ripples through every chunk; 0 - 1 borrows through all
    ones = np.full((2, length), 0xFFFFFFFF, dtype=np.uint32)
    one = np.zeros((2, 1), dtype=np.uint32)
    one[:, 0] = 1
    got = dev_op(ctx, "add", ones, one, 0, 0, length + 1)
    assert np.all(got[:, :length] == 0) and np.all(got[:, length] == 1)
    zeros = np.zeros((2, length), dtype=np.uint32)
    got = dev_op(ctx, "sub", zeros, one, 0, 0, length)
    assert np.all(got == 0xFFFFFFFF)
    got = dev_op(ctx, "add_small", ones, None, 1, 0, length)
    assert np.all(got == 0)


def test_inplace_and_broadcast(ctx):
    rng = np.random.default_rng(9)
    A = rand_rows(rng, 40, 1500, "top")
    row = rand_rows(rng, 1, 700, "ones")
    ba, Aa = g.upload_rows(A)
    bb, Ba = g.upload_rows(row)
    Bb = Ba.view(batch_size=40, stride=0)
    g.check(ctx.add(Aa, Bb, Aa))
    g.check(ctx.sub(Aa, Bb, Aa))
    g.check(ctx.sub(Bb, Aa, Aa))  # C == B
    g.synchronize()
    got = g.download_rows(Aa, ba)
    r = to_int(row[0])
    for i in range(40):
        assert to_int(got[i]) == (r - to_int(A[i])) % (1 << (32 * 1500))
    # negation through a zero-length broadcast operand (pi_bs_gpu.cu:1286-1291)
    b1, X = g.upload_rows(A[:1])
    g.check(ctx.sub(X.view(length=0, stride=0), X, X))
    g.synchronize()
    assert to_int(g.download_rows(X, b1)[0]) == (-to_int(A[0])) % (1 << (32 * 1500))


def test_invalid_arguments(ctx):
    rng = np.random.default_rng(1)
    ba, A = g.upload_rows(rand_rows(rng, 4, 10))
    bb, B = g.upload_rows(rand_rows(rng, 3, 10))
    bc, C = g.upload_rows(rand_rows(rng, 4, 21))
    inv = g.HIP_ERROR_INVALID_VALUE
    assert ctx.mul(A, B, C) == inv                      # batch mismatch
    assert ctx.mul(A, A, C) == inv                      # C.length != La + Lb
    assert ctx.shift_add(A, A, 3, A) == inv             # alias
    assert ctx.exactdiv_small(A, 4) == inv              # even divisor
    assert ctx.exactdiv_small(C.view(length=129, stride=129, batch_size=1), 3) == inv
    assert ctx.mul_small(A, 3, A) == inv                # C == A
    bad = A.view(length=11)                             # length > stride
    assert ctx.add(bad, A, A) == inv


@pytest.mark.parametrize("divisor", [1, 3, 5, 0xFFFFFFFF, 0xFFFFFFFD, 0x13579BDF, 830297, 156590819])
@pytest.mark.parametrize("length,n", [(1, 5), (8, 100), (33, 70), (128, 130), (66, 5000)])
def test_exactdiv(ctx, divisor, length, n):
    rng = np.random.default_rng(0x6FDDC0FFEE123456 % (1 << 63))
    q = rand_rows(rng, n, length)
    q[:, -1] &= 0x7FFFFFFF >> max(0, divisor.bit_length() - 1)
    hq = of.HostArray(q)
    prod = of.HostArray.empty(n, length + 1)
    assert of.lib().o_mul_small(hq.oarray(), divisor, prod.oarray()) == 0
    rows = prod.rows()[:, :length]
    b, X = g.upload_rows(rows, length + 1, pad_value=PAD)
    g.check(ctx.exactdiv_small(X, divisor))
    g.synchronize()
    full = b.download(np.uint32, count=n * (length + 1)).reshape(n, length + 1)
    assert np.all(full[:, length] == PAD)
    h = of.HostArray(rows, length + 1)
    assert of.lib().o_exactdiv_small(h.oarray(), divisor) == 0
    assert np.array_equal(full[:, :length], h.rows())


@pytest.mark.parametrize("shift", [0, 1, 31, 32, 33, 100, -1, -31, -32, -33, -100, -5000])
def test_shift_bits(ctx, shift):
    rng = np.random.default_rng(11)
    for n, lin, lout in [(3, 7, 12), (50, 9, 4), (1, 100000, 100010), (2, 3000, 2500)]:
        A = rand_rows(rng, n, lin)
        ba, Aa = g.upload_rows(A, lin + 1)
        bb, Bb = g.upload_rows(np.full((n, lout + 1), PAD, dtype=np.uint32), lout + 1)
        Bb.length = lout
        g.check(ctx.shift_bits(Aa, Bb, shift))
        g.synchronize()
        full = bb.download(np.uint32, count=n * (lout + 1)).reshape(n, lout + 1)
        assert np.all(full[:, lout] == PAD)
        ho = of.HostArray.empty(n, lout)
        assert of.lib().o_shift_bits(of.HostArray(A, lin + 1).oarray(), ho.oarray(), shift) == 0
        assert np.array_equal(full[:, :lout], ho.rows())


@pytest.mark.parametrize("n,length", [(1, 1), (7, 40), (1000, 64), (5, 65), (3, 5000), (1, 1 << 20)])
def test_length_max(ctx, n, length):
    rng = np.random.default_rng(21)
    for mode in ("random", "sparse", "zeros"):
        A = rand_rows(rng, n, length, mode)
        if mode == "sparse":
            A[:, length // 2:] = 0
        b, X = g.upload_rows(A, length + 2)
        h = of.HostArray(A, length + 2)
        assert ctx.limblength_max_value(X) == of.lib().o_limblength_max(h.oarray())
        import ctypes
        r = ctypes.c_uint32(0)
        g.check(ctx.bitlength_max(X, ctypes.byref(r)))
        assert r.value == of.lib().o_bitlength_max(h.oarray())


def test_length_max_long_rows_of_different_lengths(ctx):
    # the chunked scan (rows over 16384 limbs): rows end at different
    # heights, so blocks must walk down past several all-zero chunks
    import ctypes
    rng = np.random.default_rng(22)
    n, length = 6, 200000
    for longest in (5, 2, 0):
        A = rand_rows(rng, n, length)
        for i in range(n):
            top = length - 30000 * (i if i != longest else 0) - 7 * i - (0 if i == longest else 50000)
            A[i, max(top, 0):] = 0
        b, X = g.upload_rows(A, length + 3)
        h = of.HostArray(A, length + 3)
        assert ctx.limblength_max_value(X) == of.lib().o_limblength_max(h.oarray())
        r = ctypes.c_uint32(0)
        g.check(ctx.bitlength_max(X, ctypes.byref(r)))
        assert r.value == of.lib().o_bitlength_max(h.oarray())


@pytest.mark.parametrize("L_limb,L_dec,n", [(1, 9, 3), (2, 18, 5), (51, 477, 300), (52, 489, 70),
                                             (61, 576, 10), (65, 600, 4), (2047, 19000, 2), (300, 20000, 3)])
def test_decimal_naive(ctx, L_limb, L_dec, n):
    import ctypes
    rng = np.random.default_rng(3)
    A = rand_rows(rng, n, L_limb)
    A[0] = np.where(np.arange(L_limb) % 2 == 0, 0, 0xFFFFFFFF)
    ba, Aa = g.upload_rows(A, L_limb + 1)
    out = g.DeviceBuffer(n * (L_dec + 3))
    out.fill(ord("?"))
    g.lib().batch_decimal_naive(Aa.data, out.ptr, n, L_limb, L_limb + 1, L_dec, L_dec + 3)
    g.synchronize()
    got = out.download(np.uint8).reshape(n, L_dec + 3)
    assert np.all(got[:, L_dec:] == ord("?"))
    want = np.zeros(n * L_dec, dtype=np.uint8)
    Ah = np.ascontiguousarray(A)
    of.lib().o_decimal_naive(Ah.ctypes.data, want.ctypes.data, n, L_limb, L_limb, L_dec, L_dec)
    assert np.array_equal(got[:, :L_dec], want.reshape(n, L_dec))
