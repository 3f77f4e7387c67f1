"""GPU: the per-op L1 entry points of include/batch_kernels.h called directly
through the C ABI (device pointers, caller-provided workspace, null stream),
bit-exact against the oracle, with padding guards around every output row
(the reference's per-op tests: batch_add_test.cu:33-35, batch_mul_ntt_test.cu
:509-577, batch_mul_naive_test.cu, batch_exactdiv_small_test.cu:279-305).
Also the array and workspace calls of batch_arith.h (create/release, and the
grow-only workspace of batch_arith_test.cu:142-152)."""
import ctypes

import numpy as np
import pytest

import gpupi_amd as g
import oracle_ffi as of
from helpers import closed_form, rand_rows, to_int
from test_oracle_ops import run_op as oracle_op

pytestmark = pytest.mark.gpu
PAD = 0xDEADBEEF


def padded_out(n, lc, sc):
    return g.DeviceBuffer.from_array(np.full(n * sc, PAD, dtype=np.uint32))


def read_rows(buf, n, lc, sc):
    full = buf.download(np.uint32, count=n * sc).reshape(n, sc)
    assert np.all(full[:, lc:] == PAD), "write past the row length"
    return full[:, :lc]


def rows_buffer(rows, stride):
    n, length = rows.shape
    host = np.full((n, stride), 0x5A5A5A5A, dtype=np.uint32)
    host[:, :length] = rows
    return g.DeviceBuffer.from_array(host.reshape(-1))


L1_ADDSUB = {
    # kind: (entry point, workspace-size function, takes B rows, takes a scalar, takes a shift)
    "add": ("batch_add_simple", "batch_add_simple_workspace_size", True, False, False),
    "sub": ("batch_sub_simple", "batch_sub_simple_workspace_size", True, False, False),
    "add_small": ("batch_add_small", "batch_add_small_workspace_size", False, True, False),
    "sub_small": ("batch_sub_small", "batch_sub_small_workspace_size", False, True, False),
    "mul_small": ("batch_mul_small", "batch_mul_small_workspace_size", False, True, False),
    "shift_add": ("batch_shift_add_simple", "batch_shift_add_simple_workspace_size", True, False, True),
    "shift_sub": ("batch_shift_sub_simple", "batch_shift_sub_simple_workspace_size", True, False, True),
}

L1_CASES = [(5, 0, 0, 1, "random"), (7, 1, 1, 2, "ones"), (300, 6, 4, 7, "random"), (6, 17, 5, 18, "top"),
            (3, 65, 1, 66, "top"), (2, 200, 150, 120, "random"), (4, 1030, 900, 1031, "top"),
            (3, 2048, 1500, 2052, "random"), (1, 100000, 99999, 100001, "ones")]


@pytest.mark.parametrize("kind", list(L1_ADDSUB))
@pytest.mark.parametrize("case", L1_CASES, ids=[f"N{c[0]}_La{c[1]}_Lb{c[2]}_Lc{c[3]}_{c[4]}" for c in L1_CASES])
def test_l1_addsub_family(kind, case):
    L = g.lib()
    entry, wsfn, has_b, has_scalar, has_shift = L1_ADDSUB[kind]
    n, la, lb, lc, mode = case
    rng = np.random.default_rng(4242)
    A = rand_rows(rng, n, la, mode)
    B = rand_rows(rng, n, lb, mode)
    sa, sb, sc = la + 3, lb + 1, lc + 2
    ba, bb = rows_buffer(A, sa), rows_buffer(B, sb)
    for shift in ([0] if not has_shift else [0, 15, 33, 1000]):
        for scalar in ([0] if not has_scalar else [1, 10005, 0xFFFFFFFF]):
            bc = padded_out(n, lc, sc)
            if has_shift:
                ws_bytes = getattr(L, wsfn)(n, la, lb, lc, shift)
            elif has_b:
                ws_bytes = getattr(L, wsfn)(n, la, lb, lc)
            else:
                ws_bytes = getattr(L, wsfn)(n, la, lc)
            ws = g.DeviceBuffer(ws_bytes)
            fn = getattr(L, entry)
            if has_shift:
                fn(ba.ptr, bb.ptr, bc.ptr, ws.ptr, shift, n, la, lb, lc, sa, sb, sc)
            elif has_b:
                fn(ba.ptr, bb.ptr, bc.ptr, ws.ptr, n, la, lb, lc, sa, sb, sc)
            else:
                fn(ba.ptr, scalar, bc.ptr, ws.ptr, n, la, lc, sa, sc)
            g.check(g.hip().hipGetLastError(), entry)
            g.synchronize()
            got = read_rows(bc, n, lc, sc)
            want = oracle_op(kind, A, B if has_b else None, scalar, shift, lc, sa, sb if has_b else None, sc)
            assert np.array_equal(got, want), (kind, case, shift, scalar)
            if n * lc <= 3000:
                for i in range(n):
                    assert to_int(got[i]) == closed_form(kind, to_int(A[i]), to_int(B[i]), scalar, shift, lc)


def test_l1_add_in_place():
    # batch_add.h:15-16: C may be A when it has C's length and stride
    L = g.lib()
    rng = np.random.default_rng(5)
    n, length = 9, 3000
    A = rand_rows(rng, n, length, "top")
    B = rand_rows(rng, n, 40, "ones")
    ba, bb = rows_buffer(A, length), rows_buffer(B, 40)
    ws = g.DeviceBuffer(L.batch_add_simple_workspace_size(n, length, 40, length))
    L.batch_add_simple(ba.ptr, bb.ptr, ba.ptr, ws.ptr, n, length, 40, length, length, 40, length)
    g.synchronize()
    got = ba.download(np.uint32).reshape(n, length)
    for i in range(n):
        assert to_int(got[i]) == (to_int(A[i]) + to_int(B[i])) % (1 << (32 * length))


@pytest.mark.parametrize("shift", [0, 1, 31, 32, 33, -1, -32, -33, -1000])
def test_l1_shift_bits(shift):
    L = g.lib()
    rng = np.random.default_rng(17)
    for n, lin, lout in [(3, 7, 12), (40, 9, 4), (1, 50000, 50010)]:
        A = rand_rows(rng, n, lin)
        ba = rows_buffer(A, lin + 1)
        bb = padded_out(n, lout, lout + 1)
        L.batch_shift_bits(ba.ptr, bb.ptr, n, lin, lout, lin + 1, lout + 1, shift)
        g.synchronize()
        got = read_rows(bb, n, lout, lout + 1)
        ho = of.HostArray.empty(n, lout)
        assert of.lib().o_shift_bits(of.HostArray(A, lin + 1).oarray(), ho.oarray(), shift) == 0
        assert np.array_equal(got, ho.rows())


@pytest.mark.parametrize("divisor", [1, 3, 0xFFFFFFFF, 830297, 156590819])
@pytest.mark.parametrize("length,n", [(1, 5), (33, 70), (128, 130)])
def test_l1_exactdiv(divisor, length, n):
    L = g.lib()
    rng = np.random.default_rng(99)
    q = rand_rows(rng, n, length)
    q[:, -1] &= 0x7FFFFFFF >> max(0, divisor.bit_length() - 1)
    prod = of.HostArray.empty(n, length + 1)
    assert of.lib().o_mul_small(of.HostArray(q).oarray(), divisor, prod.oarray()) == 0
    rows = prod.rows()[:, :length]
    buf = rows_buffer(rows, length + 1)
    host = buf.download(np.uint32).reshape(n, length + 1)
    L.batch_exactdiv_small(buf.ptr, divisor, n, length, length + 1)
    g.synchronize()
    got = buf.download(np.uint32).reshape(n, length + 1)
    assert np.array_equal(got[:, length], host[:, length]), "write past the row length"
    assert np.array_equal(got[:, :length], q)


@pytest.mark.parametrize("n,length", [(1, 1), (7, 40), (1000, 64), (3, 5000), (1, 1 << 20)])
def test_l1_length_max(n, length):
    L = g.lib()
    rng = np.random.default_rng(23)
    ws = g.DeviceBuffer(L.batch_bitlength_workspace_size(n, length))
    for mode in ("random", "sparse", "zeros"):
        A = rand_rows(rng, n, length, mode)
        if mode == "sparse":
            A[:, length // 2:] = 0
        buf = rows_buffer(A, length + 2)
        h = of.HostArray(A, length + 2)
        assert L.batch_limblength_max(buf.ptr, ws.ptr, n, length, length + 2) == of.lib().o_limblength_max(h.oarray())
        assert L.batch_bitlength_max(buf.ptr, ws.ptr, n, length, length + 2) == of.lib().o_bitlength_max(h.oarray())


def oracle_products(A, B):
    n = A.shape[0]
    hc = of.HostArray.empty(n, A.shape[1] + B.shape[1])
    assert of.lib().o_mul(of.HostArray(A).oarray(), of.HostArray(B).oarray(), hc.oarray()) == 0
    return hc.rows()


@pytest.mark.parametrize("n,la,lb,mode", [(5, 1, 1, "random"), (300, 3, 3, "ones"), (64, 16, 9, "random"),
                                          (20, 100, 77, "ones"), (7, 513, 511, "random"), (3, 2000, 1500, "top"),
                                          (2, 16384, 3, "ones")])
def test_l1_mul_naive(n, la, lb, mode):
    L = g.lib()
    rng = np.random.default_rng(31)
    A = rand_rows(rng, n, la, mode)
    B = rand_rows(rng, n, lb, mode)
    sa, sb, sr = la + 1, lb + 2, la + lb + 3
    ba, bb = rows_buffer(A, sa), rows_buffer(B, sb)
    bc = padded_out(n, la + lb, sr)
    L.batch_mul_naive(ba.ptr, bb.ptr, bc.ptr, n, la, lb, sa, sb, sr)
    g.check(g.hip().hipGetLastError(), "batch_mul_naive")
    g.synchronize()
    assert np.array_equal(read_rows(bc, n, la + lb, sr), oracle_products(A, B))


@pytest.fixture(scope="module")
def ntt_tables():
    t = g.NTTPrecomputedTables()
    g.lib().init_ntt_precomputed_tables(ctypes.byref(t))
    assert t.roots and t.max_log2 >= 28
    yield t
    g.lib().release_ntt_precomputed_tables(ctypes.byref(t))
    assert not t.roots


@pytest.mark.parametrize("n,la,lb,mode", [(3, 8, 8, "ones"), (17, 300, 212, "random"), (5, 2048, 2048, "ones"),
                                          (4, 3000, 1100, "top"), (3, 40000, 25000, "random"),
                                          (1, 1 << 20, (1 << 20) - 5, "ones"), (2, 1 << 21, 300, "random")])
def test_l1_mul_ntt(ntt_tables, n, la, lb, mode):
    # batch_mul_ntt_test.cu:509-577: products of every transform regime
    # (one tile, two levels, three levels) with strided rows
    L = g.lib()
    rng = np.random.default_rng(100)
    A = rand_rows(rng, n, la, mode)
    B = rand_rows(rng, n, lb, mode)
    sa, sb, sr = la + 5, lb + 1, la + lb + 2
    ba, bb = rows_buffer(A, sa), rows_buffer(B, sb)
    bc = padded_out(n, la + lb, sr)
    ws = g.DeviceBuffer(L.batch_mul_ntt_workspace_size(n, la, lb))
    L.batch_mul_ntt(ba.ptr, bb.ptr, bc.ptr, ws.ptr, ntt_tables, n, la, lb, sa, sb, sr)
    g.check(g.hip().hipGetLastError(), "batch_mul_ntt")
    g.synchronize()
    assert np.array_equal(read_rows(bc, n, la + lb, sr), oracle_products(A, B))


def test_array_create_and_release():
    L = g.lib()
    a = L.batch_mp_array_create(7, 100, 0)
    assert a.data and (a.length, a.batch_size, a.stride) == (100, 7, 100)
    rows = np.arange(700, dtype=np.uint32).reshape(7, 100)
    g.check(g.hip().hipMemcpy(ctypes.c_void_p(a.data), rows.ctypes.data, rows.nbytes, 1), "upload")
    assert np.array_equal(g.rows_of(a), rows)
    L.batch_mp_array_release(a)
    b = L.batch_mp_array_create(3, 10, 16)
    assert b.data and b.stride == 16
    L.batch_mp_array_release(b)
    bad = L.batch_mp_array_create(3, 10, 5)  # stride below length
    assert not bad.data and bad.length == 0 and bad.batch_size == 0
    empty = L.batch_mp_array_create(0, 10, 0)
    assert not empty.data
    L.batch_mp_array_release(empty)


def test_workspace_grows_and_never_shrinks():
    # batch_arith_test.cu:142-152: a schoolbook product leaves the workspace
    # alone, an NTT product may only grow it; ensure_workspace grows to
    # exactly the request and ignores smaller ones
    rng = np.random.default_rng(8)
    with g.Context() as c:
        assert c.workspace_size() == 0
        prev = 0
        for la, lb in [(8, 8), (100, 90), (4000, 3000), (1000, 900), (30000, 20000), (50, 40), (2000, 2000)]:
            A = rand_rows(rng, 2, la)
            B = rand_rows(rng, 2, lb)
            ba, Aa = g.upload_rows(A)
            bb, Ba = g.upload_rows(B)
            bc = g.DeviceBuffer(4 * 2 * (la + lb))
            Ca = g.BatchMPArray(bc.ptr, la + lb, 2, la + lb)
            g.check(c.mul(Aa, Ba, Ca), "mul")
            g.synchronize()
            assert np.array_equal(g.download_rows(Ca, bc), oracle_products(A, B))
            now = c.workspace_size()
            if la + lb < 384:
                assert now == prev, (la, lb)
            else:
                assert now >= prev, (la, lb)
            prev = now
        assert prev > 0
        g.check(c.ensure_workspace(prev + 4096), "ensure_workspace")
        assert c.workspace_size() == prev + 4096
        g.check(c.ensure_workspace(16), "ensure_workspace")
        assert c.workspace_size() == prev + 4096
