"""GPU: batch_mp_mul_noalloc (the multiply that allocates nothing) against
batch_mp_mul on the same inputs, bit for bit, on calls in a row that share
the static scratch (the second smaller than the first), in chunks, with a
broadcast operand, and the rejection of a row larger than the scratch."""
import ctypes

import numpy as np
import pytest

import gpupi_amd as g
import oracle_ffi as of
from helpers import rand_rows

pytestmark = pytest.mark.gpu
PAD = 0xDEADBEEF
SCRATCH_BYTES = 128 << 20


def free_bytes():
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    g.check(g.hip().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)), "hipMemGetInfo")
    return free.value


def product_pair(c, A, B, b_broadcast=False):
    """(noalloc result, batch_mp_mul result) for the rows of A and B."""
    n = A.shape[0]
    la, lb = A.shape[1], B.shape[1]
    ba, Aa = g.upload_rows(A)
    bb, Ba = g.upload_rows(B)
    if b_broadcast:
        Ba = Ba.view(batch_size=n, stride=0)
    outs = []
    for fn in (c.mul_noalloc, c.mul):
        bc, Ca = g.upload_rows(np.full((n, la + lb + 1), PAD, dtype=np.uint32), la + lb + 1)
        Ca.length = la + lb
        g.check(fn(Aa, Ba, Ca), "multiply")
        g.synchronize()
        full = bc.download(np.uint32, count=n * (la + lb + 1)).reshape(n, la + lb + 1)
        assert np.all(full[:, la + lb] == PAD), "write past the row length"
        outs.append(full[:, :la + lb])
    return outs


def test_noalloc_matches_mul_on_calls_in_a_row():
    rng = np.random.default_rng(77)
    with g.Context() as c:
        # first call: three rows of a 2^19-point transform; second, smaller:
        # five rows of 2^16 points; both through the one scratch array
        for n, la, lb, mode in [(3, 300000, 200000, "random"), (5, 20000, 15000, "ones"), (2, 700, 500, "top")]:
            A = rand_rows(rng, n, la, mode)
            B = rand_rows(rng, n, lb, mode)
            got, want = product_pair(c, A, B)
            assert np.array_equal(got, want), (n, la, lb)
        hc = of.HostArray.empty(2, 1200)
        assert of.lib().o_mul(of.HostArray(A).oarray(), of.HostArray(B).oarray(), hc.oarray()) == 0
        assert np.array_equal(got, hc.rows())


def test_noalloc_runs_rows_in_chunks_and_allocates_nothing():
    rng = np.random.default_rng(78)
    with g.Context() as c:
        # eight rows of 2^20 points need ~256 MiB of planes: several chunks
        A = rand_rows(rng, 8, 600000)
        B = rand_rows(rng, 8, 400000)
        got, want_big = product_pair(c, A, B)
        assert np.array_equal(got, want_big)
        # a broadcast B (one row at stride 0) against many rows
        A2 = rand_rows(rng, 40, 9000, "ones")
        B2 = rand_rows(rng, 1, 7000, "ones")
        got, want = product_pair(c, A2, B2, b_broadcast=True)
        assert np.array_equal(got, want)
        # with the operands in place, the calls leave device memory as it was
        ws_before = c.workspace_size()
        ba, Aa = g.upload_rows(A)
        bb, Ba = g.upload_rows(B)
        bc = g.DeviceBuffer(4 * 8 * 1000000)
        Ca = g.BatchMPArray(bc.ptr, 1000000, 8, 1000000)
        g.synchronize()
        before = free_bytes()
        for _ in range(3):
            g.check(c.mul_noalloc(Aa, Ba, Ca), "batch_mp_mul_noalloc")
        g.synchronize()
        assert free_bytes() == before
        assert c.workspace_size() == ws_before
        assert np.array_equal(bc.download(np.uint32).reshape(8, 1000000), want_big)


def test_noalloc_rejects_rows_beyond_the_scratch():
    rng = np.random.default_rng(79)
    with g.Context() as c:
        A = rand_rows(rng, 1, 1500000)
        B = rand_rows(rng, 1, 1000000)
        ba, Aa = g.upload_rows(A)
        bb, Ba = g.upload_rows(B)
        bc, Ca = g.upload_rows(np.full((1, 2500000), PAD, dtype=np.uint32))
        assert c.mul_noalloc(Aa, Ba, Ca) == g.HIP_ERROR_INVALID_VALUE
        g.synchronize()
        assert np.all(bc.download(np.uint32) == PAD)
        # the same validation as batch_mp_mul
        bd, Da = g.upload_rows(rand_rows(rng, 2, 10))
        assert c.mul_noalloc(Aa, Da, Ca) == g.HIP_ERROR_INVALID_VALUE
