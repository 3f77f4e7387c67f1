"""GPU: batched products (schoolbook kernels and the two-prime binary64 NTT)
bit-exact against GMP's mpn_mul through the oracle.

Shapes: the schoolbook size classes, every NTT plan class (one tile; column +
row; two columns + row), the 1024-limb boundary of the reference
(batch_arith_test.cu:99-103), random shapes with random strides in the style
of batch_mul_ntt_test.cu:531-556, and broadcast operands."""
import numpy as np
import pytest

import gpupi_amd as g
import oracle_ffi as of
from helpers import rand_rows

pytestmark = pytest.mark.gpu
PAD = 0xDEADBEEF


def oracle_mul(A, B, sa=None, sb=None):
    n = A.shape[0]
    hc = of.HostArray.empty(n, A.shape[1] + B.shape[1])
    assert of.lib().o_mul(of.HostArray(A, sa).oarray(), of.HostArray(B, sb).oarray(), hc.oarray()) == 0
    return hc.rows()


def device_mul(ctx, A, B, sa=None, sb=None, sc_extra=1, b_broadcast=False):
    n = A.shape[0]
    lc = A.shape[1] + B.shape[1]
    ba, Aa = g.upload_rows(A, sa)
    bb, Ba = g.upload_rows(B if not b_broadcast else B[:1], sb)
    if b_broadcast:
        Ba = Ba.view(batch_size=n, stride=0)
    sc = lc + sc_extra
    bc, Ca = g.upload_rows(np.full((n, sc), PAD, dtype=np.uint32), sc)
    Ca.length = lc
    g.check(ctx.mul(Aa, Ba, Ca), "mul")
    g.synchronize()
    full = bc.download(np.uint32, count=n * sc).reshape(n, sc)
    assert np.all(full[:, lc:] == PAD), "write past the product"
    return full[:, :lc]


SCHOOLBOOK = [(1000, 1, 1), (777, 3, 3), (500, 4, 6), (300, 16, 9), (64, 100, 16), (40, 17, 17),
              (33, 64, 64), (9, 200, 150), (5, 512, 512), (3, 1000, 24), (2, 5000, 3000), (2, 8192, 8192)]
NTT = [(1, 8, 8), (100, 9, 7), (64, 32, 32), (17, 100, 100), (9, 513, 512), (5, 1024, 1024), (3, 2048, 2047),
       (2, 2049, 2047), (4, 3000, 1000), (3, 4097, 4095), (2, 10000, 9000), (2, 65536, 65536),
       (1, 300000, 200000), (1, 1048576, 1048576), (3, 2097152, 1), (1, 2097153, 2097152)]


@pytest.mark.parametrize("n,la,lb", SCHOOLBOOK, ids=[f"N{n}_{a}x{b}" for n, a, b in SCHOOLBOOK])
def test_schoolbook(ctx, n, la, lb):
    g.check(ctx.set_ntt_threshold(1 << 30))
    try:
        rng = np.random.default_rng(100)
        for mode in ("random", "ones"):
            A = rand_rows(rng, n, la, mode)
            B = rand_rows(rng, n, lb, mode)
            assert np.array_equal(device_mul(ctx, A, B, la + 3, lb + 1), oracle_mul(A, B, la + 3, lb + 1))
    finally:
        g.check(ctx.set_ntt_threshold(384))


@pytest.mark.parametrize("n,la,lb", NTT, ids=[f"N{n}_{a}x{b}" for n, a, b in NTT])
def test_ntt(ctx, n, la, lb):
    g.check(ctx.set_ntt_threshold(2))
    try:
        rng = np.random.default_rng(1234)
        for mode in ("random", "ones"):
            A = rand_rows(rng, n, la, mode)
            B = rand_rows(rng, n, lb, mode)
            assert np.array_equal(device_mul(ctx, A, B), oracle_mul(A, B)), mode
    finally:
        g.check(ctx.set_ntt_threshold(384))


SQUARES = [(5, 900, None), (3, 20000, None), (2, 20000, 20003), (1, 3000000, None), (4, 3000, 0)]


@pytest.mark.parametrize("n,la,stride", SQUARES, ids=[f"N{n}_{a}_s{s}" for n, a, s in SQUARES])
def test_ntt_square(ctx, n, la, stride):
    # A * A through one operand (the library transforms it once): one tile,
    # two and three levels, rows at a stride, and one row at stride 0
    g.check(ctx.set_ntt_threshold(2))
    try:
        rng = np.random.default_rng(la + n)
        for mode in ("random", "ones"):
            A = rand_rows(rng, 1 if stride == 0 else n, la, mode)
            buf, Aa = g.upload_rows(A, None if stride in (None, 0) else stride)
            if stride == 0:
                Aa = Aa.view(batch_size=n, stride=0)
                A = np.repeat(A, n, axis=0)
            lc = 2 * la
            bc, Ca = g.upload_rows(np.full((n, lc + 1), PAD, dtype=np.uint32), lc + 1)
            Ca.length = lc
            g.check(ctx.mul(Aa, Aa, Ca), "mul")
            g.synchronize()
            full = bc.download(np.uint32, count=n * (lc + 1)).reshape(n, lc + 1)
            assert np.all(full[:, lc:] == PAD), "write past the product"
            assert np.array_equal(full[:, :lc], oracle_mul(A, A)), mode
    finally:
        g.check(ctx.set_ntt_threshold(384))


ALIGNED = [(4, 3000, 2900), (3, 40000, 30000), (2, 300000, 200000), (1, 1 << 20, 1 << 20)]


@pytest.mark.parametrize("n,la,lb", ALIGNED, ids=[f"N{n}_{a}x{b}" for n, a, b in ALIGNED])
def test_ntt_aligned_rows(ctx, n, la, lb):
    # output rows 16-byte aligned (stride a multiple of 4 limbs): the final
    # carry pass runs four limbs per lane; all-ones operands make carry
    # chains cross its 1024-limb chunks
    g.check(ctx.set_ntt_threshold(2))
    try:
        rng = np.random.default_rng(la)
        for mode in ("random", "ones", "top"):
            A = rand_rows(rng, n, la, mode)
            B = rand_rows(rng, n, lb, mode)
            extra = (-(la + lb)) % 4 + 4
            assert np.array_equal(device_mul(ctx, A, B, sc_extra=extra), oracle_mul(A, B)), mode
    finally:
        g.check(ctx.set_ntt_threshold(384))


def test_boundary_shapes(ctx):
    # batch_arith_test.cu:99-103: 512+512 (naive there), 513+512, 2200+2200
    rng = np.random.default_rng(7)
    for la, lb in [(512, 512), (513, 512), (2200, 2200)]:
        A = rand_rows(rng, 6, la)
        B = rand_rows(rng, 6, lb)
        assert np.array_equal(device_mul(ctx, A, B), oracle_mul(A, B))


def test_random_shapes_and_strides(ctx):
    # batch_mul_ntt_test.cu:531-556 style: log-uniform lengths, random split, strides L+1..8
    rng = np.random.default_rng(100)
    for _ in range(24):
        total = int(2 ** rng.uniform(1, 15))
        la = max(1, int(rng.integers(1, total + 1)))
        lb = max(1, total - la) if total > la else 1
        n = max(1, int(50000 // (la + lb)))
        n = min(n, 200)
        sa = la + int(rng.integers(0, 8))
        sb = lb + int(rng.integers(0, 8))
        A = rand_rows(rng, n, la)
        B = rand_rows(rng, n, lb)
        assert np.array_equal(device_mul(ctx, A, B, sa, sb, sc_extra=int(rng.integers(1, 5))),
                              oracle_mul(A, B, sa, sb)), (n, la, lb)


def test_broadcast_operand(ctx):
    # convert_decimal multiplies every row by one power (stride 0); above
    # one tile the power's last level is transformed once for every row
    rng = np.random.default_rng(3)
    for n, la, lb, mode in [(33, 10, 7, "random"), (33, 300, 300, "random"), (33, 1500, 2500, "random"),
                            (33, 9000, 7000, "random"), (5, 70000, 60000, "ones"), (3, 600000, 400000, "random")]:
        A = rand_rows(rng, n, la, mode)
        B = np.repeat(rand_rows(rng, 1, lb, mode), n, axis=0)
        assert np.array_equal(device_mul(ctx, A, B, b_broadcast=True), oracle_mul(A, B)), (n, la, lb)


def every_size_case(k):
    # La + Lb in (2^(k-1), 2^k]: transform size 2^k, unequal operands
    total = (1 << k) - (k % 3)
    la = total // 2 + k
    return max(1, min(64, (1 << 22) >> k)), la, total - la


EVERY_K = list(range(4, 25))
SLOW_K = [25, 26, 27]


@pytest.mark.parametrize("k", EVERY_K + [pytest.param(k, marks=pytest.mark.slow) for k in SLOW_K],
                         ids=[f"K2^{k}" for k in EVERY_K + SLOW_K])
def test_ntt_every_transform_size(ctx, k):
    # one case per transform size: every in-tile DFT size and level split
    g.check(ctx.set_ntt_threshold(2))
    try:
        n, la, lb = every_size_case(k)
        rng = np.random.default_rng(k)
        A = rand_rows(rng, n, la)
        B = rand_rows(rng, n, lb)
        assert np.array_equal(device_mul(ctx, A, B), oracle_mul(A, B)), (n, la, lb)
    finally:
        g.check(ctx.set_ntt_threshold(384))


def device_mul_shared(ctx, A1, A2, B, b_rows=None, b_broadcast=False):
    # C1 = A1 B, C2 = A2 B through one batch_mp_mul_shared call; B has
    # b_rows >= max(rows) rows, or one row at stride 0
    n1, n2 = A1.shape[0], A2.shape[0]
    lb = B.shape[1]
    b1, A1a = g.upload_rows(A1, A1.shape[1] + 2)
    b2, A2a = g.upload_rows(A2, A2.shape[1] + 1)
    bb, Ba = g.upload_rows(B[:1] if b_broadcast else B, lb + 3)
    if b_broadcast:
        Ba = Ba.view(batch_size=max(n1, n2), stride=0)
    outs = []
    arrs = []
    for n, la in ((n1, A1.shape[1]), (n2, A2.shape[1])):
        lc = la + lb
        bc, Ca = g.upload_rows(np.full((n, lc + 1), PAD, dtype=np.uint32), lc + 1)
        Ca.length = lc
        outs.append((bc, n, lc))
        arrs.append(Ca)
    g.check(ctx.mul_shared(A1a, A2a, Ba, arrs[0], arrs[1]), "mul_shared")
    g.synchronize()
    res = []
    for bc, n, lc in outs:
        full = bc.download(np.uint32, count=n * (lc + 1)).reshape(n, lc + 1)
        assert np.all(full[:, lc:] == PAD), "write past the product"
        res.append(full[:, :lc])
    return res


SHARED = [  # (N1, L1, N2, L2, rows of B, L_b)
    (7, 3000, 8, 4000, 8, 2500),          # two levels, one transform size, N1 < N2
    (3, 300000, 3, 290000, 4, 240000),    # two levels, one row of B unused
    (2, 2500000, 3, 2200000, 3, 2000000),  # three levels
    (1, 40000, 2, 30000, 2, 20000),
    (2, 100000, 3, 90000, 3, 80000),      # K = 2^18: last level 2^10
    (2, 200000, 2, 180000, 2, 150000),    # K = 2^19: last level 2^11
    (1, 10000000, 1, 9000000, 1, 8000000),  # K = 2^25: three levels, last 2^11
    (5, 100, 5, 90, 5, 80),               # schoolbook: two plain products
    (4, 1000, 4, 5000, 4, 700),           # different transform sizes: two plain products
    (6, 600, 6, 700, 6, 900),             # one tile each: two plain products
]


@pytest.mark.parametrize("n1,l1,n2,l2,nb,lb", SHARED, ids=[f"{a}x{b}_{c}x{d}_B{e}x{f}" for a, b, c, d, e, f in SHARED])
def test_mul_shared(ctx, n1, l1, n2, l2, nb, lb):
    rng = np.random.default_rng(n1 * 31 + l1)
    A1 = rand_rows(rng, n1, l1)
    A2 = rand_rows(rng, n2, l2)
    B = rand_rows(rng, nb, lb)
    C1, C2 = device_mul_shared(ctx, A1, A2, B)
    assert np.array_equal(C1, oracle_mul(A1, B[:n1]))
    assert np.array_equal(C2, oracle_mul(A2, B[:n2]))


def test_mul_shared_broadcast_and_tree_shape(ctx):
    rng = np.random.default_rng(77)
    # B one row at stride 0
    A1 = rand_rows(rng, 5, 9000)
    A2 = rand_rows(rng, 4, 7000)
    B = rand_rows(rng, 1, 6000)
    C1, C2 = device_mul_shared(ctx, A1, A2, B, b_broadcast=True)
    assert np.array_equal(C1, oracle_mul(A1, np.repeat(B, 5, axis=0)))
    assert np.array_equal(C2, oracle_mul(A2, np.repeat(B, 4, axis=0)))
    # carry chains: all-ones operands
    A1 = rand_rows(rng, 2, 50000, "ones")
    A2 = rand_rows(rng, 3, 60000, "ones")
    B = rand_rows(rng, 3, 40000, "ones")
    C1, C2 = device_mul_shared(ctx, A1, A2, B)
    assert np.array_equal(C1, oracle_mul(A1, B[:2]))
    assert np.array_equal(C2, oracle_mul(A2, B))


def test_shared_and_broadcast_across_chunks(ctx):
    # more rows than one launch chunk holds (rows * K > 2^28): the shared
    # operand's transform per chunk, and a broadcast operand transformed once
    # in the first chunk and reused by the next
    rng = np.random.default_rng(11)
    n, la, lb = 4100, 40000, 25000  # K = 2^16: 4096 rows per chunk
    A1 = rand_rows(rng, n, la)
    A2 = rand_rows(rng, n - 3, la - 7000)
    B = rand_rows(rng, n, lb)
    C1, C2 = device_mul_shared(ctx, A1, A2, B)
    assert np.array_equal(C1, oracle_mul(A1, B))
    assert np.array_equal(C2, oracle_mul(A2, B[:n - 3]))
    Bb = rand_rows(rng, 1, lb)
    C1, C2 = device_mul_shared(ctx, A1, A2, Bb, b_broadcast=True)
    assert np.array_equal(C1, oracle_mul(A1, np.repeat(Bb, n, axis=0)))
    assert np.array_equal(C2, oracle_mul(A2, np.repeat(Bb, n - 3, axis=0)))


def test_mul_shared_rejects_bad_shapes(ctx):
    rng = np.random.default_rng(3)
    A = rand_rows(rng, 4, 5000)
    B = rand_rows(rng, 2, 5000)  # fewer rows than the products
    ba, Aa = g.upload_rows(A)
    bb, Ba = g.upload_rows(B)
    bc1, C1 = g.upload_rows(np.zeros((4, 10000), dtype=np.uint32))
    bc2, C2 = g.upload_rows(np.zeros((4, 10000), dtype=np.uint32))
    assert ctx.mul_shared(Aa, Aa, Ba, C1, C2) != 0
    assert ctx.mul_shared(Aa, Aa, Aa, C1, C1) != 0  # outputs alias each other


def valid_splits(k):
    # every column split the kernels cover: (m0, m1) with m1 = 0 for two levels
    out = []
    for m0 in range(5, 10):
        for m1 in ([0] if k <= 21 else range(5, 10)):
            rest = k - m0 - m1
            if 6 <= rest <= 12:
                out.append((m0, m1))
    return out


SPLIT_K = [13, 16, 18, 20, 22, 23, 24]


@pytest.mark.parametrize("k", SPLIT_K, ids=[f"K2^{k}" for k in SPLIT_K])
def test_ntt_every_split(ctx, k):
    # the same product under every level split: bit-identical to the
    # built-in plan's product and to GMP
    g.check(ctx.set_ntt_threshold(2))
    default = g.ntt_get_split(k)
    try:
        n, la, lb = every_size_case(k)
        n = min(n, 4)
        rng = np.random.default_rng(100 + k)
        A = rand_rows(rng, n, la)
        B = rand_rows(rng, n, lb)
        want = oracle_mul(A, B)
        splits = valid_splits(k)
        assert len(splits) >= 2
        for m0, m1 in splits:
            g.ntt_set_split(k, m0, m1)
            levels = g.ntt_get_split(k)
            assert levels[0] == m0 and sum(levels) == k and (m1 == 0 or levels[1] == m1)
            assert np.array_equal(device_mul(ctx, A, B), want), (m0, m1)
    finally:
        g.ntt_set_split(k, 0)
        g.check(ctx.set_ntt_threshold(384))
    assert g.ntt_get_split(k) == default


SPLIT = [  # one row with L + 2 S <= K: the longer operand runs as two pieces at K / 2
    (3000000, 500000, "random"),    # K = 2^22
    (500000, 3000000, "random"),    # the longer operand second
    (2200000, 900000, "ones"),      # carries through the seam
    (6000000, 1000001, "top"),      # K = 2^23
]
SPLIT_MANY = [  # one row with S far shorter than L: many pieces at a small transform
    (4194309, 3000, "random"),      # K = 2^23, pieces of 2^16 - 2^17 points
    (3000, 4194309, "ones"),        # the longer operand second, carries through every seam
    (1500000, 40000, "top"),        # K = 2^21
    (1048600, 200000, "random"),    # K = 2^21, a few pieces
    (2500001, 70001, "sparse"),
]


@pytest.mark.parametrize("la,lb,mode", SPLIT, ids=[f"{a}x{b}_{m}" for a, b, m in SPLIT])
def test_unbalanced_product_splits_and_stays_exact(ctx, la, lb, mode):
    rng = np.random.default_rng(la ^ lb)
    A = rand_rows(rng, 1, la, mode)
    B = rand_rows(rng, 1, lb, mode)
    want = oracle_mul(A, B)
    g.kernel_stats_reset(True)
    try:
        got = device_mul(ctx, A, B, sa=la + 3, sb=lb + 1)
        sizes = {name.split("@2^")[1] for name, *_ in g.kernel_stats(by_size=True)}
    finally:
        g.kernel_stats_reset(False)
    k = (la + lb - 1).bit_length()
    assert str(k - 1) in sizes and str(k) not in sizes, sizes
    assert np.array_equal(got, want)
    # the plain form keeps the reference's one product at K points
    with g.Context() as plain:
        plain.set_form(g.FORM_PLAIN)
        assert np.array_equal(device_mul(plain, A, B), want)


@pytest.mark.parametrize("la,lb,mode", SPLIT_MANY, ids=[f"{a}x{b}_{m}" for a, b, m in SPLIT_MANY])
def test_very_unbalanced_product_runs_in_pieces_and_stays_exact(ctx, la, lb, mode):
    rng = np.random.default_rng(la + 7 * lb)
    A = rand_rows(rng, 1, la, mode)
    B = rand_rows(rng, 1, lb, mode)
    want = oracle_mul(A, B)
    g.kernel_stats_reset(True)
    try:
        got = device_mul(ctx, A, B, sa=la + 1, sb=lb + 2)
        sizes = {int(name.split("@2^")[1]) for name, *_ in g.kernel_stats(by_size=True)}
    finally:
        g.kernel_stats_reset(False)
    k = (la + lb - 1).bit_length()
    assert sizes and max(sizes) <= k - 2, sizes
    assert np.array_equal(got, want)
    # the same pairs through the ragged call (pieces too) and mixed with a
    # pair that stays whole
    C = rand_rows(rng, 1, 5000)
    D = rand_rows(rng, 1, 4000)
    bufs = [g.DeviceBuffer.from_array(x[0]) for x in (A, B, C, D)]
    outs = [g.DeviceBuffer.from_array(np.full(n + 1, PAD, dtype=np.uint32)) for n in (la + lb, 9000)]
    g.check(ctx.mul_ragged([(bufs[0].ptr, bufs[1].ptr, outs[0].ptr, la, lb),
                            (bufs[2].ptr, bufs[3].ptr, outs[1].ptr, 5000, 4000)]), "mul_ragged")
    g.synchronize()
    o0, o1 = outs[0].download(np.uint32), outs[1].download(np.uint32)
    assert o0[-1] == PAD and o1[-1] == PAD
    assert np.array_equal(o0[:-1], want[0])
    assert np.array_equal(o1[:-1], oracle_mul(C, D)[0])


LAUNCH_K = [10, 16, 20, 23, 24]


@pytest.mark.parametrize("k", LAUNCH_K, ids=[f"K2^{k}" for k in LAUNCH_K])
def test_ntt_every_launch_shape(ctx, k):
    # every tile-pairing mask (each column level kind paired or not) and
    # XCD chunks from the plain order to 256 tiles: bit-identical to GMP;
    # K = 2^23 and 2^24 have 2^7- and 2^8-point second levels, so every kind
    # meets both block shapes
    g.check(ctx.set_ntt_threshold(2))
    builtin = g.ntt_get_launch()
    try:
        n, la, lb = every_size_case(k)
        n = min(n, 4)
        rng = np.random.default_rng(300 + k)
        A = rand_rows(rng, n, la)
        B = rand_rows(rng, n, lb)
        want = oracle_mul(A, B)
        for mask in range(16):
            for chunk in ((0, 3, 5, 8) if mask in (0, 15) else (5,)):
                g.ntt_set_launch(mask, chunk)
                assert g.ntt_get_launch() == (mask, chunk)
                assert np.array_equal(device_mul(ctx, A, B), want), (mask, chunk)
    finally:
        g.ntt_set_launch(*builtin)
        g.check(ctx.set_ntt_threshold(384))
    assert g.ntt_get_launch() == builtin


def test_shared_products_every_launch_shape(ctx):
    # the paired last level (two products sharing B) under the extreme shapes
    rng = np.random.default_rng(77)
    A1, A2 = rand_rows(rng, 3, 70000), rand_rows(rng, 2, 65000)
    B = rand_rows(rng, 3, 60000)
    want1, want2 = oracle_mul(A1, B[:3]), oracle_mul(A2, B[:2])
    builtin = g.ntt_get_launch()
    try:
        for mask, chunk in ((0, 0), (15, 8), (5, 5), (10, 2)):
            g.ntt_set_launch(mask, chunk)
            c1, c2 = device_mul_shared(ctx, A1, A2, B)
            assert np.array_equal(c1, want1) and np.array_equal(c2, want2), (mask, chunk)
    finally:
        g.ntt_set_launch(*builtin)


def test_launch_shape_rejects_bad_values():
    builtin = g.ntt_get_launch()
    for mask, chunk in ((16, 5), (5, 9), (5, -1)):
        with pytest.raises(ValueError):
            g.ntt_set_launch(mask, chunk)
    assert g.ntt_get_launch() == builtin


@pytest.mark.slow
def test_ntt_three_level(ctx):
    # K = 2^23 and 2^25: two column levels + the contiguous level
    g.check(ctx.set_ntt_threshold(2))
    try:
        rng = np.random.default_rng(5)
        for la, lb in [(1 << 22, (1 << 22) - 5), (1 << 24, 1 << 24)]:
            A = rand_rows(rng, 1, la)
            B = rand_rows(rng, 1, lb)
            assert np.array_equal(device_mul(ctx, A, B), oracle_mul(A, B))
    finally:
        g.check(ctx.set_ntt_threshold(384))


def ragged_case(ctx, lengths, seed, mode="random"):
    """Upload pairs (rand_rows mode), run one ragged call, return (device C rows, host A, B)."""
    rng = np.random.default_rng(seed)
    bufs, items, As, Bs, Cs = [], [], [], [], []
    for la, lb in lengths:
        A = rand_rows(rng, 1, la, mode) if la else np.zeros((1, 0), np.uint32)
        B = rand_rows(rng, 1, lb, mode) if lb else np.zeros((1, 0), np.uint32)
        ba = g.DeviceBuffer.from_array(A.reshape(-1)) if la else None
        bb = g.DeviceBuffer.from_array(B.reshape(-1)) if lb else None
        bc = g.DeviceBuffer.from_array(np.full(la + lb + 1, PAD, dtype=np.uint32))
        bufs += [ba, bb, bc]
        items.append((ba.ptr if ba else None, bb.ptr if bb else None, bc.ptr, la, lb))
        As.append(A)
        Bs.append(B)
        Cs.append(bc)
    err = ctx.mul_ragged(items)
    g.synchronize()
    out = [c.download(np.uint32, count=la + lb + 1) for c, (la, lb) in zip(Cs, lengths)]
    return err, out, As, Bs, bufs


def test_ragged_products(ctx):
    # one call with schoolbook, one-tile, column+row and three-level sizes,
    # zero-length operands and repeated transform sizes
    lengths = [(0, 0), (5, 0), (0, 7), (1, 1), (3, 17), (100, 90), (1000, 900), (2048, 2048), (2049, 1),
               (3000, 3000), (5000, 1200), (4095, 1), (65536, 30000), (40000, 50000), (1 << 20, 1 << 20),
               (2097153, 5), (12, 70000), (700, 700), (1000, 900)]
    err, out, As, Bs, _ = ragged_case(ctx, lengths, 21)
    assert err == 0
    for (la, lb), C, A, B in zip(lengths, out, As, Bs):
        assert C[-1] == PAD, (la, lb)
        if la and lb:
            assert np.array_equal(C[:-1], oracle_mul(A, B)[0]), (la, lb)
        else:
            assert not np.any(C[:-1]), (la, lb)


def test_ragged_pieces_and_shared_operands(ctx):
    # one ragged call where rows share B operands: items naming one buffer
    # (whole products of one transform size, a prefix of it, the buffer on
    # the A side) and split products (two pieces, many pieces, the longer
    # operand on either side, two products with one shorter operand, all-ones
    # carries through every seam); every C ends at its length
    rng = np.random.default_rng(41)
    ops = {
        "b": rand_rows(rng, 1, 30000)[0], "a1": rand_rows(rng, 1, 40000)[0], "a2": rand_rows(rng, 1, 50000)[0],
        "a3": rand_rows(rng, 1, 35000)[0], "s": rand_rows(rng, 1, 3000)[0], "l1": rand_rows(rng, 1, 4194309)[0],
        "l2": rand_rows(rng, 1, 1500000)[0], "t1": rand_rows(rng, 1, 3000000)[0],
        "t2": rand_rows(rng, 1, 500000)[0], "o1": rand_rows(rng, 1, 2500001, "ones")[0],
        "o2": rand_rows(rng, 1, 70001, "ones")[0],
    }
    bufs = {k: g.DeviceBuffer.from_array(v) for k, v in ops.items()}
    # (A, B, A length, B length): lengths up to the buffer's
    pairs = [("a1", "b", 40000, 30000), ("a2", "b", 50000, 30000), ("a3", "b", 35000, 30000),
             ("a1", "b", 40000, 20000), ("b", "a3", 30000, 35000), ("l1", "s", 4194309, 3000),
             ("s", "l2", 3000, 1500000), ("t1", "t2", 3000000, 500000), ("t2", "t1", 500000, 3000000),
             ("o1", "o2", 2500001, 70001), ("a2", "b", 50000, 30000)]
    outs, items = [], []
    for a, b, la, lb in pairs:
        c = g.DeviceBuffer.from_array(np.full(la + lb + 1, PAD, dtype=np.uint32))
        outs.append(c)
        items.append((bufs[a].ptr, bufs[b].ptr, c.ptr, la, lb))
    g.kernel_stats_reset(True)
    try:
        g.check(ctx.mul_ragged(items), "mul_ragged")
        g.synchronize()
        sizes = {int(name.split("@2^")[1]) for name, *_ in g.kernel_stats(by_size=True)}
    finally:
        g.kernel_stats_reset(False)
    assert max(sizes) <= 21, sizes  # every long product ran in pieces
    for (a, b, la, lb), c in zip(pairs, outs):
        got = c.download(np.uint32, count=la + lb + 1)
        assert got[-1] == PAD, (a, b, la, lb)
        want = oracle_mul(ops[a][:la].reshape(1, -1), ops[b][:lb].reshape(1, -1))[0]
        assert np.array_equal(got[:-1], want), (a, b, la, lb)


def test_ragged_one_tile_products_of_all_ones(ctx):
    # the largest coefficients a one-tile product meets (2^11 products of
    # two all-ones limbs: above 2^74), through the per-item operand path
    lengths = [(2048, 2048), (2047, 2049), (1024, 1000), (300, 200), (2000, 96), (129, 127)]
    err, out, As, Bs, _ = ragged_case(ctx, lengths, 9, "ones")
    assert err == 0
    for (la, lb), C, A, B in zip(lengths, out, As, Bs):
        assert C[-1] == PAD, (la, lb)
        assert np.array_equal(C[:-1], oracle_mul(A, B)[0]), (la, lb)


@pytest.mark.parametrize("k", [4, 5, 6, 7, 8, 9, 10, 11, 12])
def test_one_tile_products_of_all_ones_every_size(ctx, k):
    # the one-tile kernel at every transform size 2^4..2^12 (with the
    # crossover lowered, so that the small sizes reach it) on all-ones
    # operands, whose coefficients are the largest a product of that size
    # meets; uniform and ragged calls
    g.check(ctx.set_ntt_threshold(2))
    try:
        K = 1 << k
        lengths = [(K // 2, K // 2), (K - 1, 1), (K // 2 + 1, K // 2 - 1), (3 * K // 8, 3 * K // 8 + 1)]
        err, out, As, Bs, _ = ragged_case(ctx, lengths, 9, "ones")
        assert err == 0
        for (la, lb), C, A, B in zip(lengths, out, As, Bs):
            assert C[-1] == PAD, (la, lb)
            assert np.array_equal(C[:-1], oracle_mul(A, B)[0]), (la, lb)
        A = np.full((5, K // 2), 0xFFFFFFFF, dtype=np.uint32)
        assert np.array_equal(device_mul(ctx, A, A.copy()), oracle_mul(A, A))
    finally:
        g.check(ctx.set_ntt_threshold(384))


def test_ragged_matches_uniform_call(ctx):
    # the same pairs through batch_mp_mul one by one
    lengths = [(3000, 2000), (777, 4000), (20000, 1), (9000, 9000)]
    err, out, As, Bs, _ = ragged_case(ctx, lengths, 5)
    assert err == 0
    for (la, lb), C, A, B in zip(lengths, out, As, Bs):
        assert np.array_equal(C[:-1], device_mul(ctx, A, B)[0])


def test_ragged_rejects_aliasing(ctx):
    buf = g.DeviceBuffer.from_array(np.ones(64, dtype=np.uint32))
    assert ctx.mul_ragged([(buf.ptr, buf.ptr + 64, buf.ptr, 8, 8)]) != 0
    assert ctx.mul_ragged([(None, buf.ptr, buf.ptr + 128, 8, 8)]) != 0


def test_tuner_times_every_launch_shape_and_restores_it(ctx):
    from gpupi_amd import tune
    ops = tune.Operands(20, seed=8)
    before = (g.ntt_get_launch(), g.group_stage_get_shape())
    res = tune.tune_launch(ctx, ops, ks=(14, 17, 20), reps=1)
    assert sorted(v for v, _ in res["pair_mask"]) == list(range(16))
    assert sorted(v for v, _ in res["xcd_chunk_log2"]) == list(range(9))
    assert sorted(v for v, _ in res["group_blocks"]) == [1, 2, 4, 8, 16]
    assert all(ms > 0 for key in ("pair_mask", "xcd_chunk_log2", "group_blocks") for _, ms in res[key])
    assert (g.ntt_get_launch(), g.group_stage_get_shape()) == before


def test_tuner_times_every_split_and_restores_the_plan(ctx):
    from gpupi_amd import tune
    ops = tune.Operands(20, seed=5)
    before = g.ntt_get_split(15)
    res = tune.tune_splits(ctx, ops, 15, reps=1)
    assert sorted(lv for lv, _ in res) == sorted((m0, 15 - m0) for m0, _ in tune.splits(15))
    assert all(ms > 0 for _, ms in res) and res[0][1] <= res[-1][1]
    assert g.ntt_get_split(15) == before
    # the tree's pattern: two products sharing an operand, every split exact
    ops = tune.Operands(20, seed=6, shared=True)
    before = g.ntt_get_split(16)
    res = tune.tune_splits(ctx, ops, 16, reps=1, shared=True)
    assert sorted(lv for lv, _ in res) == sorted((m0, 16 - m0) for m0, _ in tune.splits(16))
    assert g.ntt_get_split(16) == before
