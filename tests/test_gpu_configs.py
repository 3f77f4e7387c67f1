"""GPU: the measured workloads at their full sizes, bit-exact.

  ragged 256 pairs  seed 1234, lengths floor(2^U(10, 24)) limbs, one
                    gpupi_mul_ragged call; every product against GMP mpn_mul
                    (batch_mul_ntt_test.cu:509-577 checks its products the
                    same way)
  pi 10^9           README precision convention: RET length, hash, preview
                    and the decimal head and tail of README.md:72-74
  pi 10^8           requested-digit convention: RET limbs and every digit
                    against the oracle (run on the host's cores)
  serving stream    the Poisson stream of the serving benchmark (seed 7):
                    every job alone equals the job in the mix, every multiply
                    equals GMP, every pi job up to 10^7 digits equals the
                    oracle and every larger one starts with the golden digits
The group stage at 65,536 groups (the 10^7-digit run) is in
test_gpu_leaf_stage.py."""
import json
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import gpupi_amd as g
import oracle_ffi as of

pytestmark = [pytest.mark.gpu, pytest.mark.slow]
HERE = os.path.dirname(os.path.abspath(__file__))
HOST_THREADS = max(1, min(16, os.cpu_count() or 1))


def log_uniform_pairs(seed, count, lo, hi):
    rng = np.random.default_rng(seed)
    u = rng.uniform(lo, hi, size=(count, 2))
    return rng, [(int(2 ** a), int(2 ** b)) for a, b in u]


def test_ragged_256_pairs_full_size():
    rng, lengths = log_uniform_pairs(1234, 256, 10.0, 24.0)
    As = [rng.integers(0, 1 << 32, size=la, dtype=np.uint64).astype(np.uint32) for la, _ in lengths]
    Bs = [rng.integers(0, 1 << 32, size=lb, dtype=np.uint64).astype(np.uint32) for _, lb in lengths]
    with g.Context() as ctx:
        bufs, items = [], []
        for (la, lb), A, B in zip(lengths, As, Bs):
            ba, bb = g.DeviceBuffer.from_array(A), g.DeviceBuffer.from_array(B)
            bc = g.DeviceBuffer.from_array(np.full(la + lb + 1, 0xDEADBEEF, dtype=np.uint32))
            bufs.append((ba, bb, bc))
            items.append((ba.ptr, bb.ptr, bc.ptr, la, lb))
        g.check(ctx.mul_ragged(items), "gpupi_mul_ragged")
        g.synchronize()
        got = [bc.download(np.uint32) for _, _, bc in bufs]
    assert all(c[-1] == 0xDEADBEEF for c in got), "write past a product"
    order = sorted(range(len(lengths)), key=lambda i: -(lengths[i][0] + lengths[i][1]))
    with ThreadPoolExecutor(HOST_THREADS) as pool:
        want = dict(zip(order, pool.map(lambda i: of.mul(As[i], Bs[i]), order)))
    bad = [lengths[i] for i in range(len(lengths)) if not np.array_equal(got[i][:-1], want[i])]
    assert not bad, bad


def test_pi_1e9_readme_convention():
    golden = json.load(open(os.path.join(HERE, "golden", "readme_1e9.json")))
    with g.Context() as ctx:
        ret, dec, res = ctx.pi(golden["target_digits"], g.PREC_ROUNDED)
    assert res.convert_digits == golden["convert_digits"]
    assert len(ret) == golden["ret_len"]
    assert g.ret_mod_hash(ret) == golden["ret_hash"]
    top = format(int(ret[-1]), "x") + format(int(ret[-2]), "08x")
    low = format(int(ret[1]), "08x") + format(int(ret[0]), "08x")
    assert f"{top}...{low}" == golden["ret_preview"]
    assert "3" + dec[:63] == golden["dec_head"]
    assert dec[-64:] == golden["dec_tail"]
    # the first 20000 digits against an independent pi
    assert dec[:20000] == open(os.path.join(HERE, "golden", "pi_digits.txt")).read().strip()[1:20001]


def test_pi_1e8_matches_oracle():
    with g.Context() as ctx:
        ret, dec, res = ctx.pi(100_000_000)
    of.set_threads(HOST_THREADS)
    try:
        oret, odec, ocd = of.pi_digits(100_000_000)
    finally:
        of.set_threads(1)
    assert res.convert_digits == ocd
    assert len(ret) == len(oret) == 10381252
    assert np.array_equal(ret, oret)
    assert dec == odec


def test_serving_stream_alone_mix_and_oracle():
    import threading
    import time
    rng = np.random.default_rng(7)
    n = 60
    jobs = []
    for _ in range(n):
        if rng.random() < 0.4:
            jobs.append(("pi", int(10 ** rng.uniform(5, 8))))
        else:
            la, lb = int(2 ** rng.uniform(12, 26)), int(2 ** rng.uniform(12, 26))
            jobs.append(("mul", (rng.integers(0, 1 << 32, size=la, dtype=np.uint64).astype(np.uint32),
                                 rng.integers(0, 1 << 32, size=lb, dtype=np.uint64).astype(np.uint32))))
    gaps = rng.exponential(1.0, size=n)
    srv = g.Server(budget_bytes=200 * 10 ** 9)
    try:
        alone, service = {}, []
        for i, (kind, x) in enumerate(jobs):
            jid = srv.submit_pi(x, keep_digits=True) if kind == "pi" else srv.submit_mul(*x)
            assert jid >= 0
            (r, data), = srv.results(wait=True)
            assert r.status == 0
            alone[i] = (r.ret_length, r.ret_hash, r.dec_hash, bytes(data)) if kind == "pi" else data
            service.append(r.finish_ms - r.start_ms)
        rate = 0.7 / (sum(service) / n / 1e3)
        arrivals = np.cumsum(gaps / rate)
        results, id2job = {}, {}

        def collect():
            while len(results) < n:
                got = srv.results(wait=True)
                for r, data in got:
                    results[r.id] = (r, data)
                if not got:
                    time.sleep(0.0005)

        th = threading.Thread(target=collect)
        th.start()
        t0 = time.perf_counter()
        for i, (kind, x) in enumerate(jobs):
            delay = arrivals[i] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            jid = srv.submit_pi(x, keep_digits=True) if kind == "pi" else srv.submit_mul(*x)
            assert jid >= 0
            id2job[jid] = i
        th.join(timeout=600)
        assert len(results) == n
    finally:
        srv.close()
    mix = {id2job[jid]: (r, data) for jid, (r, data) in results.items()}
    for i, (kind, x) in enumerate(jobs):
        r, data = mix[i]
        assert r.status == 0
        if kind == "pi":
            assert (r.ret_length, r.ret_hash, r.dec_hash, bytes(data)) == alone[i], ("pi", x)
        else:
            assert np.array_equal(data, alone[i]), ("mul", x[0].size, x[1].size)
    golden = open(os.path.join(HERE, "golden", "pi_digits.txt")).read().strip()[1:]
    small_pi = [i for i, (k, x) in enumerate(jobs) if k == "pi" and x <= 10_000_000]
    muls = sorted((i for i, (k, _) in enumerate(jobs) if k == "mul"), key=lambda i: -jobs[i][1][0].size)

    def check(i):
        kind, x = jobs[i]
        if kind == "mul":
            return np.array_equal(mix[i][1], of.mul(*x))
        _, odec, _ = of.pi_digits(x)
        return bytes(mix[i][1]).decode() == odec

    with ThreadPoolExecutor(HOST_THREADS) as pool:
        ok = dict(zip(muls + small_pi, pool.map(check, muls + small_pi)))
    assert all(ok.values()), [jobs[i][0] for i, v in ok.items() if not v]
    for i, (kind, x) in enumerate(jobs):
        if kind == "pi" and x > 10_000_000:
            assert bytes(mix[i][1][:20000]).decode() == golden[:20000]
