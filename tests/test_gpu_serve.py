"""GPU: the serving loop (gpupi_serve.h).  A mix of pi-digit and multiply
jobs submitted together must give every job the result it gets alone, and
those results must equal the oracle (GMP mpn_mul; the pi restatement)."""
import numpy as np
import pytest

import gpupi_amd as g
import oracle_ffi as of
from helpers import rand_rows, run_jobs

pytestmark = pytest.mark.gpu

PI_SIZES = [1000, 20000, 100000]
MUL_SIZES = [(4096, 4096), (100, 90), (70000, 3000), (1 << 17, 1 << 16), (5000, 5000), (1 << 20, 1 << 20)]


def oracle_mul(A, B):
    hc = of.HostArray.empty(1, A.size + B.size)
    assert of.lib().o_mul(of.HostArray(A.reshape(1, -1)).oarray(), of.HostArray(B.reshape(1, -1)).oarray(),
                          hc.oarray()) == 0
    return hc.rows()[0]


def test_mix_equals_alone_and_oracle():
    rng = np.random.default_rng(77)
    muls = [(rand_rows(rng, 1, a)[0], rand_rows(rng, 1, b)[0]) for a, b in MUL_SIZES]
    srv = g.Server(pi_workers=3)
    try:
        mix = run_jobs(srv, PI_SIZES, muls)
        alone = {}
        for d in PI_SIZES:
            alone.update(run_jobs(srv, [d], []))
        for A, B in muls:
            alone.update(run_jobs(srv, [], [(A, B)]))
    finally:
        srv.close()
    for key, (r, data) in mix.items():
        ra, da = alone[key]
        if key[0] == "pi":
            assert (r.ret_hash, r.dec_hash, r.ret_length) == (ra.ret_hash, ra.dec_hash, ra.ret_length)
            assert data == da
            _, odec, _ = of.pi_digits(key[1])
            assert data.decode() == odec
        else:
            assert np.array_equal(data, da)
    for (A, B) in muls:
        assert np.array_equal(mix[("mul", (A.size, B.size))][1], oracle_mul(A, B))
    # the multiply jobs submitted together shared launches
    assert max(r.batch_size for k, (r, _) in mix.items() if k[0] == "mul") > 1


def test_pi_jobs_of_one_tree_size_share_a_tree():
    # one pi worker: the jobs queued behind the first, all of one tree size,
    # join one batch (one shared tree); every result equals the job served
    # with batching off and the oracle
    sizes = [40000, 45000, 50000, 55000, 60000]
    assert len({g.lib().gpupi_pi_groups_for_digits(d) for d in sizes}) == 1
    srv = g.Server(pi_workers=1)
    try:
        batched = run_jobs(srv, sizes, [])
    finally:
        srv.close()
    srv = g.Server(pi_workers=1, pi_batch_max=1)
    try:
        single = run_jobs(srv, sizes, [])
    finally:
        srv.close()
    assert max(r.batch_size for r, _ in batched.values()) > 1
    assert all(r.batch_size == 1 for r, _ in single.values())
    for d in sizes:
        r, data = batched[("pi", d)]
        rs, ds = single[("pi", d)]
        assert (r.ret_hash, r.dec_hash, r.ret_length) == (rs.ret_hash, rs.dec_hash, rs.ret_length), d
        assert data == ds
        _, odec, _ = of.pi_digits(d)
        assert data.decode() == odec


def test_pool_stays_within_budget_under_load():
    # many jobs against a budget of about three jobs: the pool's peak (all
    # workers' pipeline buffers, workspaces and arenas) stays within it
    rng = np.random.default_rng(11)
    pis = [3000, 7000, 11000, 20000, 30000]
    muls = [(rand_rows(rng, 1, 1 << 17)[0], rand_rows(rng, 1, (1 << 17) - i)[0]) for i in range(5)]
    budget = 3 * max(g.lib().gpupi_serve_job_bytes(g.JOB_PI, 30000, 0),
                     g.lib().gpupi_serve_job_bytes(g.JOB_MUL, 1 << 17, 1 << 17))
    srv = g.Server(budget_bytes=budget, pi_workers=3)
    try:
        out = run_jobs(srv, pis, muls)
        used_high, _ = srv.memory()
    finally:
        srv.close()
    assert 0 < used_high <= budget, (used_high, budget)
    for d in pis:
        r, data = out[("pi", d)]
        _, odec, _ = of.pi_digits(d)
        assert data.decode() == odec
    for A, B in muls:
        assert np.array_equal(out[("mul", (A.size, B.size))][1], oracle_mul(A, B))


def test_small_budget_serializes_but_stays_exact():
    rng = np.random.default_rng(3)
    # three shapes (results are keyed by shape), each over half the budget
    muls = [(rand_rows(rng, 1, (1 << 19) - d)[0], rand_rows(rng, 1, 1 << 19)[0]) for d in range(3)]
    need = g.lib().gpupi_serve_job_bytes(g.JOB_PI, 5000, 0)
    for A, B in muls:
        est_mul = g.lib().gpupi_serve_job_bytes(g.JOB_MUL, A.size, B.size)
        assert need / 2 < est_mul <= need  # any two jobs exceed the budget
    srv = g.Server(budget_bytes=need, pi_workers=2)
    try:
        out = run_jobs(srv, [5000], muls)
        # a job larger than the whole budget is refused at submission
        assert srv.submit_mul(np.ones(1 << 22, np.uint32), np.ones(1 << 22, np.uint32)) == -1
        # every worker's device memory came from the server's one pool, whose
        # peak stayed within the budget (one job at a time)
        used_high, reserved_high = srv.memory()
        assert 0 < used_high <= need, (used_high, need)
    finally:
        srv.close()
    for A, B in muls:
        r, data = out[("mul", (A.size, B.size))]
        assert r.reserved_bytes <= need and r.batch_size == 1
        assert np.array_equal(data, oracle_mul(A, B))
    spans = sorted((r.start_ms, r.finish_ms) for r, _ in out.values())
    assert all(b[0] >= a[1] - 1e-6 for a, b in zip(spans, spans[1:]))  # one at a time
