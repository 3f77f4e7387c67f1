"""GPU: the pi pipeline and each of its stages through the C ABI, bit-exact
against the oracle's restatement of the reference; convert_decimal against
the reference's own known-answer cases."""
import ctypes

import numpy as np
import pytest

import gpupi_amd as g
import oracle_ffi as of
import refcases
from helpers import rand_rows

pytestmark = pytest.mark.gpu


def take_device(arr: g.BatchMPArray) -> np.ndarray:
    out = g.download_words(arr.data, arr.length)
    g.hip().hipFree(ctypes.c_void_p(arr.data))
    return out


@pytest.mark.parametrize("is_head,count", [(1, 1), (1, 4096), (0, 16), (0, 65536 * 16 + 5)])
def test_leaf_terms(ctx, is_head, count):
    bufs = [g.DeviceBuffer(count * w * 4) for w in (3, 4, 6)]
    g.check(g.lib().gpupi_construct_leaf_17(ctx.ptr, is_head, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, count))
    g.synchronize()
    host = [np.zeros(count * w, dtype=np.uint32) for w in (3, 4, 6)]
    of.lib().o_leaf17(is_head, host[0].ctypes.data, host[1].ctypes.data, host[2].ctypes.data, count)
    for b, h in zip(bufs, host):
        assert np.array_equal(b.download(np.uint32, count=h.size), h)


@pytest.mark.parametrize("groups", [1, 2, 8, 64, 1024])
def test_binary_split(ctx, groups):
    Q, R = g.BatchMPArray(), g.BatchMPArray()
    g.check(g.lib().gpupi_binary_split(ctx.ptr, groups, ctypes.byref(Q), ctypes.byref(R)))
    oq, orr = of.OArray(), of.OArray()
    assert of.lib().o_binary_split(groups, ctypes.byref(oq), ctypes.byref(orr)) == 0
    assert np.array_equal(take_device(Q), of.take_array(oq))
    assert np.array_equal(take_device(R), of.take_array(orr))


@pytest.mark.parametrize("digits", [5, 100, 10000, 300000])
def test_sqrt_10005(ctx, digits):
    P, Q = g.BatchMPArray(), g.BatchMPArray()
    g.check(g.lib().gpupi_sqrt_10005(ctx.ptr, digits, ctypes.byref(P), ctypes.byref(Q)))
    op, oq = of.OArray(), of.OArray()
    assert of.lib().o_sqrt_10005(digits, ctypes.byref(op), ctypes.byref(oq)) == 0
    assert np.array_equal(take_device(P), of.take_array(op))
    assert np.array_equal(take_device(Q), of.take_array(oq))


@pytest.mark.parametrize("limbs", [10, 1000, 1100, 5000, 40000])
def test_recursive_inverse(ctx, limbs):
    rng = np.random.default_rng(limbs)
    x = rand_rows(rng, 1, limbs)
    x[0, -1] |= 0x80000000
    L = 2 * (limbs * 32) - 1
    bx, X = g.upload_rows(x)
    Y = g.BatchMPArray()
    g.check(g.lib().gpupi_recursive_inverse(ctx.ptr, X, L, ctypes.byref(Y)))
    hx = of.HostArray(x)
    oy = of.OArray()
    assert of.lib().o_recursive_inverse(hx.oarray(), L, ctypes.byref(oy)) == 0
    assert np.array_equal(take_device(Y), of.take_array(oy))


@pytest.mark.parametrize("total,bits,words,expected", refcases.convert_cases(),
                         ids=[str(c[0]) for c in refcases.convert_cases()])
def test_convert_reference_cases(ctx, total, bits, words, expected):
    leaf = 477
    levels = (total // leaf).bit_length() - 1
    table = g.ConvertDecimalPowerTable()
    g.check(g.lib().convert_decimal_precompute_powers(ctx.ptr, ctypes.byref(table), leaf, levels))
    try:
        b, X = g.upload_rows(np.array([words], dtype=np.uint32))
        out = g.DeviceBuffer(total)
        g.check(g.lib().convert_decimal_equal_split(ctx.ptr, X, bits, total, leaf, ctypes.byref(table), out.ptr))
        g.synchronize()
        assert out.download(np.uint8, count=total).tobytes().decode() == expected
    finally:
        g.lib().convert_decimal_release_powers(ctypes.byref(table))


@pytest.mark.parametrize("digits", [0, 1, 9, 100, 575, 576, 1000, 5000, 20000, 100000, 1000000])
def test_pi_matches_oracle(ctx, digits):
    ret, dec, res = ctx.pi(digits)
    oret, odec, ocd = of.pi_digits(digits)
    assert res.convert_digits == ocd
    assert np.array_equal(ret, oret)
    assert dec == odec


BATCHES = [
    [1000, 5000, 5100, 30000, 100000, 100000, 250000],  # nested trees, equal ones, one size twice
    [0, 1, 9, 100, 575, 576],                           # one-group trees and the empty job
    [200000, 210000, 220000],                           # one tree size
]


@pytest.mark.parametrize("digits", BATCHES, ids=["nested", "tiny", "same_tree"])
@pytest.mark.parametrize("convention", [g.PREC_REQUESTED, g.PREC_ROUNDED], ids=["requested", "rounded"])
def test_pi_batch_equals_single_runs(ctx, digits, convention):
    # jobs sharing one binary-splitting tree: each result equals the job run
    # alone (RET limbs, digits, sizes), in input order
    batch = ctx.pi_batch(digits, convention)
    groups = [g.lib().gpupi_pi_groups_for_digits(r.convert_digits if convention == g.PREC_ROUNDED else d)
              for d, (_, _, r) in zip(digits, batch)]
    for d, (ret, dec, res), n in zip(digits, batch, groups):
        oret, odec, ores = ctx.pi(d, convention)
        assert (res.convert_digits, res.ret_length, res.groups, res.inverse_bits) == \
            (ores.convert_digits, ores.ret_length, ores.groups, ores.inverse_bits), d
        assert res.groups == n
        assert np.array_equal(ret, oret), d
        assert dec == odec, d


def test_pi_batch_matches_oracle_and_plain_form(ctx):
    digits = [3000, 40000, 41000]
    batch = ctx.pi_batch(digits)
    for d, (ret, dec, _) in zip(digits, batch):
        oret, odec, _ = of.pi_digits(d)
        assert np.array_equal(ret, oret) and dec == odec, d
    with g.Context() as plain:
        plain.set_form(g.FORM_PLAIN)
        for (ret, dec, _), (pret, pdec, _) in zip(batch, plain.pi_batch(digits)):
            assert np.array_equal(ret, pret) and dec == pdec


def test_pi_batch_rejects_bad_arguments(ctx):
    res = (g.GpupiPiResult * 1)()
    arr = (ctypes.c_uint32 * 1)(100)
    assert g.lib().gpupi_pi_run_batch(ctx.ptr, arr, 0, 0, res) != 0
    assert g.lib().gpupi_pi_run_batch(ctx.ptr, None, 1, 0, res) != 0
    assert g.lib().gpupi_pi_run_batch(ctx.ptr, arr, 1, 0, None) != 0


def test_repeat_run_reuses_cached_blocks():
    # the second run of one size takes its buffers from the context's block
    # cache (RET's block leaves the cache with the result; best fit may pair
    # requests and blocks differently from the first run)
    with g.Context() as c:
        ret1, dec1, _ = c.pi(200000)
        cached1, allocs1 = c.cache_stats()
        assert cached1 > 0
        ret2, dec2, _ = c.pi(200000)
        cached2, allocs2 = c.cache_stats()
        print(f"allocations: first run {allocs1}, second run {allocs2 - allocs1}; cached {cached2} bytes")
        assert allocs2 - allocs1 <= 4
        assert np.array_equal(ret1, ret2) and dec1 == dec2
        c.cache_trim()
        assert c.cache_stats()[0] == 0
        ret3, dec3, _ = c.pi(200000)
        assert np.array_equal(ret1, ret3) and dec1 == dec3
    oret, odec, _ = of.pi_digits(200000)
    assert np.array_equal(ret1, oret) and dec1 == odec


def test_tree_graph_replays_and_recaptures():
    # the bottom tree levels run from a hipGraph captured on first use and
    # replayed after; a new group count or crossover captures anew
    sizes = [(300000, None), (300000, None), (120000, None), (300000, 256), (300000, None), (5000, None)]
    with g.Context() as c:
        for digits, threshold in sizes:
            if threshold is not None:
                g.check(c.set_ntt_threshold(threshold))
            ret, dec, _ = c.pi(digits)
            if threshold is not None:
                g.check(c.set_ntt_threshold(384))
            oret, odec, _ = of.pi_digits(digits)
            assert np.array_equal(ret, oret), (digits, threshold)
            assert dec == odec, (digits, threshold)


@pytest.mark.parametrize("digits", [1, 64, 763, 768, 20000])
def test_pi_digits_match_golden_fixture(ctx, digits):
    # tests/golden/pi_digits.txt: pi from an independent implementation
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pi_digits.txt")) as f:
        want = f.read().strip()
    for conv in (g.PREC_REQUESTED, g.PREC_ROUNDED):
        _, dec, _ = ctx.pi(digits, conv)
        assert "3" + dec == want[:digits + 1], conv


@pytest.mark.parametrize("digits", [1000, 123457])
def test_pi_readme_convention(ctx, digits):
    ret, dec, res = ctx.pi(digits, g.PREC_ROUNDED)
    oret, odec, _ = of.pi_digits(digits, readme_prec=True)
    assert np.array_equal(ret, oret)
    assert dec == odec


# the 17-term group stage runs as one fused kernel (one "leaf" entry; its
# exact divisions happen inside it)
STAGE_KINDS = {"leaf", "mul", "mul_small", "shift", "add", "sub", "compact", "decimal_powers", "decimal_convert"}


def test_stage_profile(ctx):
    # the reference's --profile-stages accounting (pi_bs_gpu.cu:133-489,
    # :2258-2325): profiling leaves the result unchanged and every stage kind
    # of the pipeline shows up
    ret0, dec0, _ = ctx.pi(100000)
    ctx.stage_profile(True)
    try:
        ret1, dec1, res1 = ctx.pi(100000)
        entries = ctx.stage_entries()
    finally:
        ctx.stage_profile(False)
    assert np.array_equal(ret0, ret1) and dec0 == dec1
    assert {e["kind"] for e in entries} == STAGE_KINDS
    assert [e["label"] for e in entries if e["kind"] == "leaf"] == ["fused_group_stage"]
    assert any(e["ntt"] for e in entries if e["kind"] == "mul")
    assert not any(e["ntt"] for e in entries if e["kind"] != "mul")
    assert 0 < sum(e["ms"] for e in entries) <= res1.elapsed_ms * 1.05 + 1.0
    ctx.pi(1000)  # disabled: nothing more is recorded
    assert len(ctx.stage_entries()) == len(entries)


def decimal_hash(digits: str) -> int:
    v = 3  # pi_bs_gpu.cu:860-866
    for ch in digits:
        v = (v * 10 + ord(ch) - 48) % 1000000007
    return v


def test_cli_output_file_and_bbp_check(tmp_path):
    # --output writes "3." and the digits; --verify-bbp compares RET's hex
    # digits with the BBP digit-extraction formula evaluated on the host
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(g.LIB_PATH), "pi_bs_gpu")
    out_file = tmp_path / "pi.txt"
    r = subprocess.run([exe, "100000", "--output", str(out_file), "--verify-bbp", "1", "--verify-bbp", "1000",
                        "--verify-bbp", "40000", "--verify-bbp", "82000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stderr.count(" match") == 4 and "MISMATCH" not in r.stderr
    assert "formula 243f6a88 RET 243f6a88" in r.stderr
    text = out_file.read_text()
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pi_digits.txt")) as f:
        want = f.read().strip()
    assert len(text) == 100003 and text.startswith("3.") and text.endswith("\n")
    assert "3" + text[2:20002] == want[:20001]
    # a position past RET's exact digits is refused
    r = subprocess.run([exe, "1000", "--verify-bbp", "100000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "beyond" in r.stderr


def test_cli_profile_stages():
    import os
    import re
    import subprocess
    exe = os.path.join(os.path.dirname(g.LIB_PATH), "pi_bs_gpu")
    out = subprocess.run([exe, "100000", "--profile-stages"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    stages = [ln for ln in lines if ln.startswith("stage[")]
    assert len(stages) > 50
    assert all(re.match(r"stage\[\d{3,}\] kind=\s*\S+ label=\s*\S+ digits=\s*\d+ batch=\s*\d+ lhs_len=\s*\d+ "
                        r"rhs_len=\s*\d+ kernel_ms=\s*[\d.]+ uses_ntt=[01]$", ln) for ln in stages)
    totals = next(ln for ln in lines if ln.startswith("leaf_total_ms="))
    assert "ntt_total_ms=" in totals and "decimal_convert_total_ms=" in totals
    stats = next(ln for ln in lines if ln.startswith("target_digits=100000 "))
    assert "accounted_ms=" in stats and f"RET_len={of.lib().o_prec_limbs_for_digits(100000) + 1} " in stats
    _, odec, _ = of.pi_digits(100000)
    assert f"DEC_preview_hash = {decimal_hash(odec[:4096])}" in lines


@pytest.mark.slow
def test_pi_1e7_matches_oracle(ctx):
    ret, dec, res = ctx.pi(10_000_000)
    oret, odec, _ = of.pi_digits(10_000_000)
    assert len(ret) == 1038127
    assert np.array_equal(ret, oret)
    assert dec == odec


@pytest.mark.parametrize("digits", [0, 1000, 20000, 1000000])
def test_plain_form_matches_oracle(digits):
    # GPUPI_FORM_PLAIN: the reference's per-op form (group stage op by op,
    # a compaction after every product it compacts, separate products for
    # shared operands, hipMalloc/hipFree instead of the block cache) computes
    # the same RET limbs and digits
    with g.Context() as c:
        c.set_form(g.FORM_PLAIN)
        c.stage_profile(digits == 20000)
        ret, dec, res = c.pi(digits)
        entries = c.stage_entries()
        assert c.cache_stats() == (0, 0), "the plain form took blocks from the cache"
    oret, odec, ocd = of.pi_digits(digits)
    assert res.convert_digits == ocd
    assert np.array_equal(ret, oret)
    assert dec == odec
    if digits == 20000:
        kinds = [e["kind"] for e in entries]
        assert kinds.count("leaf") == 2 and "exactdiv" in kinds
        # one compaction per P, Q and R of every tree level (group stage
        # levels included), one per sqrt result
        assert kinds.count("compact") > 3 * 20
