"""CPU: host-side helpers of the package (no GPU calls)."""
import numpy as np
import pytest

import gpupi_amd as g
import oracle_ffi as of


def horner_hash(limbs) -> int:
    # pi_bs_gpu.cu:852-858, one limb at a time from the top
    mod = 1000000007
    base = (1 << 32) % mod
    v = 0
    for x in reversed([int(t) for t in limbs]):
        v = (v * base + x) % mod
    return v


@pytest.mark.parametrize("n", [0, 1, 2, 3, 7, 64, 1001, 65537])
def test_ret_mod_hash_matches_horner_and_oracle(n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
    want = horner_hash(a)
    assert g.ret_mod_hash(a) == want
    assert of.mod_hash(a) == want
    ones = np.full(n, 0xFFFFFFFF, dtype=np.uint32)
    assert g.ret_mod_hash(ones) == horner_hash(ones)
