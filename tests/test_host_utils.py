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


def test_bench_refuses_more_than_one_gpu():
    # one process on one GPU: --gpus above 1 is refused before anything
    # touches a device
    import pathlib
    import subprocess
    import sys
    root = pathlib.Path(__file__).resolve().parent.parent
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
