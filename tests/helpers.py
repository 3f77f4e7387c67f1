"""Shared test helpers: random limb rows, Python big-integer closed forms,
and runners that apply one batched op through the oracle or the device."""
from __future__ import annotations

import numpy as np

MASK32 = 0xFFFFFFFF


def to_int(row) -> int:
    return int.from_bytes(np.asarray(row, dtype="<u4").tobytes(), "little")


def from_int(v: int, length: int) -> np.ndarray:
    v &= (1 << (32 * length)) - 1
    return np.frombuffer(v.to_bytes(4 * length, "little"), dtype="<u4").copy() if length else np.zeros(0, np.uint32)


def rand_rows(rng: np.random.Generator, n: int, length: int, mode: str = "random") -> np.ndarray:
    """Rows of limbs.  mode: random | ones (0xffffffff) | zeros | sparse | top."""
    if mode == "random":
        return rng.integers(0, 1 << 32, size=(n, length), dtype=np.uint64).astype(np.uint32)
    if mode == "ones":
        return np.full((n, length), MASK32, dtype=np.uint32)
    if mode == "zeros":
        return np.zeros((n, length), dtype=np.uint32)
    if mode == "sparse":
        r = rng.integers(0, 1 << 32, size=(n, length), dtype=np.uint64).astype(np.uint32)
        r[rng.random((n, length)) < 0.7] = 0
        return r
    if mode == "top":  # long runs of ones with a random low limb (carry chains)
        r = np.full((n, length), MASK32, dtype=np.uint32)
        if length:
            r[:, 0] = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
        return r
    raise ValueError(mode)


def closed_form(kind: str, a: int, b: int, scalar: int, shift: int, lc: int) -> int:
    """Reference semantics of each batched op (batch_*.h), mod 2^(32 lc)."""
    m = (1 << (32 * lc)) - 1
    if kind == "add":
        return (a + b) & m
    if kind == "sub":
        return (a - b) & m
    if kind == "add_small":
        return (a + scalar) & m
    if kind == "sub_small":
        return (a - scalar) & m
    if kind == "mul_small":
        return (a * scalar) & m
    if kind == "shift_add":
        return ((a << shift) + b) & m
    if kind == "shift_sub":
        return ((a << shift) - b) & m
    if kind == "mul":
        return a * b
    raise ValueError(kind)


def run_jobs(srv, pis, muls):
    """Submits the jobs, waits for every result and returns them keyed by
    (kind, size); the keys must be distinct (jobs of one shape would share a
    key, so they are refused here rather than counted once)."""
    ids = {}
    for d in pis:
        ids[srv.submit_pi(d, keep_digits=True)] = ("pi", d)
    for A, B in muls:
        ids[srv.submit_mul(A, B)] = ("mul", (A.size, B.size))
    assert all(i >= 0 for i in ids)
    assert len(set(ids.values())) == len(ids), "jobs of one shape in one call"
    out, received = {}, 0
    while received < len(ids):
        got = srv.results(wait=True)
        assert got, "the server returned no result while jobs were outstanding"
        for r, data in got:
            assert r.status == 0, (ids[r.id], r.status)
            out[ids[r.id]] = (r, data)
            received += 1
    return out
