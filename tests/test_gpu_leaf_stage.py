"""GPU: the 17-term group stage (leaf terms, four pairwise levels, head
merge, exact division; pi_bs_gpu.cu:555-779) in its three forms: the
reference's per-op form with a compaction after every product, the batched
form with lengths fixed in advance, and the batched form captured as a
hipGraph and replayed.  Every form must leave the oracle's P, Q, R rows
(same values, same compacted lengths); graph replays are checked one by one
with eager work in between."""
import ctypes

import numpy as np
import pytest

import gpupi_amd as g
import oracle_ffi as of
from helpers import rand_rows

pytestmark = pytest.mark.gpu
MODES = {"plain": g.LEAF_PLAIN, "batched": g.LEAF_BATCHED, "graph": g.LEAF_GRAPH, "fused": g.LEAF_FUSED}


def oracle_stage(groups):
    op, oq, orr = of.OArray(), of.OArray(), of.OArray()
    assert of.lib().o_leaf_stage(groups, ctypes.byref(op), ctypes.byref(oq), ctypes.byref(orr)) == 0
    return of.take_rows(op), of.take_rows(oq), of.take_rows(orr)


def stage_rows(stage):
    return [g.rows_of(a) for a in stage.run()]


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("groups", [1, 4, 64, 4096, pytest.param(65536, marks=pytest.mark.slow)])
def test_leaf_stage_matches_oracle(ctx, groups, mode):
    want = oracle_stage(groups)
    st = g.LeafStage(ctx, groups, MODES[mode])
    try:
        for rep in range(3 if mode == "graph" else 1):
            got = stage_rows(st)
            for name, w, x in zip("PQR", want, got):
                assert w.shape == x.shape, (name, rep, w.shape, x.shape)
                assert np.array_equal(w, x), (name, rep)
    finally:
        st.close()


@pytest.mark.parametrize("groups", [4, 4096])
def test_fused_stage_every_block_shape(ctx, groups):
    # groups per block of the fused kernel: the oracle's rows for every
    # shape, including blocks with more waves than groups
    want = oracle_stage(groups)
    before = g.group_stage_get_shape()
    st = g.LeafStage(ctx, groups, g.LEAF_FUSED)
    try:
        for gpb in (1, 2, 4, 8, 16):
            g.group_stage_set_shape(gpb)
            assert g.group_stage_get_shape() == gpb
            for name, w, x in zip("PQR", want, stage_rows(st)):
                assert np.array_equal(w, x), (name, gpb)
        with pytest.raises(ValueError):
            g.group_stage_set_shape(3)
    finally:
        g.group_stage_set_shape(before)
        st.close()
    assert g.group_stage_get_shape() == before


def test_graph_replay_after_context_grows(ctx):
    # the captured graph uses the stage's own scratch: a large product that
    # regrows the context's workspace in between must not disturb replays
    want = oracle_stage(256)
    st = g.LeafStage(ctx, 256, g.LEAF_GRAPH)
    try:
        assert all(np.array_equal(w, x) for w, x in zip(want, stage_rows(st)))
        rng = np.random.default_rng(9)
        A = rand_rows(rng, 1, 1 << 20)
        B = rand_rows(rng, 1, 1 << 20)
        ba, Aa = g.upload_rows(A)
        bb, Ba = g.upload_rows(B)
        bc, Ca = g.upload_rows(np.zeros((1, 2 << 20), dtype=np.uint32))
        g.check(ctx.mul(Aa, Ba, Ca), "mul")
        g.synchronize()
        for _ in range(2):
            assert all(np.array_equal(w, x) for w, x in zip(want, stage_rows(st)))
    finally:
        st.close()


def test_bad_arguments(ctx):
    out = ctypes.c_void_p()
    lib = g.lib()
    assert lib.gpupi_leaf_stage_create(ctx.ptr, 3, g.LEAF_PLAIN, ctypes.byref(out)) != 0  # not a power of two
    assert lib.gpupi_leaf_stage_create(ctx.ptr, 4, 7, ctypes.byref(out)) != 0  # unknown form
    assert lib.gpupi_leaf_stage_create(None, 4, g.LEAF_PLAIN, ctypes.byref(out)) != 0
