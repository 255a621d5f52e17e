"""Kernel schedules and the autotuner: launch shapes change speed, never
bytes. Proofs under every schedule knob's extremes must equal the oracle's,
the tree kernels must reproduce every level, and the tuner must leave a
valid schedule under which proofs are still byte-identical."""
import numpy as np
import pytest

import oracle_ffi as o
from helpers import fri_bench_coeffs, seeded_claims

import z0p_amd as z
from z0p_amd import _ffi

pytestmark = pytest.mark.gpu

DEFAULT = dict(tail_max_log=17, tail_group_max_log=7, tail_spread_log=0, levels_fuse=2, small_max_log=10,
               grind_blocks=4096, fold_fuse_max_log=0, compose_fuse=1, trace_fuse=1, subtree_levels=0,
               compose_subtree=0, graph=0, lde_split=1, grind_pool=1024, fold_commit_max_log=18, tree_top=0,
               wave_streams=0)

SCHEDULES = [
    dict(DEFAULT, tail_group_max_log=9, tail_spread_log=8),
    dict(DEFAULT, tail_group_max_log=2, tail_max_log=12),
    dict(DEFAULT, tail_group_max_log=5, tail_spread_log=6, levels_fuse=1),
    dict(DEFAULT, tail_max_log=4, levels_fuse=2, small_max_log=0),
    dict(DEFAULT, tail_max_log=20, tail_group_max_log=8, small_max_log=4, grind_blocks=16),
    dict(DEFAULT, small_max_log=8, grind_blocks=1 << 20),
    dict(DEFAULT, fold_fuse_max_log=16, compose_fuse=0, trace_fuse=0),
    dict(DEFAULT, fold_fuse_max_log=27, small_max_log=0),
    dict(DEFAULT, subtree_levels=8),
    dict(DEFAULT, subtree_levels=3, tail_max_log=10),
    dict(DEFAULT, subtree_levels=8, tail_max_log=4, fold_fuse_max_log=22, small_max_log=0, compose_subtree=8),
    dict(DEFAULT, compose_subtree=2, tail_max_log=13),
    dict(DEFAULT, subtree_levels=1, compose_fuse=0, trace_fuse=0, tail_group_max_log=3),
    dict(DEFAULT, lde_split=0),
    dict(DEFAULT, graph=1, lde_split=0),
    dict(DEFAULT, graph=1, fold_fuse_max_log=22, subtree_levels=6, tail_max_log=10),
    dict(DEFAULT, grind_pool=0),
    dict(DEFAULT, grind_pool=1, grind_blocks=16),
    dict(DEFAULT, grind_pool=8192, small_max_log=0),
    dict(DEFAULT, fold_commit_max_log=0),
    dict(DEFAULT, fold_commit_max_log=27, small_max_log=0, tail_max_log=4),
    dict(DEFAULT, fold_commit_max_log=8, small_max_log=4, fold_fuse_max_log=22, subtree_levels=8),
    dict(DEFAULT, fold_commit_max_log=11, small_max_log=10),
    dict(DEFAULT, tree_top=1),
    dict(DEFAULT, tree_top=1, tail_max_log=4, subtree_levels=3),
    dict(DEFAULT, tree_top=1, tail_max_log=24, fold_commit_max_log=0, small_max_log=0),
    dict(DEFAULT, wave_streams=1),
    dict(DEFAULT, wave_streams=1, graph=1, grind_pool=0),
]


@pytest.fixture
def sched_ctx(ctx):
    saved = ctx.get_schedule()
    yield ctx
    ctx.set_schedule(saved)


def test_default_schedule(sched_ctx):
    assert sched_ctx.get_schedule() == DEFAULT


@pytest.mark.parametrize("sched", SCHEDULES, ids=lambda s: "g%d_s%d_t%d_f%d_sm%d_gb%d_ff%d_c%d%d_st%d_cs%d_h%d" % (
    s["tail_group_max_log"], s["tail_spread_log"], s["tail_max_log"], s["levels_fuse"], s["small_max_log"],
    s["grind_blocks"], s["fold_fuse_max_log"], s["compose_fuse"], s["trace_fuse"], s["subtree_levels"],
    s["compose_subtree"], s["graph"] + 2 * s["lde_split"]) + "_p%d_fc%d_tt%d_ws%d" % (s["grind_pool"], s["fold_commit_max_log"], s["tree_top"],
                                                     s["wave_streams"]))
def test_proofs_identical_under_schedule(sched_ctx, sched):
    ctx = sched_ctx
    ctx.set_schedule(sched)
    print("schedule", sched, flush=True)
    # one pipelined AIR proof
    n = 1 << 16
    a0, b0 = seeded_claims(5)
    bl = o.fib_b_last(a0, b0, n)
    got = z.air.prove(z.air.for_security(n, a0, b0, bl), z.Transcript(b"sched"), ctx)
    assert got == o.air_prove(o.air_params_for_security(n, a0, b0, bl), o.Transcript(b"sched"))
    # a mixed batch
    sizes = [1 << 10, 1 << 13, 1 << 11, 1 << 14]
    ps = [z.air.for_security(k, 1, 2, o.fib_b_last(1, 2, k)) for k in sizes]
    got = z.air.prove_batch(ps, [z.Transcript(b"b%d" % i) for i in range(len(ps))], ctx)
    for i, k in enumerate(sizes):
        want = o.air_prove(o.air_params_for_security(k, 1, 2, o.fib_b_last(1, 2, k)), o.Transcript(b"b%d" % i))
        assert got[i] == want, k
    # a FRI commitment (base-field layer 0)
    coeffs = fri_bench_coeffs(1 << 12)
    zp = z.fri.for_security(1 << 12)
    cm, pf = z.fri.commit_open(coeffs, zp, z.Transcript(b"f"), ctx)
    ocm, opf = o.fri_commit_open(coeffs, o.fri_params_for_security(1 << 12), o.Transcript(b"f"))
    assert (cm, pf) == (ocm, opf)


@pytest.mark.parametrize("leaves", [2, 4, 64, 1 << 9, 1 << 11, 1 << 16, 1 << 19])
@pytest.mark.parametrize("tail", [(2, 0, 4, 8), (9, 0, 17, 8), (9, 8, 17, 0), (6, 6, 10, 8), (8, 0, 2, 8),
                                  (8, 0, 12, 2)])
def test_merkle_levels_under_tail_shapes(sched_ctx, leaves, tail):
    """Every level of a 64-byte-leaf tree: leaves with their workgroups'
    subtrees (up to 9 levels), node-level jobs reducing 10 levels per
    workgroup, and quad-lane tails, in every split the knobs allow."""
    ctx = sched_ctx
    ctx.set_schedule(dict(DEFAULT, tail_group_max_log=tail[0], tail_spread_log=tail[1], tail_max_log=tail[2],
                          subtree_levels=tail[3]))
    rng = np.random.default_rng(leaves + tail[0])
    flat = rng.integers(0, 256, leaves * 64, dtype=np.uint8).tobytes()
    out = (np.zeros(32 * (2 * leaves - 1), dtype=np.uint8))
    print("merkle", leaves, tail, flush=True)
    rc = _ffi.lib().z0p_test_merkle64(ctx.handle, _ffi.buf(flat), leaves,
                                      out.ctypes.data_as(_ffi.u8p))
    assert rc == 0
    assert out.tobytes() == o.merkle_flat64(flat)


def test_set_schedule_rejects_out_of_range(sched_ctx):
    for bad in [dict(DEFAULT, tail_group_max_log=10), dict(DEFAULT, levels_fuse=3), dict(DEFAULT, small_max_log=11),
                dict(DEFAULT, small_max_log=2), dict(DEFAULT, grind_blocks=8), dict(DEFAULT, tail_max_log=1),
                dict(DEFAULT, fold_fuse_max_log=2), dict(DEFAULT, fold_fuse_max_log=28), dict(DEFAULT, compose_fuse=2),
                dict(DEFAULT, trace_fuse=-1), dict(DEFAULT, subtree_levels=9), dict(DEFAULT, subtree_levels=-1),
                dict(DEFAULT, compose_subtree=9), dict(DEFAULT, graph=2), dict(DEFAULT, lde_split=2),
                dict(DEFAULT, grind_pool=-1), dict(DEFAULT, grind_pool=8193), dict(DEFAULT, fold_commit_max_log=7),
                dict(DEFAULT, fold_commit_max_log=28), dict(DEFAULT, tree_top=2), dict(DEFAULT, wave_streams=2),
                dict(DEFAULT, wave_streams=-1)]:
        with pytest.raises(z.Z0pError) as e:
            sched_ctx.set_schedule(bad)
        assert e.value.name == "InvalidArgument"
    assert sched_ctx.get_schedule() == DEFAULT


def test_autotune_keeps_proofs_identical(sched_ctx):
    ctx = sched_ctx
    n = 1 << 12
    p = z.air.for_security(n, 1, 1, o.fib_b_last(1, 1, n))
    sched, best_ms, start_ms, log = ctx.autotune([p], reps=1)
    print("tuned", sched, best_ms, start_ms, flush=True)
    assert ctx.get_schedule() == sched
    assert best_ms > 0 and start_ms > 0
    assert log.count(";") >= 10  # every knob tried at least once
    assert z.air.prove(p, z.Transcript(b"t"), ctx) == o.air_prove(o.air_params_for_security(n, 1, 1, o.fib_b_last(
        1, 1, n)), o.Transcript(b"t"))
    # a batch workload tunes too
    ps = [z.air.for_security(k, 1, 1, o.fib_b_last(1, 1, k)) for k in (1 << 10, 1 << 11)]
    sched2, _, _, _ = ctx.autotune(ps, reps=1)
    assert ctx.get_schedule() == sched2
