"""Batched and served proving: every proof must be byte-identical to the one
the single-instance path (and the oracle) produces for the same request,
whatever it ran beside, however the scheduler admitted it, and each
transcript must end in the state the single call leaves."""
import random
import threading
import time

import pytest

import oracle_ffi as o
from helpers import seeded_claims, splitmix64

import z0p_amd as z

pytestmark = pytest.mark.gpu


def _instances(count, kmin, kmax, seed):
    """Log-uniform step counts 2^kmin..2^kmax and seeded (a0, b0), as the
    batched workload draws them."""
    out = []
    st = seed
    for i in range(count):
        st, r = splitmix64(st)
        k = kmin + r % (kmax - kmin + 1)
        n = 1 << k
        a0, b0 = seeded_claims(seed * 1000 + i)
        bl = o.fib_b_last(a0, b0, n)
        out.append((n, a0, b0, bl, b"batch/%d" % i))
    return out


def _oracle(n, a0, b0, bl, label):
    return o.air_prove(o.air_params_for_security(n, a0, b0, bl), o.Transcript(label))


def test_batch_mixed_sizes_byte_identical(ctx):
    inst = _instances(24, 6, 14, 7)
    print("batch sizes", [i[0] for i in inst], flush=True)
    params = [z.air.for_security(n, a0, b0, bl) for n, a0, b0, bl, _ in inst]
    trs = [z.Transcript(lab) for *_, lab in inst]
    got = z.air.prove_batch(params, trs, ctx)
    for (n, a0, b0, bl, lab), g, t in zip(inst, got, trs):
        assert isinstance(g, bytes), (n, g)
        ot = o.Transcript(lab)
        want = o.air_prove(o.air_params_for_security(n, a0, b0, bl), ot)
        assert g == want, n
        assert t.challenge_bytes(b"after") == ot.challenge_bytes(b"after")


def test_batch_large_and_small_together(ctx):
    """A 2^20 proof beside small ones equals the single-instance proof."""
    inst = [(1 << 20, 1, 1, o.fib_b_last(1, 1, 1 << 20), b"bench/air-scale")]
    inst += _instances(6, 10, 16, 3)
    params = [z.air.for_security(n, a0, b0, bl) for n, a0, b0, bl, _ in inst]
    got = z.air.prove_batch(params, [z.Transcript(lab) for *_, lab in inst], ctx)
    for (n, a0, b0, bl, lab), g, p in zip(inst, got, params):
        print("compare n", n, flush=True)
        assert g == z.air.prove(p, z.Transcript(lab), ctx), n
    assert len(got[0]) == 203553


def test_batch_under_tight_budget(ctx):
    """A budget that holds only a few proofs at once: later instances are
    admitted as earlier ones finish, with identical results."""
    inst = _instances(12, 8, 12, 11)
    params = [z.air.for_security(n, a0, b0, bl) for n, a0, b0, bl, _ in inst]
    want = [z.air.prove(p, z.Transcript(lab), ctx) for p, (*_, lab) in zip(params, inst)]
    ctx.set_hbm_budget(24 << 20)
    try:
        got = z.air.prove_batch(params, [z.Transcript(lab) for *_, lab in inst], ctx)
    finally:
        ctx.set_hbm_budget(0)
    assert got == want


def test_batch_reports_per_instance_errors(ctx):
    inst = _instances(3, 6, 8, 5)
    params = [z.air.for_security(n, a0, b0, bl) for n, a0, b0, bl, _ in inst]
    bad = z.air.Params()
    bad.n_steps, bad.blowup, bad.num_queries, bad.final_size = 3, 16, 4, 4  # not a power of two
    got = z.air.prove_batch(params[:1] + [bad] + params[1:], [z.Transcript(b"e%d" % i) for i in range(4)], ctx)
    assert isinstance(got[1], z.Z0pError) and got[1].name == "InvalidDomainSize"
    for g, p, i in zip([got[0], got[2], got[3]], params, [0, 2, 3]):
        assert g == z.air.prove(p, z.Transcript(b"e%d" % i), ctx)


def test_batch_rejects_shared_transcript(ctx):
    p = z.air.for_security(64, 1, 1, o.fib_b_last(1, 1, 64))
    t = z.Transcript(b"shared")
    with pytest.raises(z.Z0pError) as e:
        z.air.prove_batch([p, p], [t, t], ctx)
    assert e.value.name == "InvalidArgument"


def test_batch_pooled_grind_across_windows(ctx):
    """Several 26-bit proofs-of-work in one launch go through the pooled
    search (every wave takes chunks of any unfinished proof); their smallest
    nonces lie beyond the first window of 2^24 nonces for some of them, so
    the search also continues window by window. Each proof must equal its
    single-instance proof (whose 26-bit grind is pinned to the oracle) and
    the per-proof workgroup search of the same batch."""
    n = 64
    inst = [(2, 3, b"g26a"), (5, 8, b"g26b"), (1, 1, b"g26c"), (7, 9, b"g26d")]
    params = []
    for a0, b0, _ in inst:
        p = z.air.params_new(n, 8, 6, 8, a0, b0, o.fib_b_last(a0, b0, n))
        p.grinding_bits = 26
        params.append(p)
    saved = ctx.get_schedule()
    try:
        ctx.set_schedule(dict(saved, grind_pool=1024))
        print("pooled 26-bit batch", flush=True)
        pooled = z.air.prove_batch(params, [z.Transcript(lab) for *_, lab in inst], ctx)
        ctx.set_schedule(dict(saved, grind_pool=0))
        print("per-proof 26-bit batch", flush=True)
        split = z.air.prove_batch(params, [z.Transcript(lab) for *_, lab in inst], ctx)
    finally:
        ctx.set_schedule(saved)
    assert pooled == split
    for p, (_, _, lab), g in zip(params, inst, pooled):
        assert g == z.air.prove(p, z.Transcript(lab), ctx), lab
    # The first instance's smallest nonce (103,769,303) is pinned by the
    # single-proof test against the oracle.
    assert pooled[0] == z.air.prove(params[0], z.Transcript(b"g26a"), ctx)


def test_batch_and_server_zero_copy_views(ctx):
    """zero_copy hands the C ABI's buffers over as read-only memoryviews:
    equal to the bytes results, and released when dropped."""
    import gc
    inst = _instances(6, 6, 12, 5)
    params = [z.air.for_security(n, a0, b0, bl) for n, a0, b0, bl, _ in inst]
    want = z.air.prove_batch(params, [z.Transcript(lab) for *_, lab in inst], ctx)
    views = z.air.prove_batch(params, [z.Transcript(lab) for *_, lab in inst], ctx, zero_copy=True)
    for v, w in zip(views, want):
        assert isinstance(v, memoryview) and v.readonly
        assert v == w and bytes(v) == w
    assert z.air.verify_batch(params, views, [z.Transcript(lab) for *_, lab in inst]) == [None] * len(inst)
    del views
    gc.collect()
    sctx = z.Context(0)
    with z.Server(sctx, hbm_budget=1 << 30) as srv:
        ids = {srv.submit(p, z.Transcript(lab)): i for i, (p, (*_, lab)) in enumerate(zip(params, inst))}
        got = {}
        deadline = time.time() + 60
        while len(got) < len(inst) and time.time() < deadline:
            r = srv.poll(200, zero_copy=True)
            if r is not None:
                got[r[0]] = r[1]
        assert len(got) == len(inst)
        for rid, v in got.items():
            assert isinstance(v, memoryview) and v == want[ids[rid]]
    sctx.close()


def _serve(requests, budget, gaps_s, waves=None, large_schedule=None):
    """Submit from a producer thread with the given inter-arrival gaps, poll
    from this thread; returns {request index: (proof, latency, queue)}."""
    sctx = z.Context(0)
    if waves is not None:
        sctx.set_waves(waves)
    out = {}
    with z.Server(sctx, hbm_budget=budget, large_schedule=large_schedule) as srv:
        ids = {}

        def producer():
            for i, (p, lab) in enumerate(requests):
                time.sleep(gaps_s[i])
                ids[srv.submit(p, z.Transcript(lab))] = i

        th = threading.Thread(target=producer)
        th.start()
        deadline = time.time() + 120
        while len(out) < len(requests) and time.time() < deadline:
            r = srv.poll(200)
            if r is None:
                continue
            rid, res, lat, q = r
            while rid not in ids:  # submit() returned after the worker finished it
                time.sleep(0.001)
            out[ids[rid]] = (res, lat, q)
        th.join()
        assert srv.pending() == 0
    sctx.close()
    return out


def test_server_same_alone_and_in_mix(ctx):
    inst = _instances(20, 6, 16, 21)
    params = [z.air.for_security(n, a0, b0, bl) for n, a0, b0, bl, _ in inst]
    alone = [z.air.prove(p, z.Transcript(lab), ctx) for p, (*_, lab) in zip(params, inst)]
    rng = random.Random(4)
    gaps = [rng.expovariate(1000.0) for _ in inst]  # ~1 ms mean gap: arrivals overlap running work
    print("serving", [i[0] for i in inst], flush=True)
    out = _serve([(p, lab) for p, (*_, lab) in zip(params, inst)], 0, gaps)
    assert sorted(out) == list(range(len(inst)))
    for i, a in enumerate(alone):
        res, lat, q = out[i]
        assert res == a, i
        assert lat >= q >= 0


def test_server_tight_budget_queues_and_matches(ctx):
    inst = _instances(10, 10, 12, 33)
    params = [z.air.for_security(n, a0, b0, bl) for n, a0, b0, bl, _ in inst]
    alone = [z.air.prove(p, z.Transcript(lab), ctx) for p, (*_, lab) in zip(params, inst)]
    out = _serve([(p, lab) for p, (*_, lab) in zip(params, inst)], 24 << 20, [0.0] * len(inst))
    assert [out[i][0] for i in range(len(inst))] == alone
    assert max(out[i][2] for i in range(len(inst))) > 0.0  # some request waited for memory


def test_server_rejects_invalid_request(ctx):
    sctx = z.Context(0)
    with z.Server(sctx) as srv:
        bad = z.air.Params()
        bad.n_steps, bad.blowup, bad.num_queries, bad.final_size = 64, 3, 4, 4
        with pytest.raises(z.Z0pError) as e:
            srv.submit(bad, z.Transcript(b"x"))
        assert e.value.name == "InvalidBlowup"
        assert srv.pending() == 0
    sctx.close()


@pytest.mark.parametrize("waves", [0, 2])
def test_batch_waves_on_and_off(ctx, waves):
    """One launch for all proofs, or two alternating waves (from 2 active
    proofs, also under a budget that holds only a few): the same bytes and
    the same transcript states."""
    inst = _instances(10, 6, 13, 17)
    params = [z.air.for_security(n, a0, b0, bl) for n, a0, b0, bl, _ in inst]
    ots = [o.Transcript(lab) for *_, lab in inst]
    want = [o.air_prove(o.air_params_for_security(n, a0, b0, bl), t) for (n, a0, b0, bl, _), t in zip(inst, ots)]
    ctx.set_waves(waves)
    try:
        trs = [z.Transcript(lab) for *_, lab in inst]
        got = z.air.prove_batch(params, trs, ctx)
        ctx.set_hbm_budget(16 << 20)
        got_tight = z.air.prove_batch(params, [z.Transcript(lab) for *_, lab in inst], ctx)
    finally:
        ctx.set_hbm_budget(0)
        ctx.set_waves(8)
    assert got == want
    assert got_tight == want
    for t, ot in zip(trs, ots):
        assert t.challenge_bytes(b"after") == ot.challenge_bytes(b"after")


def test_server_two_waves_match_alone(ctx):
    """Both lanes alternating two waves from 2 active proofs: every result is
    the proof the request gets alone."""
    inst = _instances(24, 8, 17, 41)
    params = [z.air.for_security(n, a0, b0, bl) for n, a0, b0, bl, _ in inst]
    alone = [z.air.prove(p, z.Transcript(lab), ctx) for p, (*_, lab) in zip(params, inst)]
    out = _serve([(p, lab) for p, (*_, lab) in zip(params, inst)], 0, [0.0] * len(inst), waves=2)
    assert [out[i][0] for i in range(len(inst))] == alone


def test_wave_streams_match_alone(ctx):
    """The two waves on two streams (schedule wave_streams): kernels of both
    waves run side by side, proofs are admitted on the idle stream, and
    every batch and served result is the proof the request gets alone, also
    under a budget that admits proofs while a wave is in flight."""
    inst = _instances(16, 8, 17, 53)
    params = [z.air.for_security(n, a0, b0, bl) for n, a0, b0, bl, _ in inst]
    alone = [z.air.prove(p, z.Transcript(lab), ctx) for p, (*_, lab) in zip(params, inst)]
    saved = ctx.get_schedule()
    ctx.set_waves(2)
    try:
        ctx.set_schedule(dict(saved, wave_streams=1))
        print("two streams, free budget", flush=True)
        got = z.air.prove_batch(params, [z.Transcript(lab) for *_, lab in inst], ctx)
        ctx.set_hbm_budget(48 << 20)
        print("two streams, tight budget", flush=True)
        got_tight = z.air.prove_batch(params, [z.Transcript(lab) for *_, lab in inst], ctx)
    finally:
        ctx.set_hbm_budget(0)
        ctx.set_waves(8)
        ctx.set_schedule(saved)
    assert got == alone
    assert got_tight == alone
    print("server, two streams", flush=True)
    sctx = z.Context(0)
    sctx.set_schedule(dict(sctx.get_schedule(), wave_streams=1))
    sctx.set_waves(2)
    out = {}
    # 1 GiB: the large lane's 512 MiB holds a few 2^17-step proofs at once.
    with z.Server(sctx, hbm_budget=1 << 30) as srv:
        ids = {srv.submit(p, z.Transcript(lab)): i for i, (p, (*_, lab)) in enumerate(zip(params, inst))}
        deadline = time.time() + 120
        while len(out) < len(inst) and time.time() < deadline:
            r = srv.poll(200)
            if r is not None:
                out[ids[r[0]]] = r[1]
    sctx.close()
    assert [out.get(i) for i in range(len(inst))] == alone


def test_server_large_lane_schedule(ctx):
    """A large lane under a schedule of its own (every fusion and launch knob
    away from the default) returns each request's proof byte-identical to
    z0p_air_prove under the context's schedule; an out-of-range large-lane
    schedule is refused at creation."""
    inst = _instances(14, 10, 18, 33)
    params = [z.air.for_security(n, a0, b0, bl) for n, a0, b0, bl, _ in inst]
    alone = [z.air.prove(p, z.Transcript(lab), ctx) for p, (*_, lab) in zip(params, inst)]
    large = dict(ctx.get_schedule())
    large.update(tail_max_log=12, tail_group_max_log=7, fold_fuse_max_log=22, subtree_levels=2, compose_subtree=1,
                 levels_fuse=1, fold_commit_max_log=0, wave_streams=1, grind_pool=512)
    assert large != ctx.get_schedule()
    rng = random.Random(9)
    gaps = [rng.expovariate(2000.0) for _ in inst]
    print("serving with a large-lane schedule", sorted(p.n_steps.bit_length() - 1 for p in params), flush=True)
    out = _serve([(p, lab) for p, (*_, lab) in zip(params, inst)], 4 << 30, gaps, large_schedule=large)
    assert sorted(out) == list(range(len(inst)))
    assert any(p.n_steps >= 1 << 16 for p in params) and any(p.n_steps < 1 << 16 for p in params)
    for i, a in enumerate(alone):
        assert out[i][0] == a, params[i].n_steps
    bad = dict(large, levels_fuse=3)
    sctx = z.Context(0)
    with pytest.raises(z.Z0pError):
        z.Server(sctx, hbm_budget=1 << 30, large_schedule=bad)
    sctx.close()
