"""The batched and served workloads at their stated sizes, every proof
checked: 256 Fibonacci AIR proofs with 2^10..2^18 steps proved as one batch,
and a Poisson stream of 2^10..2^20-step requests served under a fixed device
memory budget. Each result must equal the single-proof GPU path's bytes for
the same request (whose 2^10 and 2^20 outputs the golden fixtures pin) and be
accepted by the host verifier; a sample is also compared with the oracle."""
import random
import threading
import time

import pytest

import oracle_ffi as o
from helpers import P, splitmix64

import z0p_amd as z

pytestmark = pytest.mark.gpu


def _instance(k, seed, label):
    """(params, label) of 2^k steps with (a0, b0) drawn from the seed."""
    st, x = splitmix64(seed)
    st, y = splitmix64(st)
    a0, b0 = x % P, y % P
    n = 1 << k
    return z.air.for_security(n, a0, b0, z.air.b_last(a0, b0, n)), label


def _batch_workload(count=256, seed=2024):
    """The batched benchmark's draw: log2 steps 10 + splitmix % 9."""
    st, out = seed, []
    for i in range(count):
        st, r = splitmix64(st)
        out.append(_instance(10 + r % 9, seed * 7919 + i, b"batch/%d" % i))
    return out


def _oracle_check(params, label, proof):
    op = o.air_params_new(params.n_steps, params.blowup, params.num_queries, params.final_size, params.a0,
                          params.b0, params.b_last)
    op.grinding_bits, op.cap_height = params.grinding_bits, params.cap_height
    assert o.air_prove(op, o.Transcript(label)) == proof, params.n_steps


def test_config4_batch_256_proofs(ctx):
    inst = _batch_workload()
    ks = sorted({p.n_steps.bit_length() - 1 for p, _ in inst})
    print("batch: 256 proofs, log2 steps", ks, "total steps", sum(p.n_steps for p, _ in inst), flush=True)
    assert ks[0] >= 10 and ks[-1] <= 18 and len(ks) >= 8
    trs = [z.Transcript(l) for _, l in inst]
    got = z.air.prove_batch([p for p, _ in inst], trs, ctx)
    print("batch proved; comparing with the single-proof path", flush=True)
    for (p, l), g, t in zip(inst, got, trs):
        assert isinstance(g, bytes), (p.n_steps, g)
        single_t = z.Transcript(l)
        assert g == z.air.prove(p, single_t, ctx), p.n_steps
        assert t.challenge_bytes(b"after") == single_t.challenge_bytes(b"after")
    ok = z.air.verify_batch([p for p, _ in inst], got, [z.Transcript(l) for _, l in inst])
    assert all(v is None for v in ok)
    # The oracle on the smallest and the largest instances of the batch.
    small = min(range(len(inst)), key=lambda i: inst[i][0].n_steps)
    large = max(range(len(inst)), key=lambda i: inst[i][0].n_steps)
    for i in (small, large):
        print("oracle check n", inst[i][0].n_steps, flush=True)
        _oracle_check(inst[i][0], inst[i][1], got[i])


def test_config5_served_stream_with_2p20_requests(ctx):
    """Open-loop Poisson arrivals (log2 steps uniform in 10..20, plus two
    2^20-step requests) against a 4 GiB budget; every served proof equals
    the single-proof path's and verifies."""
    rng = random.Random(5)
    reqs = [_instance(rng.randint(10, 20), rng.getrandbits(63), b"serve/t/%d" % i) for i in range(46)]
    reqs[7] = _instance(20, 1, b"serve/t/big0")
    reqs[30] = _instance(20, 2, b"serve/t/big1")
    reqs[12] = _instance(18, 3, b"serve/t/mid")
    assert sum(1 for p, _ in reqs if p.n_steps == 1 << 20) >= 2
    assert {p.n_steps.bit_length() - 1 for p, _ in reqs} >= {10, 18, 19, 20}
    qps, t, arrivals = 600.0, 0.0, []
    for _ in reqs:
        t += rng.expovariate(qps)
        arrivals.append(t)
    sctx = z.Context(0)
    results = {}
    with z.Server(sctx, hbm_budget=4 << 30) as srv:
        ids = {}

        def producer():
            t0 = time.perf_counter()
            for i, ((p, l), at) in enumerate(zip(reqs, arrivals)):
                d = t0 + at - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ids[srv.submit(p, z.Transcript(l))] = i

        th = threading.Thread(target=producer)
        th.start()
        deadline = time.time() + 120
        while len(results) < len(reqs) and time.time() < deadline:
            r = srv.poll(1000)
            if r is not None:
                results[r[0]] = r[1]
        th.join()
    sctx.close()
    assert len(results) == len(reqs)
    print("served", len(results), "requests; comparing", flush=True)
    by_index = {ids[rid]: proof for rid, proof in results.items()}
    proofs = []
    for i, (p, l) in enumerate(reqs):
        got = by_index[i]
        assert isinstance(got, bytes), (i, got)
        want = z.air.prove(p, z.Transcript(l), ctx)
        assert got == want, (i, p.n_steps)
        proofs.append(got)
    ok = z.air.verify_batch([p for p, _ in reqs], proofs, [z.Transcript(l) for _, l in reqs])
    assert all(v is None for v in ok)
    assert len(by_index[7]) == len(by_index[30]) == 203553
