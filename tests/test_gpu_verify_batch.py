"""Batched verification with the Merkle path hashing on the GPU
(z0p_air_verify_batch_device): for honest proofs and for forgeries of every
kind (path digests, FRI values, final layer, trace values, nonce, statement,
truncation), each instance's status must equal what the host verifier
returns for it alone, including which error comes first."""
import pytest

import oracle_ffi as o
import wire
from helpers import P, seeded_claims

import z0p_amd as z

pytestmark = pytest.mark.gpu


def _instances():
    out = []
    for k, seed in [(10, 1), (11, 2), (12, 3), (13, 4), (14, 5), (10, 6)]:
        a0, b0 = seeded_claims(seed)
        n = 1 << k
        bl = o.fib_b_last(a0, b0, n)
        label = b"vb/%d" % seed
        proof = o.air_prove(o.air_params_for_security(n, a0, b0, bl), o.Transcript(label))
        out.append((z.air.for_security(n, a0, b0, bl), proof, label))
    return out


def _forgeries(params, proof, label):
    d = lambda: wire.AirProof.from_bytes(proof)
    out = []
    bad = d()
    dg, left = bad.fri_proof.queries[3][2].path[1]
    bad.fri_proof.queries[3][2].path[1] = (bytes([dg[0] ^ 4]) + dg[1:], left)
    out.append(bad.to_bytes())
    bad = d()
    fw = bad.fri_proof.queries[5][1]
    fw.values[0] = wire.ext_add_one(fw.values[0])
    out.append(bad.to_bytes())
    bad = d()
    fw = bad.fri_proof.queries[7][3]
    fw.values[2] = wire.ext_add_one(fw.values[2])
    out.append(bad.to_bytes())
    bad = d()
    bad.fri_proof.final_poly[1] = wire.ext_add_one(bad.fri_proof.final_poly[1])
    out.append(bad.to_bytes())
    bad = d()
    bad.openings[4].b_next = (bad.openings[4].b_next + 1) % P
    out.append(bad.to_bytes())
    bad = d()
    dg, left = bad.openings[9].shift_path[0]
    bad.openings[9].shift_path[0] = (bytes([dg[0] ^ 1]) + dg[1:], left)
    out.append(bad.to_bytes())
    bad = d()
    leaf = bytearray(bad.openings[2].leaf)
    leaf[63] ^= 0x80
    bad.openings[2].leaf = bytes(leaf)
    out.append(bad.to_bytes())
    bad = d()
    bad.fri_proof.pow_nonce ^= 1 << 40
    out.append(bad.to_bytes())
    out.append(proof[:-3])
    out.append(proof[:1000] + bytes([proof[1000] ^ 1]) + proof[1001:])
    return out


def _host_status(params, proof, label):
    try:
        z.air.verify(params, proof, z.Transcript(label))
        return 0
    except z.Z0pError as e:
        return e.code


def test_device_batch_verify_matches_host(ctx):
    inst = _instances()
    ps, proofs, labels = [], [], []
    for params, proof, label in inst:
        ps.append(params)
        proofs.append(proof)
        labels.append(label)
        for bad in _forgeries(params, proof, label):
            ps.append(params)
            proofs.append(bad)
            labels.append(label)
    # a false statement and a wrong label
    p0, pr0, l0 = inst[0]
    ps.append(z.air.for_security(p0.n_steps, p0.a0, p0.b0, (p0.b_last + 1) % P))
    proofs.append(pr0)
    labels.append(l0)
    ps.append(p0)
    proofs.append(pr0)
    labels.append(b"other")
    want = [_host_status(p, pr, l) for p, pr, l in zip(ps, proofs, labels)]
    assert want.count(0) == len(inst) and 6 in want and 7 in want and 8 in want, want
    print("instances", len(ps), "statuses", sorted(set(want)), flush=True)
    for threads in (0, 1, 5):
        got = z.air.verify_batch_device(ps, proofs, [z.Transcript(l) for l in labels], ctx, threads=threads)
        assert [0 if g is None else g.code for g in got] == want, threads
    cpu = z.air.verify_batch(ps, proofs, [z.Transcript(l) for l in labels])
    assert [0 if g is None else g.code for g in cpu] == want


def test_device_batch_verify_chunks_and_large_proofs(ctx):
    """Several 64-proof chunks (the host verifies one chunk while the
    device hashes the other), with two 2^20-step proofs among them."""
    ps, proofs, labels = [], [], []
    for i in range(600):
        k = 10 + (i % 3)
        a0, b0 = seeded_claims(100 + i % 7)
        n = 1 << k
        ps.append(z.air.for_security(n, a0, b0, o.fib_b_last(a0, b0, n)))
        labels.append(b"vbc/%d" % (i % 7))
    cache = {}
    for p, l in zip(ps, labels):
        key = (p.n_steps, p.a0, l)
        if key not in cache:
            cache[key] = z.air.prove(p, z.Transcript(l), ctx)
        proofs.append(cache[key])
    big = z.air.for_security(1 << 20, 1, 1, o.fib_b_last(1, 1, 1 << 20))
    big_proof = z.air.prove(big, z.Transcript(b"bench/air-scale"), ctx)
    for at in (17, 431):
        ps[at], proofs[at], labels[at] = big, big_proof, b"bench/air-scale"
    bad = bytearray(proofs[300])
    bad[len(bad) // 2] ^= 2
    proofs[300] = bytes(bad)
    print("verifying", len(ps), "proofs", flush=True)
    got = z.air.verify_batch_device(ps, proofs, [z.Transcript(l) for l in labels], ctx)
    want = z.air.verify_batch(ps, proofs, [z.Transcript(l) for l in labels])
    assert [None if g is None else g.code for g in got] == [None if g is None else g.code for g in want]
    assert sum(g is None for g in got) == 599
