"""Pins the oracle's IVC accumulator (src/fold.rs) without running the
reference: the wire size of the reference bench's proof (tools/baseline.json
"fold 4x8 fib": 16845 bytes, benches/speed_size.rs:432-480), the tuner's
(blowup, queries, grinding) for the reference's for_security case, and the
reference's acceptance and rejection tests (fold.rs:907-1233): chain links,
the out_a boundary (a forged A jump with a consistent FRI is still refused),
a poisoned segment, tampered and shifted openings, caps and transcripts."""
import pytest

import oracle_ffi as o


def _params(out_b, **kw):
    return o.fold_params_new(8, 8, 8, 4, 1, 1, out_b).replace(**kw)


def _last_b(seg, count):
    return o.chained_traces(seg, count)[-1][1][-1]


def _prove(params, traces, label):
    t = o.Transcript(label)
    acc = o.FoldAccumulator(params)
    for ca, cb in traces:
        acc.add_segment(ca, cb, t)
    return acc, acc.finalize(t)


def _ok(params, label, proof):
    return o.fold_verify(params, o.Transcript(label), proof) == 0


def test_bench_proof_size_matches_reference_baseline():
    traces = o.chained_traces(8, 4)
    p = _params(traces[3][1][7])
    _, proof = _prove(p, traces, b"bench/fold")
    assert len(proof) == 16845
    assert _ok(p, b"bench/fold", proof)


def test_rejects_bad_params():
    # fold.rs:917-924
    for args, code in [((3, 8, 8, 4), 3), ((8, 2, 8, 4), 4), ((8, 8, 0, 4), 5), ((1 << 63, 4, 8, 4), 3)]:
        with pytest.raises(o.OracleError) as e:
            o.fold_params_new(*args, 1, 1, 1)
        assert e.value.code == code, args


def test_empty_accumulator_and_segment_shape():
    # fold.rs:984-995
    acc = o.FoldAccumulator(_params(1))
    assert len(acc) == 0
    with pytest.raises(o.OracleError) as e:
        acc.finalize(o.Transcript(b"fold-empty"))
    assert e.value.code == 1
    with pytest.raises(o.OracleError) as e:
        acc.add_segment([1] * 7, [1] * 7, o.Transcript(b"x"))
    assert e.value.code == 8


def test_honest_ivc_and_for_security():
    # fold.rs:997-1031
    traces = o.chained_traces(8, 3)
    p = _params(_last_b(8, 3))
    acc, proof = _prove(p, traces, b"fold-ivc")
    assert len(acc) == 3 and int.from_bytes(proof[1:5], "little") == 3
    assert _ok(p, b"fold-ivc", proof)
    q = o.fold_params_for_security(8, 1, 1, _last_b(8, 2), 12)
    assert q.blowup == 16 and (q.num_queries, q.grinding_bits) == (3, 6) and q.cap_height <= 4
    _, proof = _prove(q, o.chained_traces(8, 2), b"fold-forsec")
    assert _ok(q, b"fold-forsec", proof)
    with pytest.raises(o.OracleError):
        o.fold_params_for_security(3, 1, 1, 1, 12)


def test_forged_a_jump_rejected():
    # fold.rs:926-981: with out_a unbound, A could jump between segments.
    s = 8
    first = o.chained_traces(s, 1)[0]
    second = o.chained_traces(s, 1, 99, first[1][-1])[0]
    target = second[1][-1]
    assert target != _last_b(2 * s, 1)
    p = _params(target)
    acc = o.FoldAccumulator(p)
    t = o.Transcript(b"fold-forge-a")
    acc.add_segment(*first, t)
    acc.add_segment(*second, t)
    forged = acc.finalize_forged_out_a(0, 99, b"fold-forge-a")
    assert not _ok(p, b"fold-forge-a", forged)


def test_chain_segment_and_claims_rejected():
    # fold.rs:1137-1170: a broken chain (out_b claim) and a poisoned segment.
    traces = o.chained_traces(8, 3)
    p = _params(_last_b(8, 3))
    _, proof = _prove(p, traces, b"fold-chain")
    assert not _ok(p.replace(out_b=p.out_b + 1), b"fold-chain", proof)
    assert not _ok(p.replace(in_a=2), b"fold-chain", proof)
    bad = [traces[0], ([*traces[1][0][:4], traces[1][0][4] + 1, *traces[1][0][5:]], traces[1][1]), traces[2]]
    _, proof = _prove(p, bad, b"fold-badseg")
    assert not _ok(p, b"fold-badseg", proof)


def test_tampering_caps_and_transcripts_rejected():
    # fold.rs:1046-1128, 1172-1232
    traces = o.chained_traces(8, 2)
    p = _params(_last_b(8, 2))
    _, proof = _prove(p, traces, b"fold-tamper")
    assert _ok(p, b"fold-tamper", proof)
    assert not _ok(p, b"fold-B", proof)
    assert o.fold_verify(p, o.Transcript(b"fold-tamper"), proof[:-1]) == 8
    assert o.fold_verify(p, o.Transcript(b"fold-tamper"), proof + b"\0") == 8
    for off in range(5, len(proof), 211):
        bad = bytearray(proof)
        bad[off] ^= 2
        assert not _ok(p, b"fold-tamper", bytes(bad)), off
    capped = p.replace(cap_height=2)
    _, cproof = _prove(capped, traces, b"fold-cap")
    assert _ok(capped, b"fold-cap", cproof)
    assert not _ok(p, b"fold-cap", cproof)
