"""Pins the oracle's range-check AIR (src/range.rs) without running the
reference: the wire size of the reference bench's proof (tools/baseline.json
"range 16-bit": 5217 bytes, benches/speed_size.rs:267-290), the tuner's
(queries, grinding) for the reference's own for_security case, and the
reference's acceptance and rejection tests (range.rs:785-1022). The verifier
recomputes the composition at each query from the opened row, so honest
acceptance pins the batched composition against composition_at; a dropped
term or a wrong sign shows in the rejection tests (an unbound value or bit
still verifies, a bound one does not)."""
import pytest

import oracle_ffi as o
from helpers import P


def _plain(value=12_345):
    return o.range_params_new(8, 16, 8, 8, 4, value)


def _zk_params(value):
    # range.rs:990-994: final size 8 for the 128-point masked LDE.
    return o.range_params_new(8, 16, 8, 8, 8, value)


def _rng():
    return o.ChaCha20(bytes([33] * 32), bytes([44] * 12))


def _ok(params, label, proof):
    return o.range_verify(params, o.Transcript(label), proof) == 0


def test_bench_proof_size_matches_reference_baseline():
    p = _plain()
    proof = o.range_prove(p, o.Transcript(b"bench/range"))
    assert len(proof) == 5217
    assert _ok(p, b"bench/range", proof)


def test_rejects_bad_params():
    # range.rs:791-799
    for args, code in [((3, 16, 8, 8, 4), 3), ((8, 0, 8, 8, 4), 8), ((8, 25, 8, 8, 4), 8), ((8, 16, 1, 8, 4), 4),
                       ((8, 16, 8, 0, 4), 5), ((1 << 63, 16, 4, 8, 4), 3), ((8, 16, 8, 8, 1), 3)]:
        with pytest.raises(o.OracleError) as e:
            o.range_params_new(*args, 1)
        assert e.value.code == code, args
    with pytest.raises(o.OracleError) as e:
        o.range_prove(_plain().replace(cap_height=9), o.Transcript(b"x"))
    assert e.value.code == 8


def test_for_security_roundtrip():
    # range.rs:812-826
    p = o.range_params_for_security(8, 16, 12_345, 12)
    assert p.blowup == 16 and (p.num_queries, p.grinding_bits) == (3, 6) and p.cap_height <= 4
    proof = o.range_prove(p, o.Transcript(b"range-forsec"))
    assert _ok(p, b"range-forsec", proof)
    for args in [(3, 16), (8, 0), (8, 25)]:
        with pytest.raises(o.OracleError):
            o.range_params_for_security(*args, 1, 12)


def test_honest_and_boundary_values_verify():
    # range.rs:802-810, 895-904
    for value in [12_345, 0, 1, 65_535]:
        p = _plain(value)
        proof = o.range_prove(p, o.Transcript(b"range-bound"))
        assert _ok(p, b"range-bound", proof), value


def test_out_of_range_and_wrong_claim_rejected():
    # range.rs:906-915, 953-962
    p = _plain(70_000)
    assert not _ok(p, b"range-oor", o.range_prove(p, o.Transcript(b"range-oor")))
    p = _plain(999)
    proof = o.range_prove(p, o.Transcript(b"range-claim"))
    assert _ok(p, b"range-claim", proof)
    assert not _ok(p.replace(value=1000), b"range-claim", proof)


def test_non_bit_column_rejected():
    # range.rs:917-928: a constant 2 in bit column 3 keeps the decomposition
    # of 5 only if booleanity is enforced.
    p = _plain(5)
    cols = [[5] * 8] + [[(5 >> j) & 1] * 8 for j in range(16)]
    honest = o.range_prove_with_columns(p, cols, o.Transcript(b"range-nonbit"))
    assert honest == o.range_prove(p, o.Transcript(b"range-nonbit"))
    cols[3] = [2] * 8
    assert not _ok(p, b"range-nonbit", o.range_prove_with_columns(p, cols, o.Transcript(b"range-nonbit")))
    with pytest.raises(o.OracleError) as e:
        o.range_prove_with_columns(p, cols[:-1], o.Transcript(b"x"))
    assert e.value.code == 8


def test_non_constant_but_valid_columns_verify():
    # The constraints hold row by row, not only for constants: a varying V
    # with matching bits and V[0] = v verifies; V[0] = 7 over bits that still
    # decode 6 breaks decomposition (and the boundary) at row 0.
    p = _plain(6)
    vals = [6, 1, 2, 3, 4, 5, 7, 0]
    cols = [vals] + [[(v >> j) & 1 for v in vals] for j in range(16)]
    assert _ok(p, b"rows", o.range_prove_with_columns(p, cols, o.Transcript(b"rows")))
    cols[0] = [7] + vals[1:]
    assert not _ok(p, b"rows", o.range_prove_with_columns(p, cols, o.Transcript(b"rows")))


def test_tampering_rejected():
    # range.rs:828-840, 940-951: truncated, extended and flipped bytes.
    p = _plain(999)
    proof = o.range_prove(p, o.Transcript(b"range-tamper"))
    assert _ok(p, b"range-tamper", proof)
    assert o.range_verify(p, o.Transcript(b"range-tamper"), proof[:-1]) == 8
    assert o.range_verify(p, o.Transcript(b"range-tamper"), proof + b"\0") == 8
    for off in [40, len(proof) // 2, len(proof) - 40]:
        bad = bytearray(proof)
        bad[off] ^= 1
        assert not _ok(p, b"range-tamper", bytes(bad)), off


def test_cap_mismatch_and_capped_roundtrip():
    # range.rs:842-857, 877-893
    capped = _plain().replace(cap_height=2)
    proof = o.range_prove(capped, o.Transcript(b"range-cap"))
    assert _ok(capped, b"range-cap", proof)
    assert proof[0] == 6 and int.from_bytes(proof[1:5], "little") == 4  # version, 4 cap digests
    assert not _ok(_plain(), b"range-cap", proof)
    plain = o.range_prove(_plain(), o.Transcript(b"range-cap"))
    assert not _ok(capped, b"range-cap", plain)
    bad = bytearray(proof)
    bad[5] ^= 1  # first cap digest
    assert not _ok(capped, b"range-cap", bytes(bad))


def test_downgraded_composition_bound_rejected():
    # range.rs:964-983: interpolated honest columns under bound n.
    p = _plain(999)
    cols = [[999] + [0] * 7] + [[(999 >> j) & 1] + [0] * 7 for j in range(16)]  # constants' coefficients
    honest = o.range_prove_coeffs(p, cols, 7, o.Transcript(b"range-downgrade"))
    assert honest == o.range_prove(p, o.Transcript(b"range-downgrade"))
    proof = o.range_prove_coeffs(p, cols, 8, o.Transcript(b"range-downgrade"))
    assert not _ok(p, b"range-downgrade", proof)


def test_zk_proofs():
    # range.rs:996-1021
    pz = _zk_params(12_345)
    proof = o.range_prove_zk(pz, 8, _rng(), o.Transcript(b"range-zk"))
    assert _ok(pz, b"range-zk", proof)
    plain = o.range_prove(_plain(999), o.Transcript(b"range-diverge"))
    blinded = o.range_prove_zk(_zk_params(999), 8, _rng(), o.Transcript(b"range-diverge"))
    assert blinded != plain
    assert blinded == o.range_prove_zk(_zk_params(999), 8, _rng(), o.Transcript(b"range-diverge"))
    oor = o.range_prove_zk(_zk_params(70_000), 8, _rng(), o.Transcript(b"range-zkoor"))
    assert not _ok(_zk_params(70_000), b"range-zkoor", oor)
    with pytest.raises(o.OracleError) as e:
        o.range_prove_zk(_zk_params(7), 7, _rng(), o.Transcript(b"range-fewmask"))
    assert e.value.code == 8


def test_composition_vanishes_exactly_on_honest_rows():
    # Each constraint is a separate term: honest rows give 0 at any point off
    # the trace domain; breaking one constraint gives that term alone.
    n, x = 8, 31 * 7
    al = [(k + 1, 2 * k + 3, 5, 7 * k) for k in range(18)]
    bits = [(12_345 >> j) & 1 for j in range(16)]
    assert o.range_composition_at(12_345, bits, x, 12_345, n, al) == (0, 0, 0, 0)
    # V off by one: decomposition and boundary fire; their sum is linear in
    # the alphas, so doubling the two alphas doubles the value.
    r1 = o.range_composition_at(12_346, bits, x, 12_345, n, al)
    al2 = [tuple(2 * c % P for c in al[0])] + al[1:17] + [tuple(2 * c % P for c in al[17])]
    r2 = o.range_composition_at(12_346, bits, x, 12_345, n, al2)
    assert r1 != (0, 0, 0, 0) and r2 == tuple(2 * c % P for c in r1)
    # a bit column of 2 with V still decomposing (bit 1 := 2 replaces bits
    # 1..0 = (0, 1) of 5 → 5 = 1 + 2*2): only booleanity of bit 1 fires, so
    # the value scales with alpha_bits[1] alone.
    b5 = [1, 2] + [0] * 14
    r3 = o.range_composition_at(5, b5, x, 5, n, al)
    al3 = list(al)
    al3[2] = (0, 0, 0, 0)
    assert r3 != (0, 0, 0, 0) and o.range_composition_at(5, b5, x, 5, n, al3) == (0, 0, 0, 0)
