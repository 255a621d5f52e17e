"""Pins the oracle's table-membership AIR (src/table.rs) without running the
reference: the wire size of the reference bench's proof (tools/baseline.json
"table 4-entry": 4673 bytes, benches/speed_size.rs:301-329), the tuner's
(blowup, queries, grinding) for the reference's for_security case, the
membership bound formula, and the reference's acceptance and rejection tests
(table.rs:685-935). The verifier recomputes the composition at every query
from the opened value, so honest acceptance pins the batched composition
against composition_at; a missing factor of the vanishing product or a
wrong boundary sign shows in the non-member and wrong-claim rejections."""
import pytest

import oracle_ffi as o

TABLE = [10, 20, 30, 40]


def _plain(value):
    return o.table_params_new(8, 8, 8, 4, value)


def _zk(value):
    return o.table_params_new(8, 8, 8, 8, value)


def _rng():
    return o.ChaCha20(bytes([55] * 32), bytes([66] * 12))


def _ok(params, label, proof, table=TABLE):
    return o.table_verify(params, table, o.Transcript(label), proof) == 0


def test_bench_proof_size_matches_reference_baseline():
    p = _plain(30)
    proof = o.table_prove(p, TABLE, o.Transcript(b"bench/table"))
    assert len(proof) == 4673
    assert _ok(p, b"bench/table", proof)


def test_rejects_bad_params_and_tables():
    # table.rs:690-703
    for args, code in [((3, 8, 8, 4), 3), ((8, 1, 8, 4), 4), ((8, 8, 0, 4), 5), ((1 << 63, 4, 8, 4), 3)]:
        with pytest.raises(o.OracleError) as e:
            o.table_params_new(*args, 1)
        assert e.value.code == code, args
    p = _plain(20)
    for table in [[1], []]:
        with pytest.raises(o.OracleError) as e:
            o.table_prove(p, table, o.Transcript(b"table-shape"))
        assert e.value.code == 8
    with pytest.raises(o.OracleError) as e:
        o.table_prove(p.replace(n_rows=1), TABLE, o.Transcript(b"table-shape"))
    assert e.value.code == 3


def test_honest_members_verify_and_non_members_do_not():
    # table.rs:705-714, 822-830
    for value in TABLE:
        p = _plain(value)
        assert _ok(p, b"table-e2e", o.table_prove(p, TABLE, o.Transcript(b"table-e2e"))), value
    for value in [99, 0, 15, 41]:
        p = _plain(value)
        assert not _ok(p, b"table-nonmember", o.table_prove(p, TABLE, o.Transcript(b"table-nonmember"))), value


def test_for_security_and_bound():
    # table.rs:716-730, 862-872
    p = o.table_params_for_security(8, 4, 30, 12)
    assert p.blowup == 16 and (p.num_queries, p.grinding_bits) == (6, 6) and p.cap_height <= 4
    assert _ok(p, b"table-forsec", o.table_prove(p, TABLE, o.Transcript(b"table-forsec")))
    for args in [(3, 4), (8, 1), (8, (1 << 64) - 1)]:
        with pytest.raises(o.OracleError):
            o.table_params_for_security(*args, 1, 12)
    assert o.table_membership_bound(4, 8, 0) == 20 and o.table_membership_bound(4, 8, 8) == 52
    assert o.table_membership_bound(1 << 63, 1 << 62, 0) == (1 << 64) - 1 - (1 << 62)  # saturates, then - n


def test_wrong_table_claim_and_tampering_rejected():
    # table.rs:832-860
    p = _plain(20)
    proof = o.table_prove(p, TABLE, o.Transcript(b"table-wrong"))
    assert _ok(p, b"table-wrong", proof)
    assert not _ok(p, b"table-wrong", proof, [10, 21, 30, 40])
    assert not _ok(p.replace(value=30), b"table-wrong", proof)
    assert o.table_verify(p, TABLE, o.Transcript(b"table-wrong"), proof[:-1]) == 8
    assert o.table_verify(p, TABLE, o.Transcript(b"table-wrong"), proof + b"\0") == 8
    for off in [40, len(proof) // 2, len(proof) - 30]:
        bad = bytearray(proof)
        bad[off] ^= 4
        assert not _ok(p, b"table-wrong", bytes(bad)), off


def test_cap_mismatch_and_capped_roundtrip():
    # table.rs:745-803
    capped = _plain(30).replace(cap_height=2)
    proof = o.table_prove(capped, TABLE, o.Transcript(b"table-cap"))
    assert _ok(capped, b"table-cap", proof)
    assert int.from_bytes(proof[1:5], "little") == 4
    assert not _ok(_plain(30), b"table-cap", proof)
    assert not _ok(capped, b"table-cap", o.table_prove(_plain(30), TABLE, o.Transcript(b"table-cap")))


def test_downgraded_composition_bound_rejected():
    # table.rs:874-886
    p = _plain(20)
    coeffs = [20] + [0] * 7
    honest = o.table_prove_coeffs(p, TABLE, coeffs, 20, o.Transcript(b"table-downgrade"))
    assert honest == o.table_prove(p, TABLE, o.Transcript(b"table-downgrade"))
    assert not _ok(p, b"table-downgrade", o.table_prove_coeffs(p, TABLE, coeffs, 21, o.Transcript(b"table-downgrade")))


def test_zk_proofs():
    # table.rs:896-934
    pz = _zk(40)
    assert _ok(pz, b"table-zk", o.table_prove_zk(pz, TABLE, 8, _rng(), o.Transcript(b"table-zk")))
    plain = o.table_prove(_plain(10), TABLE, o.Transcript(b"table-diverge"))
    blinded = o.table_prove_zk(_zk(10), TABLE, 8, _rng(), o.Transcript(b"table-diverge"))
    assert blinded != plain
    assert blinded == o.table_prove_zk(_zk(10), TABLE, 8, _rng(), o.Transcript(b"table-diverge"))
    non = o.table_prove_zk(_zk(99), TABLE, 8, _rng(), o.Transcript(b"table-zknon"))
    assert not _ok(_zk(99), b"table-zknon", non)
    with pytest.raises(o.OracleError) as e:
        o.table_prove_zk(_zk(10), TABLE, 7, _rng(), o.Transcript(b"table-fewmask"))
    assert e.value.code == 8
