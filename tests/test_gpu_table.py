"""Table-membership AIR (src/table.rs) on the GPU: every proof must be
byte-identical to the oracle's, advance the transcript as the oracle does,
and verify exactly when the oracle's does. Cases cover members and
non-members, tables of 2 to 257 entries, blowups, cap heights, grinding,
the masked prover and the error codes."""
import numpy as np
import pytest

import oracle_ffi as o
from helpers import P

import z0p_amd as z

pytestmark = pytest.mark.gpu

TABLE = [10, 20, 30, 40]


def _pair(n_rows, blowup, nq, fs, value, **kw):
    return (o.table_params_new(n_rows, blowup, nq, fs, value).replace(**kw),
            z.table.params_new(n_rows, blowup, nq, fs, value).replace(**kw))


def _verdict(op, zp, table, label, proof):
    ok_o = o.table_verify(op, table, o.Transcript(label), proof) == 0
    try:
        z.table.verify(zp, table, proof, z.Transcript(label))
        ok_z = True
    except z.Z0pError:
        ok_z = False
    assert ok_o == ok_z
    return ok_o


def test_bench_instance(ctx):
    # benches/speed_size.rs:301-329: the reference's 4673-byte proof.
    op, zp = _pair(8, 8, 8, 4, 30)
    zt, ot = z.Transcript(b"bench/table"), o.Transcript(b"bench/table")
    got = z.table.prove(zp, TABLE, zt, ctx)
    want = o.table_prove(op, TABLE, ot)
    assert len(got) == 4673 and got == want
    assert zt.challenge_bytes(b"after") == ot.challenge_bytes(b"after")
    assert _verdict(op, zp, TABLE, b"bench/table", got)


@pytest.mark.parametrize("value", [10, 20, 30, 40, 99, 0])
def test_members_and_non_members(ctx, value):
    op, zp = _pair(8, 8, 8, 4, value)
    got = z.table.prove(zp, TABLE, z.Transcript(b"table-e2e"), ctx)
    assert got == o.table_prove(op, TABLE, o.Transcript(b"table-e2e"))
    assert _verdict(op, zp, TABLE, b"table-e2e", got) == (value in TABLE)


@pytest.mark.parametrize("n_rows,table_len,member", [(8, 2, True), (1 << 10, 16, True), (1 << 12, 5, False),
                                                     (1 << 14, 3, True), (64, 257, True), (1 << 16, 4, True)])
def test_for_security_byte_identical(ctx, n_rows, table_len, member):
    rng = np.random.default_rng(n_rows + table_len)
    table = [int(v) for v in rng.integers(0, P, table_len)]
    value = table[table_len // 2] if member else (table[0] + 1) % P
    op = o.table_params_for_security(n_rows, table_len, value)
    zp = z.table.for_security(n_rows, table_len, value)
    print("table for_security", n_rows, table_len, zp.shape(), flush=True)
    got = z.table.prove(zp, table, z.Transcript(b"ts"), ctx)
    assert got == o.table_prove(op, table, o.Transcript(b"ts"))
    assert _verdict(op, zp, table, b"ts", got) == (member or value in table)


@pytest.mark.parametrize("blowup,fs,cap,grind", [(8, 2, 1, 3), (16, 16, 3, 0), (32, 8, 8, 5)])
def test_shapes_byte_identical(ctx, blowup, fs, cap, grind):
    n_rows = 128
    m = n_rows * blowup
    if ((m // fs).bit_length() - 1) % 2:
        fs *= 2
    op, zp = _pair(n_rows, blowup, 9, fs, 20, cap_height=cap, grinding_bits=grind)
    print("table shape", zp.shape(), flush=True)
    got = z.table.prove(zp, TABLE, z.Transcript(b"shape"), ctx)
    assert got == o.table_prove(op, TABLE, o.Transcript(b"shape"))
    assert _verdict(op, zp, TABLE, b"shape", got)


@pytest.mark.parametrize("n_rows,blowup,nq,fs,n_rand,value", [
    (8, 8, 8, 8, 8, 40),          # table.rs:896-902
    (8, 8, 8, 8, 8, 99),          # table.rs:925-931: still rejected
    (1 << 10, 16, 20, 8, 40, 30),
])
def test_zk_byte_identical(ctx, n_rows, blowup, nq, fs, n_rand, value):
    op, zp = _pair(n_rows, blowup, nq, fs, value)
    print("table zk", n_rows, blowup, n_rand, flush=True)
    key, nonce = bytes([55] * 32), bytes([66] * 12)
    orng, zrng = o.ChaCha20(key, nonce), z.ChaCha20(key, nonce)
    ot, zt = o.Transcript(b"zk"), z.Transcript(b"zk")
    got = z.table.prove_zk(zp, TABLE, n_rand, zrng, zt, ctx)
    want = o.table_prove_zk(op, TABLE, n_rand, orng, ot)
    assert got == want
    assert zrng.next_u64() == orng.next_u64()
    assert zt.challenge_bytes(b"after") == ot.challenge_bytes(b"after")
    assert _verdict(op, zp, TABLE, b"zk", got) == (value in TABLE)


def test_error_codes_match_oracle(ctx):
    key, nonce = bytes([55] * 32), bytes([66] * 12)
    op, zp = _pair(8, 8, 8, 4, 30)
    cases = [
        (lambda: o.table_prove(op, [1], o.Transcript(b"e")), lambda: z.table.prove(zp, [1], z.Transcript(b"e"), ctx)),
        (lambda: o.table_prove(op, [], o.Transcript(b"e")), lambda: z.table.prove(zp, [], z.Transcript(b"e"), ctx)),
        (lambda: o.table_prove(op.replace(cap_height=9), TABLE, o.Transcript(b"e")),
         lambda: z.table.prove(zp.replace(cap_height=9), TABLE, z.Transcript(b"e"), ctx)),
        (lambda: o.table_prove(op.replace(n_rows=1), TABLE, o.Transcript(b"e")),
         lambda: z.table.prove(zp.replace(n_rows=1), TABLE, z.Transcript(b"e"), ctx)),
        # |T| = 40 over an LDE of 64: the bound exceeds the domain (rate >= 1)
        (lambda: o.table_prove(op, list(range(40)), o.Transcript(b"e")),
         lambda: z.table.prove(zp, list(range(40)), z.Transcript(b"e"), ctx)),
        (lambda: o.table_prove_zk(op, TABLE, 7, o.ChaCha20(key, nonce), o.Transcript(b"e")),
         lambda: z.table.prove_zk(zp, TABLE, 7, z.ChaCha20(key, nonce), z.Transcript(b"e"), ctx)),
    ]
    for k, (want, got) in enumerate(cases):
        with pytest.raises(o.OracleError) as oe:
            want()
        with pytest.raises(z.Z0pError) as ze:
            got()
        assert ze.value.code == oe.value.code, k


def test_smallest_domain(ctx):
    # Two rows, a two-entry table, a 16-point LDE.
    op, zp = _pair(2, 8, 4, 4, 7)
    got = z.table.prove(zp, [7, 9], z.Transcript(b"tiny"), ctx)
    assert got == o.table_prove(op, [7, 9], o.Transcript(b"tiny"))
    assert _verdict(op, zp, [7, 9], b"tiny", got)


def test_cap_covers_every_row(ctx):
    op, zp = _pair(8, 8, 8, 4, 30, cap_height=8)
    got = z.table.prove(zp, TABLE, z.Transcript(b"c8"), ctx)
    assert got == o.table_prove(op, TABLE, o.Transcript(b"c8"))
    assert _verdict(op, zp, TABLE, b"c8", got)
