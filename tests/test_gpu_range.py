"""Range-check AIR (src/range.rs) on the GPU: every proof must be
byte-identical to the oracle's, advance the transcript as the oracle does,
and verify exactly when the oracle's does. Cases cover one- and two-block
row leaves (n_bits + 1 columns of 4 bytes behind a 0x00 prefix: one block up
to 15 columns), blowups 2..16, cap heights, grinding, explicit columns
(valid, out of range, non-boolean) and the masked prover."""
import numpy as np
import pytest

import oracle_ffi as o
from helpers import P

import z0p_amd as z

pytestmark = pytest.mark.gpu


def _pair(n_rows, n_bits, blowup, nq, fs, value, **kw):
    return (o.range_params_new(n_rows, n_bits, blowup, nq, fs, value).replace(**kw),
            z.range.params_new(n_rows, n_bits, blowup, nq, fs, value).replace(**kw))


def _verdict(op, zp, label, proof):
    ok_o = o.range_verify(op, o.Transcript(label), proof) == 0
    try:
        z.range.verify(zp, proof, z.Transcript(label))
        ok_z = True
    except z.Z0pError:
        ok_z = False
    assert ok_o == ok_z
    return ok_o


def test_bench_instance(ctx):
    # benches/speed_size.rs:267-290: the reference's 5217-byte proof.
    op, zp = _pair(8, 16, 8, 8, 4, 12_345)
    zt, ot = z.Transcript(b"bench/range"), o.Transcript(b"bench/range")
    got = z.range.prove(zp, zt, ctx)
    want = o.range_prove(op, ot)
    assert len(got) == 5217 and got == want
    assert zt.challenge_bytes(b"after") == ot.challenge_bytes(b"after")
    assert _verdict(op, zp, b"bench/range", got)


@pytest.mark.parametrize("n_rows,n_bits,value", [(8, 1, 1), (16, 14, 0x3ABC), (16, 15, 0x7FFF), (64, 24, 0xFFFFFF),
                                                 (1 << 10, 8, 0xA5), (1 << 12, 20, 0x12345)])
def test_for_security_byte_identical(ctx, n_rows, n_bits, value):
    op = o.range_params_for_security(n_rows, n_bits, value)
    zp = z.range.for_security(n_rows, n_bits, value)
    print("range for_security", n_rows, n_bits, zp.shape(), flush=True)
    got = z.range.prove(zp, z.Transcript(b"rs"), ctx)
    assert got == o.range_prove(op, o.Transcript(b"rs"))
    assert _verdict(op, zp, b"rs", got)


@pytest.mark.parametrize("blowup,fs,cap,grind", [(2, 4, 0, 0), (4, 2, 1, 3), (8, 16, 3, 0), (16, 8, 4, 6),
                                                 (2, 8, 8, 2)])
def test_shapes_byte_identical(ctx, blowup, fs, cap, grind):
    n_rows = 256
    m = n_rows * blowup
    if ((m // fs).bit_length() - 1) % 2:
        fs *= 2  # a power-of-four fold ratio
    op, zp = _pair(n_rows, 16, blowup, 9, fs, 40_000, cap_height=cap, grinding_bits=grind)
    print("range shape", zp.shape(), flush=True)
    got = z.range.prove(zp, z.Transcript(b"shape"), ctx)
    assert got == o.range_prove(op, o.Transcript(b"shape"))
    assert _verdict(op, zp, b"shape", got)


def test_large_instance(ctx):
    # 2^16 rows, 24 bits: 25 columns of 2^20 LDE values, two-block leaves.
    op = o.range_params_for_security(1 << 16, 24, 0xC0FFEE)
    zp = z.range.for_security(1 << 16, 24, 0xC0FFEE)
    print("range large", zp.shape(), flush=True)
    got = z.range.prove(zp, z.Transcript(b"big"), ctx)
    assert got == o.range_prove(op, o.Transcript(b"big"))
    assert _verdict(op, zp, b"big", got)


def test_explicit_columns(ctx):
    rng = np.random.default_rng(7)
    n, nb = 1 << 9, 12
    op, zp = _pair(n, nb, 8, 12, 16, 1234)
    vals = rng.integers(0, 1 << nb, n, dtype=np.uint32)
    vals[0] = 1234
    bits = [((vals >> j) & 1).astype(np.uint32) for j in range(nb)]
    cases = [("valid", vals, bits, True)]
    bad = vals.copy()
    bad[5] = 1 << nb  # row 5 out of range: its bits drop the top one
    cases.append(("oor", bad, bits, False))
    nonbit = [b.copy() for b in bits]
    nonbit[2][7] = 2
    cases.append(("nonbit", vals, nonbit, False))
    cases.append(("random", rng.integers(0, P, n, dtype=np.uint32), [rng.integers(0, P, n, dtype=np.uint32)
                                                                     for _ in range(nb)], False))
    for name, v, bs, honest in cases:
        print("range columns", name, flush=True)
        got = z.range.prove_with_columns(zp, v, bs, z.Transcript(b"cols"), ctx)
        want = o.range_prove_with_columns(op, [list(map(int, v))] + [list(map(int, b)) for b in bs],
                                          o.Transcript(b"cols"))
        assert got == want, name
        assert _verdict(op, zp, b"cols", got) == honest, name


def _rng():
    return bytes([33] * 32), bytes([44] * 12)


@pytest.mark.parametrize("n_rows,n_bits,blowup,nq,fs,n_rand,value", [
    (8, 16, 8, 8, 8, 8, 12_345),          # range.rs:990-1003
    (8, 16, 8, 8, 8, 8, 70_000),          # range.rs:1015-1021: still rejected
    (1 << 10, 16, 16, 20, 8, 40, 50_000),
    (1 << 12, 24, 4, 30, 32, 200, 0xFFFF),
])
def test_zk_byte_identical(ctx, n_rows, n_bits, blowup, nq, fs, n_rand, value):
    op, zp = _pair(n_rows, n_bits, blowup, nq, fs, value)
    print("range zk", n_rows, n_bits, blowup, n_rand, flush=True)
    key, nonce = _rng()
    orng, zrng = o.ChaCha20(key, nonce), z.ChaCha20(key, nonce)
    ot, zt = o.Transcript(b"zk"), z.Transcript(b"zk")
    got = z.range.prove_zk(zp, n_rand, zrng, zt, ctx)
    want = o.range_prove_zk(op, n_rand, orng, ot)
    assert got == want
    assert zrng.next_u64() == orng.next_u64()
    assert zt.challenge_bytes(b"after") == ot.challenge_bytes(b"after")
    assert _verdict(op, zp, b"zk", got) == (value < (1 << n_bits))


def test_error_codes_match_oracle(ctx):
    key, nonce = _rng()
    op, zp = _pair(8, 16, 8, 8, 4, 5)
    cases = [
        (lambda: o.range_prove(op.replace(cap_height=9), o.Transcript(b"e")),
         lambda: z.range.prove(zp.replace(cap_height=9), z.Transcript(b"e"), ctx)),
        (lambda: o.range_prove(op.replace(final_size=64), o.Transcript(b"e")),
         lambda: z.range.prove(zp.replace(final_size=64), z.Transcript(b"e"), ctx)),
        (lambda: o.range_prove(op.replace(final_size=8), o.Transcript(b"e")),
         lambda: z.range.prove(zp.replace(final_size=8), z.Transcript(b"e"), ctx)),
        (lambda: o.range_prove_with_columns(op, [[5] * 8] * 16, o.Transcript(b"e")),
         lambda: z.range.prove_with_columns(zp, [5] * 8, [[0] * 8] * 15, z.Transcript(b"e"), ctx)),
        (lambda: o.range_prove_with_columns(op, [[5] * 7] * 17, o.Transcript(b"e")),
         lambda: z.range.prove_with_columns(zp, [5] * 7, [[0] * 7] * 16, z.Transcript(b"e"), ctx)),
        (lambda: o.range_prove_zk(op, 7, o.ChaCha20(key, nonce), o.Transcript(b"e")),
         lambda: z.range.prove_zk(zp, 7, z.ChaCha20(key, nonce), z.Transcript(b"e"), ctx)),
    ]
    for k, (want, got) in enumerate(cases):
        with pytest.raises(o.OracleError) as oe:
            want()
        with pytest.raises(z.Z0pError) as ze:
            got()
        assert ze.value.code == oe.value.code, k
    # canonical columns only (Field values)
    with pytest.raises(z.Z0pError) as ze:
        z.range.prove_with_columns(zp, [P] * 8, [[0] * 8] * 16, z.Transcript(b"e"), ctx)
    assert ze.value.name == "MalformedInput"


@pytest.mark.parametrize("n_rows,n_bits,blowup,fs", [(2, 3, 8, 4), (4, 24, 2, 2)])
def test_smallest_domains(ctx, n_rows, n_bits, blowup, fs):
    # Two rows (the smallest trace) and a 2-blowup LDE of 8 points.
    op, zp = _pair(n_rows, n_bits, blowup, 4, fs, 5)
    print("range tiny", n_rows, n_bits, blowup, flush=True)
    got = z.range.prove(zp, z.Transcript(b"tiny"), ctx)
    assert got == o.range_prove(op, o.Transcript(b"tiny"))
    assert _verdict(op, zp, b"tiny", got)


def test_cap_covers_every_row(ctx):
    # cap height 8 over a 64-row LDE: the cap is the leaf level, paths are empty.
    op, zp = _pair(8, 16, 8, 8, 4, 12_345, cap_height=8)
    got = z.range.prove(zp, z.Transcript(b"c8"), ctx)
    assert got == o.range_prove(op, o.Transcript(b"c8"))
    assert _verdict(op, zp, b"c8", got)
