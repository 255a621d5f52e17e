"""The two-column AIR paths on the GPU: air::prove_zk (masked columns over a
larger LDE, air.rs:439-464) and air::prove_with_trace (explicit columns,
air.rs:385-406). Every proof must be byte-identical to the oracle's for the
same parameters, transcript and ChaCha20 stream, leave the stream and the
transcript where the oracle leaves them, and pass both verifiers."""
import numpy as np
import pytest

import oracle_ffi as o
from helpers import P, seeded_claims

import z0p_amd as z

pytestmark = pytest.mark.gpu

KEY, NONCE = bytes([11] * 32), bytes([22] * 12)  # air.rs:1365-1367

# (n, blowup, queries, final, grinding, cap, n_rand): the reference's ZK test
# shape, masks spanning many n-blocks (n_rand > n), a ragged n_rand, and every
# LDE kernel path (one 4096-point coset, four-step rows) at production rate.
ZK_CASES = [
    (16, 8, 8, 8, 0, 0, 128),
    (4, 4, 2, 4, 0, 0, 32),
    (64, 4, 4, 4, 2, 1, 100),
    (1024, 16, 40, 32, 8, 4, 640),
    (4096, 16, 40, 32, 12, 4, 640),
    (1 << 14, 16, 40, 32, 20, 4, 640),
    (1 << 16, 16, 40, 32, 20, 4, 5000),
]


def _pair(case, seed):
    n, blowup, nq, final, g, cap, _ = case
    a0, b0 = seeded_claims(seed)
    bl = o.fib_b_last(a0, b0, n)
    op = o.air_params_new(n, blowup, nq, final, a0, b0, bl)
    op.grinding_bits, op.cap_height = g, cap
    zp = z.air.params_new(n, blowup, nq, final, a0, b0, bl)
    zp.grinding_bits, zp.cap_height = g, cap
    return op, zp


@pytest.mark.parametrize("case", ZK_CASES, ids=lambda c: "n%d_B%d_r%d" % (c[0], c[1], c[6]))
def test_zk_byte_identical(ctx, case):
    n_rand = case[6]
    op, zp = _pair(case, case[0])
    print("zk case", case, flush=True)
    label = b"zk/%d" % case[0]
    ort, zrt = o.ChaCha20(KEY, NONCE), z.ChaCha20(KEY, NONCE)
    ot, zt = o.Transcript(label), z.Transcript(label)
    want = o.air_prove_zk(op, n_rand, ort, ot)
    got = z.air.prove_zk(zp, n_rand, zrt, zt, ctx)
    assert got == want
    assert zrt.next_u64() == ort.next_u64()
    assert zt.challenge_bytes(b"after") == ot.challenge_bytes(b"after")
    z.air.verify(zp, got, z.Transcript(label))
    assert o.air_verify(op, o.Transcript(label), got) == 0


def test_zk_repeat_and_fresh_stream(ctx):
    # Deterministic under a fixed stream; a stream that has moved on gives a
    # different, still valid proof (air.rs:1379-1394).
    case = ZK_CASES[3]
    op, zp = _pair(case, 1)
    r = z.ChaCha20(KEY, NONCE)
    first = z.air.prove_zk(zp, 640, r, z.Transcript(b"rep"), ctx)
    assert z.air.prove_zk(zp, 640, z.ChaCha20(KEY, NONCE), z.Transcript(b"rep"), ctx) == first
    second = z.air.prove_zk(zp, 640, r, z.Transcript(b"rep"), ctx)
    assert second != first
    z.air.verify(zp, second, z.Transcript(b"rep"))


def test_zk_errors_match_oracle(ctx):
    case = ZK_CASES[0]
    op, zp = _pair(case, 2)
    # too few maskers: MalformedInput before any draw (air.rs:447-456)
    zr, orr = z.ChaCha20(KEY, NONCE), o.ChaCha20(KEY, NONCE)
    with pytest.raises(z.Z0pError) as e:
        z.air.prove_zk(zp, 8 * 16 - 1, zr, z.Transcript(b"e"), ctx)
    assert e.value.name == "MalformedInput"
    assert zr.next_u64() == orr.next_u64()
    # a final size that leaves an odd fold ratio over the masked domain
    op2, zp2 = _pair((16, 8, 8, 16, 0, 0, 128), 2)
    with pytest.raises(o.OracleError) as oe:
        o.air_prove_zk(op2, 128, o.ChaCha20(KEY, NONCE), o.Transcript(b"e"))
    with pytest.raises(z.Z0pError) as ze:
        z.air.prove_zk(zp2, 128, z.ChaCha20(KEY, NONCE), z.Transcript(b"e"), ctx)
    assert ze.value.code == oe.value.code


@pytest.mark.parametrize("n,blowup", [(16, 4), (1024, 16), (4096, 8), (1 << 16, 16)])
def test_with_trace_equals_linked_prove(ctx, n, blowup):
    # A valid Fibonacci trace through the general path commits to the same
    # polynomials as the linked path (air.rs:1012-1022), so the bytes agree.
    a0, b0 = seeded_claims(n + 7)
    bl = o.fib_b_last(a0, b0, n)
    zp = z.air.params_new(n, blowup, 12, 4 if (n * blowup).bit_length() % 2 else 8, a0, b0, bl)
    col_a, col_b = o.build_trace(a0, b0, n)
    print("with_trace", n, blowup, flush=True)
    got = z.air.prove_with_trace(zp, col_a, col_b, z.Transcript(b"wt"), ctx)
    assert got == z.air.prove(zp, z.Transcript(b"wt"), ctx)


@pytest.mark.parametrize("n", [64, 4096, 1 << 15])
def test_with_trace_arbitrary_columns(ctx, n):
    # Columns that break the transition: still proved byte-identically, and
    # rejected by both verifiers (air.rs:374-384).
    rng = np.random.default_rng(n)
    col_a = rng.integers(0, P, n, dtype=np.uint32)
    col_b = rng.integers(0, P, n, dtype=np.uint32)
    op = o.air_params_new(n, 8, 16, 8 if (n * 8).bit_length() % 2 == 0 else 4, 1, 1, 5)
    zp = z.air.params_new(n, 8, 16, op.final_size, 1, 1, 5)
    want = o.air_prove_with_trace(op, [int(v) for v in col_a], [int(v) for v in col_b], o.Transcript(b"bad"))
    got = z.air.prove_with_trace(zp, col_a, col_b, z.Transcript(b"bad"), ctx)
    assert got == want
    with pytest.raises(z.Z0pError):
        z.air.verify(zp, got, z.Transcript(b"bad"))
    assert o.air_verify(op, o.Transcript(b"bad"), got) != 0


def test_with_trace_rejects_bad_columns(ctx):
    zp = z.air.params_new(16, 4, 4, 4, 1, 1, o.fib_b_last(1, 1, 16))
    a, b = o.build_trace(1, 1, 16)
    with pytest.raises(z.Z0pError) as e:
        z.air.prove_with_trace(zp, a[:15], b, z.Transcript(b"x"), ctx)
    assert e.value.name == "MalformedInput"
    with pytest.raises(z.Z0pError) as e:
        z.air.prove_with_trace(zp, a, [P] + list(b[1:]), z.Transcript(b"x"), ctx)
    assert e.value.name == "MalformedInput"
