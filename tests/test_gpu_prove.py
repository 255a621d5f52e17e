"""End-to-end parity on the GPU: proofs and commitments must be byte-identical
to the oracle's for the same parameters and transcript, and accepted by both
the oracle verifier and the library's host verifier."""
import pytest

import oracle_ffi as o
from helpers import fri_bench_coeffs, seeded_claims

import z0p_amd as z

pytestmark = pytest.mark.gpu


def _air_case(n, a0, b0, label, ctx, shape=None):
    bl = o.fib_b_last(a0, b0, n)
    if shape is None:
        zp = z.air.for_security(n, a0, b0, bl)
        op = o.air_params_for_security(n, a0, b0, bl)
    else:
        blowup, nq, fs = shape
        zp = z.air.params_new(n, blowup, nq, fs, a0, b0, bl)
        op = o.air_params_new(n, blowup, nq, fs, a0, b0, bl)
    print("air prove n", n, "a0", a0, "b0", b0, zp.shape(), flush=True)
    got = z.air.prove(zp, z.Transcript(label), ctx)
    want = o.air_prove(op, o.Transcript(label))
    assert len(got) == len(want)
    assert got == want
    assert o.air_verify(op, o.Transcript(label), got) == 0
    z.air.verify(zp, got, z.Transcript(label))
    return got


@pytest.mark.parametrize("n,label", [(64, b"bench/air-prod"), (1024, b"bench/air-scale")])
def test_air_bench_rows_byte_identical(ctx, n, label):
    proof = _air_case(n, 1, 1, label, ctx)
    assert len(proof) == {64: 33061, 1024: 68573}[n]  # tools/baseline.json:83-97


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_air_seeded_1024(ctx, seed):
    a0, b0 = seeded_claims(seed)
    _air_case(1024, a0, b0, b"bench/air-scale", ctx)


@pytest.mark.parametrize("n,shape,label", [(64, (8, 12, 8), b"bench/air"), (1024, (8, 12, 8), b"bench/air1024"),
                                           (256, (4, 8, 4), b"air/b4")])
def test_air_small_shapes(ctx, n, shape, label):
    _air_case(n, 1, 1, label, ctx, shape)


@pytest.mark.parametrize("n", [8192, 65536])
def test_air_medium(ctx, n):
    proof = _air_case(n, 1, 1, b"bench/air-scale", ctx)
    assert len(proof) == {8192: 101233, 65536: 141641}[n]


def test_air_transcript_continues(ctx):
    """The transcript is advanced exactly like the reference's &mut Transcript."""
    n = 256
    bl = o.fib_b_last(3, 5, n)
    zt, ot = z.Transcript(b"pre"), o.Transcript(b"pre")
    zt.absorb(b"ctx", b"caller data")
    ot.absorb(b"ctx", b"caller data")
    got = z.air.prove(z.air.for_security(n, 3, 5, bl), zt, ctx)
    want = o.air_prove(o.air_params_for_security(n, 3, 5, bl), ot)
    assert got == want
    assert zt.challenge_bytes(b"after") == ot.challenge_bytes(b"after")


@pytest.mark.parametrize("n,label", [(256, b"bench/fri-prod"), (1024, b"bench/fri-prod-big")])
def test_fri_prod_byte_identical(ctx, n, label):
    coeffs = fri_bench_coeffs(n)
    zp, op = z.fri.for_security(n), o.fri_params_for_security(n)
    print("fri n", n, zp.shape(), flush=True)
    c, p = z.fri.commit_open(coeffs, zp, z.Transcript(label), ctx)
    oc, opf = o.fri_commit_open(coeffs, op, o.Transcript(label))
    assert c == oc and p == opf
    assert len(c) + len(p) == {256: 29378, 1024: 43174}[n]  # tools/baseline.json:78-92
    assert o.fri_verify(c, p, op, o.Transcript(label)) == 0
    z.fri.verify(c, p, zp, z.Transcript(label))


def test_fri_bench_shape(ctx):
    coeffs = fri_bench_coeffs(256)
    zp = z.fri.params_new(4, 12, 4)
    op = o.FriParams()
    op.blowup, op.num_queries, op.final_size = 4, 12, 4
    c, p = z.fri.commit_open(coeffs, zp, z.Transcript(b"bench/fri"), ctx)
    oc, opf = o.fri_commit_open(coeffs, op, o.Transcript(b"bench/fri"))
    assert c == oc and p == opf
    assert len(c) + len(p) == 11522  # tools/baseline.json:16-20


def _seeded_coeffs(n, seed):
    """c_i = Field::from_u64(splitmix64(&mut st)) (the reference's test idiom)."""
    from helpers import splitmix_fields
    return splitmix_fields(n, seed)


@pytest.mark.parametrize("logp", [12, 16])
def test_fri_seeded_byte_identical(ctx, logp):
    n = 1 << logp
    coeffs = _seeded_coeffs(n, logp)
    zp, op = z.fri.for_security(n), o.fri_params_for_security(n)
    print("fri seeded n", n, zp.shape(), flush=True)
    c, p = z.fri.commit_open(coeffs, zp, z.Transcript(b"bench/fri-prod-big"), ctx)
    oc, opf = o.fri_commit_open(coeffs, op, o.Transcript(b"bench/fri-prod-big"))
    assert c == oc and p == opf
    z.fri.verify(c, p, zp, z.Transcript(b"bench/fri-prod-big"))


def test_fri_2_22_config3(ctx):
    """Config 3: 2^22 coefficients, 2^26-point LDE, byte-identical to the oracle."""
    n = 1 << 22
    coeffs = _seeded_coeffs(n, 0)
    zp, op = z.fri.for_security(n), o.fri_params_for_security(n)
    assert zp.shape() == "B16Q40G20C4F16"
    print("fri config3 n", n, flush=True)
    c, p = z.fri.commit_open(coeffs, zp, z.Transcript(b"bench/fri-prod-big"), ctx)
    assert len(c) + len(p) == 181390  # size formula of fri.rs:963-1074
    oc, opf = o.fri_commit_open(coeffs, op, o.Transcript(b"bench/fri-prod-big"))
    assert c == oc and p == opf
    z.fri.verify(c, p, zp, z.Transcript(b"bench/fri-prod-big"))


def test_air_2_20_byte_identical(ctx):
    """The headline instance (a0 = b0 = 1, N = 2^20): 203,553 bytes, identical to the oracle."""
    proof = _air_case(1 << 20, 1, 1, b"bench/air-scale", ctx)
    assert len(proof) == 203553


@pytest.mark.parametrize("kind", ["air", "fri"])
def test_grind_beyond_first_window(ctx, kind):
    # 26-bit proof-of-work: the smallest nonce usually lies beyond the first
    # window of 2^24 nonces, so the openings enqueued ahead are cancelled and
    # the search continues window by window; bytes must not change.
    if kind == "air":
        n = 64
        bl = o.fib_b_last(2, 3, n)
        zp = z.air.params_new(n, 8, 6, 8, 2, 3, bl)
        op = o.air_params_new(n, 8, 6, 8, 2, 3, bl)
        zp.grinding_bits = op.grinding_bits = 26
        print("air grind 26", flush=True)
        got = z.air.prove(zp, z.Transcript(b"g26a"), ctx)  # smallest nonce 103,769,303 > 2^24
        assert got == o.air_prove(op, o.Transcript(b"g26a"))
    else:
        coeffs = fri_bench_coeffs(256)
        zp = z.fri.params_new(4, 6, 4, 26, 1)
        op = o.fri_params_new(4, 6, 4, 26, 1)
        print("fri grind 26", flush=True)
        got = z.fri.commit_open(coeffs, zp, z.Transcript(b"g26"), ctx)  # smallest nonce 167,873,451
        assert got == o.fri_commit_open(coeffs, op, o.Transcript(b"g26"))
