"""Pins the oracle (oracle/, the CPU restatement of the reference prover)
against what the reference itself fixes: its known-answer vectors, its unit
identities, closed forms, brute force on tiny inputs, and the byte-exact proof
sizes of tools/baseline.json. Each check is chosen so a plausible restatement
mistake (wrong flag, index, sign, transposed operand) fails it."""
import struct

import numpy as np
import pytest

import oracle_ffi as o
from helpers import P, fri_bench_coeffs, splitmix64


def pattern(n):
    return bytes(i % 251 for i in range(n))


# hash.rs:733-775 (official BLAKE3 vectors)
KATS = [
    (b"", "af1349b9f5f9a1a6a0404dea36dcc9499bcb25c9adc112b7cc9a93cae41f3262"),
    (b"abc", "6437b3ac38465133ffb63b75273a8db548c558465d79db03fd359c6cd5bd9d85"),
    (pattern(1024), "42214739f095a406f3fc83deb889744ac00df831c10daa55189b5d121c855af7"),
    (pattern(1025), "d00278ae47eb27b34faecf67b4fe263f82d5412916c1ffd97c8cb7fb814b8444"),
    (pattern(2048), "e776b6028c7cd22a4d0ba182a8bf62205d2ef576467e838ed6f2529b85fba24a"),
    (pattern(3000), "5fade288bf27444bee55ba2babb98c3c922c1e84c2e445e7d1f6da24756f5060"),
]


@pytest.mark.parametrize("data,digest", KATS)
def test_blake3_kat(data, digest):
    assert o.blake3(data).hex() == digest


def test_blake3_small_and_incremental_match_oneshot():
    # hash.rs:777-823
    for n in [0, 1, 7, 63, 64, 65, 1000, 1023, 1024]:
        assert o.blake3_small(pattern(n)) == o.blake3(pattern(n))
    data = pattern(3000)
    assert o.blake3_split(data, [1, 7, 63, 64, 65, 959, 1024, 1025, 90]) == o.blake3(data)
    for n in [64, 1024, 2048]:
        assert o.blake3_split(pattern(n), [n // 2]) == o.blake3(pattern(n))


def test_leaf_and_node_digests_are_blake3_of_their_preimages():
    # merkle.rs:387-423: batch leaf/node forms equal the one-shot hash.
    rng = np.random.default_rng(1)
    for _ in range(20):
        block = rng.integers(0, 256, 64, dtype=np.uint8).tobytes()
        left = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
        right = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
        assert o.leaf64(block) == o.blake3(block)
        assert o.node(left, right) == o.blake3(b"\x01" + left + right)


def test_barrett_matches_remainder():
    # babybear.rs:353-375, including the state = splitmix64(&mut state) idiom.
    state = 0xBA771E5EED0000D1
    lib = o.lib()
    for _ in range(20000):
        _, state = splitmix64(state)
        a = state % P
        _, state = splitmix64(state)
        b = state % P
        assert lib.oracle_field_mul(a, b) == a * b % P
    corners = [0, 1, 2, P - 1000, P - 1]
    for a in corners:
        for b in corners:
            assert lib.oracle_field_mul(a, b) == a * b % P


def test_field_identities():
    lib = o.lib()
    assert lib.oracle_field_mul(2, 1006632961) == 1  # babybear.rs:379
    assert lib.oracle_field_inv(0) == 0
    assert o.field_pow(31, (P - 1) // 2) == P - 1  # generator is a non-residue
    for bits in [1, 10, 27]:
        w = o.two_adic_generator(bits)
        assert pow(w, 1 << bits, P) == 1
        assert pow(w, 1 << (bits - 1), P) != 1
        assert w == pow(31, (P - 1) >> bits, P)
    rng = np.random.default_rng(2)
    for v in rng.integers(1, P, 50):
        v = int(v)
        assert lib.oracle_field_inv(v) == pow(v, P - 2, P)
        assert (2 * lib.oracle_field_halve(v)) % P == v


def _e4(*c):
    return (o.ctypes.c_uint32 * 4)(*c)


def test_ext4_tower():
    # ext4.rs:313-367: u^2 = 11, v^2 = u, inverse = Fermat power.
    lib = o.lib()
    out = (o.ctypes.c_uint32 * 4)()
    lib.oracle_ext4_mul(_e4(0, 1, 0, 0), _e4(0, 1, 0, 0), out)
    assert list(out) == [11, 0, 0, 0]
    lib.oracle_ext4_mul(_e4(0, 0, 1, 0), _e4(0, 0, 1, 0), out)
    assert list(out) == [0, 1, 0, 0]
    e = P ** 4 - 2
    seed = 0x5EED1234ABCD9876
    for _ in range(32):
        c = []
        for _ in range(4):
            seed, z = splitmix64(seed)
            c.append(z % P)
        inv = (o.ctypes.c_uint32 * 4)()
        lib.oracle_ext4_inv(_e4(*c), inv)
        prod = (o.ctypes.c_uint32 * 4)()
        lib.oracle_ext4_mul(_e4(*c), inv, prod)
        assert list(prod) == [1, 0, 0, 0]
        fp = (o.ctypes.c_uint32 * 4)()
        lib.oracle_ext4_pow(_e4(*c), e >> 64, e & ((1 << 64) - 1), fp)
        assert list(fp) == list(inv)


def _horner(coeffs, x):
    acc = 0
    for c in reversed(coeffs):
        acc = (acc * x + int(c)) % P
    return acc


@pytest.mark.parametrize("n", [1, 2, 32, 256])
def test_ntt_against_textbook_dft(n):
    # ntt.rs:477-575: forward == Horner at every root power; inverse round-trips.
    rng = np.random.default_rng(n)
    coeffs = rng.integers(0, P, n, dtype=np.uint64).astype(np.uint32)
    evals = o.ntt_forward(coeffs)
    w = pow(31, (P - 1) // n, P) if n > 1 else 1
    for i in range(n):
        assert int(evals[i]) == _horner(coeffs, pow(w, i, P))
    np.testing.assert_array_equal(o.ntt_inverse(evals), coeffs)


@pytest.mark.parametrize("n,blowup", [(8, 4), (16, 16), (64, 2)])
def test_coset_lde_against_textbook(n, blowup):
    rng = np.random.default_rng(n + blowup)
    coeffs = rng.integers(0, P, n, dtype=np.uint64).astype(np.uint32)
    lde = o.evaluate_on_coset(coeffs, blowup)
    m = n * blowup
    om = pow(31, (P - 1) // m, P)
    for i in range(m):
        assert int(lde[i]) == _horner(coeffs, 31 * pow(om, i, P) % P)


def test_transcript_is_hash_of_framed_log():
    # transcript.rs:357-381: squeeze = H(framed history || u64 len || label || u64 counter).
    def frame(label, data):
        return struct.pack("<Q", len(label)) + label + struct.pack("<Q", len(data)) + data

    t = o.Transcript(b"proto")
    log = frame(b"domain", b"proto")
    for i, (label, data) in enumerate([(b"a", b"xyz"), (b"bb", b"\x00" * 1500), (b"c", b"")]):
        t.absorb(label, data)
        log += frame(label, data)
        ch = t.challenge_bytes(b"chal")
        want = o.blake3(log + struct.pack("<Q", 4) + b"chal" + struct.pack("<Q", i))
        assert ch == want
        log += frame(b"chal", want)
    # challenge_ext: four LE words of the squeeze mod p; index: low u64 masked.
    t2 = t.clone()
    d = t.challenge_bytes(b"e")
    assert t2.challenge_ext(b"e") == [x % P for x in struct.unpack("<4I", d[:16])]
    t3 = t.clone()
    d = t.challenge_bytes(b"q")
    assert t3.challenge_index(b"q", 1 << 20) == struct.unpack("<Q", d[:8])[0] & ((1 << 20) - 1)
    assert o.Transcript(b"x").challenge_index(b"q", 3) is None


def _lz(d):
    n = 0
    for b in d:
        if b == 0:
            n += 8
            continue
        n += 8 - b.bit_length()
        break
    return n


@pytest.mark.parametrize("bits", [0, 4, 9, 12])
def test_grind_returns_smallest_passing_nonce(bits):
    # transcript.rs:206-228, 495-504: minimality checked by brute force.
    t = o.Transcript(b"pow")
    t.absorb(b"x", bytes(range(40)))
    head = t.history_digest() + struct.pack("<Q", 7) + b"fri-pow"
    nonce = t.clone().grind(b"fri-pow", bits)
    assert nonce == o.grind_head(head, bits)
    assert _lz(o.blake3(head + struct.pack("<Q", nonce))) >= bits
    for k in range(nonce):
        assert _lz(o.blake3(head + struct.pack("<Q", k))) < bits
    assert t.clone().verify_grind(b"fri-pow", nonce, bits)


def test_fold_closed_forms():
    # fri.rs:1912-1939
    unit = o.two_adic_generator(2)
    assert unit * unit % P == P - 1
    x = 123456
    inv = o.field_inv(x)
    r = [7, 11, 13]
    ch = [[v, 0, 0, 0] for v in r]

    def quad(f):
        pts = [x, P - x, unit * x % P, (P - unit * x % P) % P]
        return [[f(t), 0, 0, 0] for t in pts]

    assert list(o.fold_values(quad(lambda t: (t + t * t % P * t) % P), inv, ch, unit)) == [(r[0] + r[2]) % P, 0, 0, 0]
    assert list(o.fold_values(quad(lambda t: t * t % P), inv, ch, unit)) == [r[1], 0, 0, 0]
    assert list(o.fold_values(quad(lambda t: 5), inv, ch, unit)) == [5, 0, 0, 0]


@pytest.mark.parametrize("lde_len,final", [(16, 1), (64, 4), (256, 4), (1024, 1)])
def test_point_at_matches_materialised_layers(lde_len, final):
    # fri.rs:1480-1505
    om = pow(31, (P - 1) // lde_len, P)
    bits = lde_len.bit_length() - 1
    from helpers import brev
    layer = [31 * pow(om, brev(i, bits), P) % P for i in range(lde_len)]
    depth = 0
    while True:
        for j, v in enumerate(layer):
            assert o.point_at(lde_len, depth, j) == v
        if len(layer) <= final:
            break
        layer = [pow(layer[4 * i], 4, P) for i in range(len(layer) // 4)]
        depth += 1


def test_linked_lde_equals_independent_ntt():
    # air.rs:1012-1022
    for n in [16, 256]:
        a, b = o.build_trace(3, 5, n)
        lde_a = o.evaluate_on_coset(o.ntt_inverse(a), 16)
        lde_b = o.evaluate_on_coset(o.ntt_inverse(b), 16)
        np.testing.assert_array_equal(o.linked_lde(lde_a, n, int(b[-1]), 3), lde_b)


def test_batched_compose_equals_scalar():
    # air.rs:1154-1242
    rng = np.random.default_rng(5)
    n, B = 64, 16
    a = rng.integers(0, P, n * B, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, P, n * B, dtype=np.uint64).astype(np.uint32)
    al = rng.integers(0, P, 20, dtype=np.uint64).astype(np.uint32)
    np.testing.assert_array_equal(o.compose(a, b, n, 1, 2, 3, al), o.compose(a, b, n, 1, 2, 3, al, scalar=True))


# tools/baseline.json byte-exact sizes (benches/speed_size.rs conventions).
AIR_SIZES = [(64, b"bench/air-prod", 33061), (1024, b"bench/air-scale", 68573), (8192, b"bench/air-scale", 101233),
             (65536, b"bench/air-scale", 141641)]


@pytest.mark.parametrize("n,label,size", AIR_SIZES)
def test_air_production_sizes_and_verify(n, label, size):
    bl = o.fib_b_last(1, 1, n)
    p = o.air_params_for_security(n, 1, 1, bl)
    proof = o.air_prove(p, o.Transcript(label))
    assert len(proof) == size
    assert o.air_verify(p, o.Transcript(label), proof) == 0


def test_air_2_20_size():
    # tools/baseline.json:108-112 — the headline row.
    n = 1 << 20
    bl = o.fib_b_last(1, 1, n)
    p = o.air_params_for_security(n, 1, 1, bl)
    assert p.shape() == "B16Q40G20C4F16"
    proof = o.air_prove(p, o.Transcript(b"bench/air-scale"))
    assert len(proof) == 203553
    assert o.air_verify(p, o.Transcript(b"bench/air-scale"), proof) == 0


def test_air_bench_shapes_sizes():
    for n, label, size in [(64, b"bench/air", 15465), (1024, b"bench/air1024", 28257)]:
        bl = o.fib_b_last(1, 1, n)
        p = o.air_params_new(n, 8, 12, 8, 1, 1, bl)
        proof = o.air_prove(p, o.Transcript(label))
        assert len(proof) == size
        assert o.air_verify(p, o.Transcript(label), proof) == 0


@pytest.mark.parametrize("n,label,size", [(256, b"bench/fri-prod", 29378), (1024, b"bench/fri-prod-big", 43174)])
def test_fri_production_sizes(n, label, size):
    p = o.fri_params_for_security(n)
    c, pf = o.fri_commit_open(fri_bench_coeffs(n), p, o.Transcript(label))
    assert len(c) + len(pf) == size
    assert o.fri_verify(c, pf, p, o.Transcript(label)) == 0


def test_fri_bench_shape_size():
    p = o.FriParams()
    p.blowup, p.num_queries, p.final_size = 4, 12, 4
    c, pf = o.fri_commit_open(fri_bench_coeffs(256), p, o.Transcript(b"bench/fri"))
    assert len(c) + len(pf) == 11522
    assert o.fri_verify(c, pf, p, o.Transcript(b"bench/fri")) == 0


def test_false_trace_is_rejected():
    # air.rs bad_trace regression: a trace breaking the transition fails FRI's final check.
    n = 64
    a, b = o.build_trace(1, 1, n)
    b = b.copy()
    b[10] = (int(b[10]) + 1) % P
    p = o.air_params_new(n, 8, 8, 8, 1, 1, int(b[-1]))
    proof = o.air_prove_with_trace(p, a, b, o.Transcript(b"bad"))
    assert o.air_verify(p, o.Transcript(b"bad"), proof) != 0


def test_thread_count_independence():
    # par.rs:70-79, ntt.rs:548-575, transcript.rs:495-504
    n = 1 << 14
    bl = o.fib_b_last(1, 1, n)
    p = o.air_params_for_security(n, 1, 1, bl)
    o.set_threads(1)
    one = o.air_prove(p, o.Transcript(b"t"))
    o.set_threads(8)
    many = o.air_prove(p, o.Transcript(b"t"))
    o.set_threads(0)
    assert one == many
