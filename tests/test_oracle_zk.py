"""Pins the oracle's zero-knowledge pieces: the ChaCha20 stream (src/rng.rs)
against RFC 8439's published vectors, the hiding mask (src/blind.rs) against
its defining identity evaluated at points, and prove_zk (src/air.rs:439-464)
against the reference's own acceptance, determinism and rejection tests
(air.rs:1365-1412)."""
import pytest

import oracle_ffi as o
from helpers import P


def test_chacha_quarter_round_rfc8439_2_1_1():
    # rng.rs:214-226 (RFC 8439 section 2.1.1)
    st = [0] * 16
    st[0:4] = [0x11111111, 0x01020304, 0x9B8D6F43, 0x01234567]
    out = o.chacha20_quarter(st, 0, 1, 2, 3)
    assert out[0:4] == [0xEA2A92F4, 0xCB1CF8CE, 0x4581472E, 0x5881C4BB]


def test_chacha_block_rfc8439_2_3_2():
    # rng.rs:228-247 (RFC 8439 section 2.3.2)
    key = bytes(range(32))
    nonce = bytes([0, 0, 0, 9, 0, 0, 0, 0x4A, 0, 0, 0, 0])
    want = bytes.fromhex(
        "10f1e7e4d13b5915500fdd1fa32071c4c7d1f4c733c068030422aa9ac3d46c4e"
        "d2826446079faa0914c2d705d98b02a2b5129cd1de164eb9cbd083e8a2503c4e")
    assert o.chacha20_block(key, 1, nonce) == want


def test_chacha_block_all_zero_key():
    # RFC 8439 appendix A.1, test vector 1: zero key, zero nonce, counter 0.
    want = bytes.fromhex(
        "76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7"
        "da41597c5157488d7724e03fb8d84a376a43b8f41518a11cc387b669b2ee6586")
    assert o.chacha20_block(bytes(32), 0, bytes(12)) == want


def test_chacha_stream_views():
    # rng.rs:249-286: the stream is consecutive blocks from counter 0; the
    # word views read the same bytes little-endian, unaligned across blocks.
    key, nonce = bytes([7] * 32), bytes([9] * 12)
    s = o.ChaCha20(key, nonce)
    assert s.fill_bytes(64) == o.chacha20_block(key, 0, nonce)
    assert s.fill_bytes(64) == o.chacha20_block(key, 1, nonce)
    a, b = o.ChaCha20(key, nonce), o.ChaCha20(key, nonce)
    stream = a.fill_bytes(200)
    got = b.fill_bytes(3) + b.next_u32().to_bytes(4, "little") + b.next_u64().to_bytes(8, "little")
    got += b.fill_bytes(61) + b.next_u32().to_bytes(4, "little")
    assert got == stream[:len(got)]


def test_chacha_next_field_rejects_the_top_interval():
    # rng.rs:157-168: words at or above 2p (= floor(2^32/p) p) are skipped,
    # the rest reduce mod p. Find a stream whose early words include a
    # rejected one, so that a missing rejection would shift the outputs.
    for seed in range(64):
        key = bytes([seed]) * 32
        words = [int.from_bytes(o.ChaCha20(key, bytes(12)).fill_bytes(64)[4 * i:4 * i + 4], "little")
                 for i in range(16)]
        if any(w >= 2 * P for w in words):
            break
    else:
        pytest.fail("no rejected word in the searched streams")
    want = [w % P for w in words if w < 2 * P]
    s = o.ChaCha20(key, bytes(12))
    assert [s.next_field() for _ in want] == want
    assert all(v < P for v in want)


def _eval(coeffs, z):
    acc = 0
    for c in reversed(coeffs):
        acc = (acc * z + c) % P
    return acc


@pytest.mark.parametrize("n,n_rand,clen", [(8, 4, 3), (16, 8, 16), (4, 13, 2), (16, 128, 16), (1024, 640, 1024)])
def test_mask_is_t_plus_vanishing_times_r(n, n_rand, clen):
    # blind.rs:5-9: M(z) = T(z) + (z^n - 1) R(z), R's coefficients being the
    # next n_rand field draws; checked at points, which fixes every index.
    key, nonce = bytes([42] * 32), bytes([7] * 12)
    coeffs = [(i * 7919 + 3) % P for i in range(clen)]
    masked = o.mask_coeffs(coeffs, n, n_rand, o.ChaCha20(key, nonce))
    total = n + n_rand
    assert len(masked) == 1 << (total - 1).bit_length()
    twin = o.ChaCha20(key, nonce)
    r = [twin.next_field() for _ in range(n_rand)]
    for z in [3, 123456, P - 2, 977 ** 3 % P]:
        want = (_eval(coeffs, z) + (pow(z, n, P) - 1) * _eval(r, z)) % P
        assert _eval(masked, z) == want
    # agrees with T on the base domain (blind.rs:100-115)
    w = o.two_adic_generator(n.bit_length() - 1)
    for k in range(min(n, 8)):
        x = pow(w, k, P)
        assert _eval(masked, x) == _eval(coeffs, x)


def test_mask_shapes_and_transparency():
    # blind.rs:90-98, 130-150
    rng = o.ChaCha20(bytes([42] * 32), bytes([7] * 12))
    for args in [([1] * 4, 3, 2), ([1] * 5, 4, 2), ([1], 1 << 27, 1)]:
        with pytest.raises(o.OracleError):
            o.mask_coeffs(args[0], args[1], args[2], rng)
    assert o.mask_coeffs([9, 2], 4, 0, rng) == [9, 2, 0, 0]
    assert len(o.mask_coeffs([1] * 16, 16, 8, rng)) == 32


def _zk_params(n=16):
    # air.rs:995-1001: n = 16, blowup 8, 8 queries, final 8.
    return o.air_params_new(n, 8, 8, 8, 1, 1, o.fib_b_last(1, 1, n))


def _zk_rng():
    return o.ChaCha20(bytes([11] * 32), bytes([22] * 12))  # air.rs:1365-1367


def test_zk_proof_verifies():
    p = _zk_params()
    proof = o.air_prove_zk(p, 128, _zk_rng(), o.Transcript(b"fib-zk"))
    assert o.air_verify(p, o.Transcript(b"fib-zk"), proof) == 0


def test_zk_deterministic_and_differs_from_plain():
    p = _zk_params()
    plain = o.air_prove(p, o.Transcript(b"fib-diverge"))
    z1 = o.air_prove_zk(p, 128, _zk_rng(), o.Transcript(b"fib-diverge"))
    z2 = o.air_prove_zk(p, 128, _zk_rng(), o.Transcript(b"fib-diverge"))
    assert z1 == z2
    assert z1 != plain
    assert len(z1) > len(plain)  # the masked LDE is larger: longer paths


def test_zk_rejects_few_maskers():
    p = _zk_params()
    with pytest.raises(o.OracleError) as e:
        o.air_prove_zk(p, 8 * 16 - 1, _zk_rng(), o.Transcript(b"fib-fewmask"))
    assert e.value.code == 8  # MalformedInput


def test_zk_tampered_proof_fails():
    p = _zk_params()
    proof = bytearray(o.air_prove_zk(p, 128, _zk_rng(), o.Transcript(b"fib-zktamper")))
    proof[-40] ^= 1  # inside the last opening's shifted-leaf bytes
    assert o.air_verify(p, o.Transcript(b"fib-zktamper"), bytes(proof)) != 0


def test_zk_production_shape_verifies():
    # n = 2^10 at blowup 16 with 640 maskers: the masked domain is 2^15, so
    # the final layer must be 32 (a power-of-four ratio).
    n = 1024
    p = o.air_params_new(n, 16, 40, 32, 5, 8, o.fib_b_last(5, 8, n))
    p.grinding_bits = 8
    p.cap_height = 4
    proof = o.air_prove_zk(p, 640, _zk_rng(), o.Transcript(b"zk/prod"))
    assert o.air_verify(p, o.Transcript(b"zk/prod"), proof) == 0
