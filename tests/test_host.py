"""Host-side parts of the product library on the CPU: the C ABI exports,
the parameter tuners, the Fiat-Shamir transcript, the fast-doubling trace
claim and the verifier (which must accept the oracle's proofs and reject
every tampered one)."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_ffi as o
from helpers import P, fri_bench_coeffs, seeded_claims

import z0p_amd as z
from z0p_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ["z0p.h", "z0p_testing.h"]:
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(z0p_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", str(_ffi.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    declared = declared_symbols()
    assert declared, "no declarations parsed"
    assert declared <= exported, sorted(declared - exported)
    assert declared == set(_ffi.SIGNATURES), sorted(declared ^ set(_ffi.SIGNATURES))
    _ffi.lib()  # loads and binds every declared symbol


@pytest.mark.parametrize("logn", list(range(1, 24)))
def test_air_tuner_matches_oracle(logn):
    n = 1 << logn
    for bits in (40, 100, 128):
        a = z.air.for_security(n, 1, 2, 3, bits)
        b = o.air_params_for_security(n, 1, 2, 3, bits)
        assert bytes(a) == bytes(b), (n, bits)


@pytest.mark.parametrize("logn", list(range(0, 24)))
def test_fri_tuner_matches_oracle(logn):
    n = 1 << logn
    for bits in (30, 100):
        assert bytes(z.fri.for_security(n, bits)) == bytes(o.fri_params_for_security(n, bits))


def test_param_validation_codes():
    with pytest.raises(z.Z0pError) as e:
        z.air.for_security(3, 1, 1, 1)
    assert e.value.name == "InvalidDomainSize"
    with pytest.raises(z.Z0pError) as e:
        z.air.params_new(64, 2, 8, 8, 1, 1, 1)
    assert e.value.name == "InvalidBlowup"
    with pytest.raises(z.Z0pError) as e:
        z.air.params_new(64, 8, 0, 8, 1, 1, 1)
    assert e.value.name == "NoQueries"
    with pytest.raises(z.Z0pError) as e:
        z.air.for_security(1 << 24, 1, 1, 1)
    assert e.value.name == "InvalidDomainSize"
    with pytest.raises(z.Z0pError) as e:
        z.fri.for_security(0)
    assert e.value.name == "EmptyInput"


def test_b_last_fast_doubling():
    for n in [1, 2, 3, 64, 1000, 1 << 16]:
        for seed in range(3):
            a0, b0 = seeded_claims(seed)
            assert z.air.b_last(a0, b0, n) == o.fib_b_last(a0, b0, n)


def test_transcript_matches_oracle_on_random_sequences():
    rng = np.random.default_rng(11)
    for trial in range(5):
        label = rng.integers(0, 256, int(rng.integers(0, 40)), dtype=np.uint8).tobytes()
        zt, ot = z.Transcript(label), o.Transcript(label)
        for _ in range(40):
            if rng.random() < 0.6:
                lab = b"l%d" % int(rng.integers(0, 9))
                data = rng.integers(0, 256, int(rng.integers(0, 3000)), dtype=np.uint8).tobytes()
                zt.absorb(lab, data)
                ot.absorb(lab, data)
            else:
                assert zt.challenge_bytes(b"c") == ot.challenge_bytes(b"c")
        zc, oc = zt.clone(), ot.clone()
        assert zc.challenge_bytes(b"x") == oc.challenge_bytes(b"x")


@pytest.fixture(scope="module")
def oracle_air_proof():
    n, label = 1024, b"bench/air-scale"
    bl = o.fib_b_last(1, 1, n)
    return n, label, bl, o.air_prove(o.air_params_for_security(n, 1, 1, bl), o.Transcript(label))


def test_verifier_accepts_oracle_proofs(oracle_air_proof):
    n, label, bl, proof = oracle_air_proof
    z.air.verify(z.air.for_security(n, 1, 1, bl), proof, z.Transcript(label))
    for seed in range(2):
        a0, b0 = seeded_claims(seed)
        b = o.fib_b_last(a0, b0, 256)
        pr = o.air_prove(o.air_params_for_security(256, a0, b0, b), o.Transcript(b"s"))
        z.air.verify(z.air.for_security(256, a0, b0, b), pr, z.Transcript(b"s"))


def _expect_reject(params, proof, label, oracle_params):
    with pytest.raises(z.Z0pError):
        z.air.verify(params, proof, z.Transcript(label))
    assert o.air_verify(oracle_params, o.Transcript(label), proof) != 0


def test_verifier_rejects_tampering(oracle_air_proof):
    # tests/security.rs:60-354 style: byte flips across every proof region.
    n, label, bl, proof = oracle_air_proof
    zp = z.air.for_security(n, 1, 1, bl)
    op = o.air_params_for_security(n, 1, 1, bl)
    rng = np.random.default_rng(3)
    positions = sorted(set([1, 5, 40, 600, 1000, 2000, len(proof) // 2, len(proof) - 1, len(proof) - 130]
                           + list(rng.integers(1, len(proof), 24))))
    for pos in positions:
        bad = bytearray(proof)
        bad[int(pos)] ^= 1 << int(rng.integers(0, 8))
        _expect_reject(zp, bytes(bad), label, op)
    # wrong statement, wrong label, truncation, trailing bytes, wrong version
    _expect_reject(z.air.for_security(n, 1, 1, (bl + 1) % P), proof, label,
                   o.air_params_for_security(n, 1, 1, (bl + 1) % P))
    _expect_reject(zp, proof, b"other-label", op)
    _expect_reject(zp, proof[:-1], label, op)
    _expect_reject(zp, proof + b"\0", label, op)
    _expect_reject(zp, b"\x05" + proof[1:], label, op)


def test_fri_verifier_accepts_and_rejects():
    n, label = 256, b"bench/fri-prod"
    op = o.fri_params_for_security(n)
    c, pf = o.fri_commit_open(fri_bench_coeffs(n), op, o.Transcript(label))
    zp = z.fri.for_security(n)
    z.fri.verify(c, pf, zp, z.Transcript(label))
    for pos in [10, 30, len(pf) // 3, len(pf) - 3]:
        bad = bytearray(pf)
        bad[pos] ^= 4
        with pytest.raises(z.Z0pError):
            z.fri.verify(c, bytes(bad), zp, z.Transcript(label))
    for pos in [3, 40, len(c) - 1]:
        bad = bytearray(c)
        bad[pos] ^= 4
        with pytest.raises(z.Z0pError):
            z.fri.verify(bytes(bad), pf, zp, z.Transcript(label))


def test_no_gpu_context_fails_loudly():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(z.Z0pError) as e:
        z.Context(0)
    assert e.value.name == "NoDevice"


def test_chacha20_stream_matches_oracle():
    # The product's host ChaCha20 (rng.rs:82-168) against the pinned oracle:
    # RFC 8439 section 2.3.2 key/nonce, unaligned reads across blocks, and
    # the field sampler's rejection over a long run of draws.
    key, nonce = bytes(range(32)), bytes([0, 0, 0, 9, 0, 0, 0, 0x4A, 0, 0, 0, 0])
    a, b = z.ChaCha20(key, nonce), o.ChaCha20(key, nonce)
    assert a.fill_bytes(64) == o.chacha20_block(key, 0, nonce)
    assert a.fill_bytes(3) == b.fill_bytes(67)[64:]
    assert a.next_u32() == b.next_u32() and a.next_u64() == b.next_u64()
    assert [a.next_field() for _ in range(20000)] == [b.next_field() for _ in range(20000)]
    assert a.fill_bytes(200) == b.fill_bytes(200)
    with pytest.raises(ValueError):
        z.ChaCha20(bytes(31), bytes(12))
    r = z.ChaCha20.from_os()
    assert 0 <= r.next_field() < P


def test_verifier_accepts_oracle_zk_proofs():
    # air.rs:607-632: masked proofs carry a larger LDE in the commitment.
    n = 256
    bl = o.fib_b_last(3, 4, n)
    for n_rand, final in [(128, 16), (640, 32)]:
        op = o.air_params_new(n, 8, 8, final, 3, 4, bl)
        pr = o.air_prove_zk(op, n_rand, o.ChaCha20(bytes([1] * 32), bytes([2] * 12)), o.Transcript(b"zk"))
        zp = z.air.params_new(n, 8, 8, final, 3, 4, bl)
        z.air.verify(zp, pr, z.Transcript(b"zk"))
        bad = bytearray(pr)
        bad[len(bad) // 2] ^= 4
        _expect_reject(zp, bytes(bad), b"zk", op)


def test_verify_batch_matches_single_verifies():
    # Batched verification on host threads: each status equals that of its
    # own verify, and each transcript ends where its own call leaves it.
    items = []
    for i, n in enumerate([64, 256, 1024, 256]):
        a0, b0 = seeded_claims(40 + i)
        bl = o.fib_b_last(a0, b0, n)
        proof = o.air_prove(o.air_params_for_security(n, a0, b0, bl), o.Transcript(b"vb%d" % i))
        items.append((z.air.for_security(n, a0, b0, bl), proof, b"vb%d" % i))
    zk_p = o.air_params_new(256, 8, 8, 16, 3, 4, o.fib_b_last(3, 4, 256))
    zk = o.air_prove_zk(zk_p, 128, o.ChaCha20(bytes(32), bytes(12)), o.Transcript(b"vbzk"))
    items.append((z.air.params_new(256, 8, 8, 16, 3, 4, o.fib_b_last(3, 4, 256)), zk, b"vbzk"))
    bad = bytearray(items[2][1])
    bad[len(bad) // 3] ^= 1
    items.append((items[2][0], bytes(bad), items[2][2]))
    items.append((items[0][0], items[0][1][:-5], items[0][2]))  # truncated
    ts = [z.Transcript(lab) for _, _, lab in items]
    got = z.air.verify_batch([p for p, _, _ in items], [b for _, b, _ in items], ts, threads=3)
    for (p, b, lab), g, t in zip(items, got, ts):
        single = z.Transcript(lab)
        try:
            z.air.verify(p, b, single)
            want = None
        except z.Z0pError as e:
            want = e.code
        assert (g.code if g else None) == want
        assert t.challenge_bytes(b"after") == single.challenge_bytes(b"after")
    assert got[:5] == [None] * 5 and got[5] is not None and got[6] is not None
    shared = z.Transcript(b"s")
    with pytest.raises(z.Z0pError) as e:
        z.air.verify_batch([items[0][0]] * 2, [items[0][1]] * 2, [shared, shared])
    assert e.value.name == "InvalidArgument"
    assert z.air.verify_batch([], [], []) == []


def test_zero_copy_view_releases_buffer():
    # z0p_free releases what malloc returned, so a malloc'd buffer stands in
    # for one the library hands over: the view reads it without a copy, is
    # read-only, compares equal to bytes, and frees it once the last view of
    # it (slices included) is gone.
    import ctypes
    import gc
    import weakref
    from z0p_amd import _ffi
    libc = ctypes.CDLL(None)
    libc.malloc.restype = ctypes.c_void_p
    libc.malloc.argtypes = [ctypes.c_size_t]
    addr = libc.malloc(4096)
    ctypes.memmove(addr, bytes(range(256)) * 16, 4096)
    v = _ffi.view(ctypes.cast(addr, _ffi.u8p), 4096)
    assert v.readonly and len(v) == 4096 and v == bytes(range(256)) * 16
    with pytest.raises(TypeError):
        v[0] = 1
    part = v[256:512]
    assert bytes(part) == bytes(range(256))
    owner = v.obj
    alive = weakref.ref(owner)
    del v, owner
    gc.collect()
    assert alive() is not None  # the slice still holds the buffer
    del part
    gc.collect()
    assert alive() is None
    assert _ffi.view(ctypes.cast(libc.malloc(8), _ffi.u8p), 0) == b""


def test_fibonacci_interpolant_closed_form():
    # The device writes the Fibonacci columns' coefficients in closed form
    # (Binet: A[i] = c1 phi^i + c2 psi^i, phi, psi the roots of x^2 = x + 1,
    # which lie in BabyBear since p = 1 mod 5; each coefficient a geometric
    # sum). The identity must equal the oracle's INTT of build_trace
    # (air.rs:341-372) coefficient for coefficient, for A and for B.
    # sqrt(5) by Tonelli-Shanks (p - 1 = 2^27 * 15)
    q, s = P - 1, 0
    while q % 2 == 0:
        q, s = q // 2, s + 1
    z = 2
    while pow(z, (P - 1) // 2, P) != P - 1:
        z += 1
    m, c, t, r5 = s, pow(z, q, P), pow(5, q, P), pow(5, (q + 1) // 2, P)
    while t != 1:
        i, t2 = 0, t
        while t2 != 1:
            t2, i = t2 * t2 % P, i + 1
        b = pow(c, 1 << (m - i - 1), P)
        m, c, t, r5 = i, b * b % P, t * b * b % P, r5 * b % P
    assert r5 * r5 % P == 5
    inv = lambda x: pow(x, P - 2, P)
    phi, psi = (1 + r5) * inv(2) % P, (1 - r5) * inv(2) % P
    assert pow(phi, 1 << 27, P) != 1 and pow(psi, 1 << 27, P) != 1  # no vanishing denominator
    for n, (a0, b0) in [(2, (1, 1)), (8, seeded_claims(1)), (1024, seeded_claims(2))]:
        cols = o.build_trace(a0, b0, n)
        c1 = (a0 * inv(phi) + b0) * inv(r5) % P
        c2 = -(a0 * inv(psi) + b0) * inv(r5) % P
        winv = inv(o.two_adic_generator(n.bit_length() - 1))
        for col, (d1, d2) in zip(cols, [(c1, c2), (c1 * phi % P, c2 * psi % P)]):
            want = o.ntt_inverse([int(v) for v in col])
            k1 = d1 * (1 - pow(phi, n, P)) * inv(n) % P
            k2 = d2 * (1 - pow(psi, n, P)) * inv(n) % P
            got = [(k1 * inv((1 - phi * pow(winv, k, P)) % P) + k2 * inv((1 - psi * pow(winv, k, P)) % P)) % P
                   for k in range(n)]
            assert got == [int(v) for v in want]


@pytest.mark.parametrize("logn,cap_height", [(20, 4), (10, 4), (14, 0)])
def test_bench_tree_work_is_schedule_invariant(logn, cap_height):
    # The bench attributes each tree's hashing to the kernel family that runs
    # it under the schedule's fusions; the proof's compressions must not
    # move, and its bytes only by the 64-byte leaf inputs a fused kernel
    # hashes without reading them back.
    import bench
    n, m = 1 << logn, 1 << (logn + 4)
    layers, L = [], m
    while L > 16:
        layers.append(L)
        L //= 4
    base = dict(tail_max_log=17, tail_group_max_log=7, tail_spread_log=0, levels_fuse=2, small_max_log=0,
                grind_blocks=4096, fold_fuse_max_log=0, compose_fuse=0, trace_fuse=1, subtree_levels=0,
                compose_subtree=0, graph=0, lde_split=1, grind_pool=1024, fold_commit_max_log=0, tree_top=0,
                wave_streams=0)
    # unfused: every tree in the merkle family, leaf + 2 per node up to the cap
    cap = 1 << cap_height
    want = sum(k + 2 * (k - min(cap, k)) for k in [m // 8] + [x // 4 for x in layers])
    alg0, comp0 = bench.algorithmic_bytes(n, m, layers, cap_height, base)
    assert comp0 == {"merkle": want}
    for sched in [dict(base, compose_fuse=1), dict(base, small_max_log=10), dict(base, fold_fuse_max_log=18),
                  dict(base, compose_fuse=1, small_max_log=8, fold_fuse_max_log=22),
                  dict(base, fold_commit_max_log=20), dict(base, small_max_log=10, fold_commit_max_log=16,
                                                            fold_fuse_max_log=22)]:
        alg, comp = bench.algorithmic_bytes(n, m, layers, cap_height, sched)
        assert sum(comp.values()) == want
        fused_leaves = 0
        for d, L in enumerate(layers):
            if d == 0:
                fused_leaves += L // 4 if sched["compose_fuse"] else 0
            elif (sched["small_max_log"] and L <= 1 << sched["small_max_log"]) or \
                    (sched["fold_commit_max_log"] and (1 << 8) <= L <= 1 << sched["fold_commit_max_log"]) or \
                    (sched["fold_fuse_max_log"] and L <= 1 << sched["fold_fuse_max_log"]):
                fused_leaves += L // 4
        assert sum(alg.values()) - sum(alg0.values()) == -64 * fused_leaves, sched


def test_bench_dominant_kernel_bytes():
    # compose_leaves: A and B in, Ext4 layer 0 out, quad-leaf digests and
    # their parents out: 36 bytes per LDE point.
    import bench
    m = 1 << 24
    assert bench.kernel_alg_bytes(m, {"compose_fuse": 1}) == {"compose_leaves": 36 * m}
    assert bench.kernel_alg_bytes(m, {"compose_fuse": 0}) == {"compose": 24 * m}


@pytest.mark.parametrize("args", [(8, 16, 12_345, 12), (1 << 10, 8, 200, 100), (1 << 14, 24, 0xABCDEF, 100),
                                  (2, 1, 1, 100), (1 << 20, 16, 7, 128), (3, 16, 1, 12), (8, 0, 1, 12),
                                  (8, 25, 1, 12), (1 << 24, 8, 1, 100)])
def test_range_params_match_oracle(args):
    try:
        want = o.range_params_for_security(*args)
    except o.OracleError as e:
        with pytest.raises(z.Z0pError) as ze:
            z.range.for_security(*args)
        assert ze.value.code == e.code
        return
    got = z.range.for_security(*args)
    assert bytes(got) == bytes(want)


def test_range_params_new_codes_match_oracle():
    for args in [(8, 16, 8, 8, 4, 5), (3, 16, 8, 8, 4, 5), (8, 0, 8, 8, 4, 5), (8, 25, 8, 8, 4, 5),
                 (8, 16, 1, 8, 4, 5), (8, 16, 8, 0, 4, 5), (8, 16, 8, 8, 1, 5), (1 << 40, 16, 4, 8, 4, 5),
                 (8, 16, 2, 8, 4, P + 3)]:
        try:
            want = o.range_params_new(*args)
        except o.OracleError as e:
            with pytest.raises(z.Z0pError) as ze:
                z.range.params_new(*args)
            assert ze.value.code == e.code, args
            continue
        assert bytes(z.range.params_new(*args)) == bytes(want), args


def test_range_verifier_matches_oracle():
    # The host verifier accepts and rejects exactly the oracle's proofs of
    # the reference's range tests (range.rs:785-1022).
    def pair(n_rows, n_bits, blowup, nq, fs, value, **kw):
        op = o.range_params_new(n_rows, n_bits, blowup, nq, fs, value).replace(**kw)
        return op, z.range.params_new(n_rows, n_bits, blowup, nq, fs, value).replace(**kw)

    def agree(op, zp, label, proof):
        ok_o = o.range_verify(op, o.Transcript(label), proof)
        try:
            z.range.verify(zp, proof, z.Transcript(label))
            ok_z = 0
        except z.Z0pError as e:
            ok_z = e.code
        assert ok_z == ok_o, (label, ok_o, ok_z)
        return ok_o

    op, zp = pair(8, 16, 8, 8, 4, 12_345)
    pr = o.range_prove(op, o.Transcript(b"bench/range"))
    assert agree(op, zp, b"bench/range", pr) == 0
    assert agree(op, zp, b"bench/range", pr[:-1]) == 8
    assert agree(op, zp, b"other", pr) != 0
    for off in range(5, len(pr), 97):
        bad = bytearray(pr)
        bad[off] ^= 0x10
        agree(op, zp, b"bench/range", bytes(bad))
    cp, zc = pair(8, 16, 8, 8, 4, 12_345, cap_height=2, grinding_bits=5)
    prc = o.range_prove(cp, o.Transcript(b"cap"))
    assert agree(cp, zc, b"cap", prc) == 0
    assert agree(op, zp, b"cap", prc) != 0
    oop, ozp = pair(8, 16, 8, 8, 4, 70_000)
    assert agree(oop, ozp, b"oor", o.range_prove(oop, o.Transcript(b"oor"))) != 0
    assert agree(op.replace(value=1), zp.replace(value=1), b"bench/range", pr) != 0
    kp, kz = pair(8, 16, 8, 8, 8, 999)
    zk = o.range_prove_zk(kp, 8, o.ChaCha20(bytes([33] * 32), bytes([44] * 12)), o.Transcript(b"zk"))
    assert agree(kp, kz, b"zk", zk) == 0
    down = o.range_prove_coeffs(op, [[12_345] + [0] * 7] + [[(12_345 >> j) & 1] + [0] * 7 for j in range(16)], 8,
                                o.Transcript(b"down"))
    assert agree(op, zp, b"down", down) != 0


@pytest.mark.parametrize("args", [(8, 4, 30, 12), (1 << 10, 16, 5, 100), (1 << 14, 3, 7, 100), (2, 2, 1, 100),
                                  (3, 4, 1, 12), (8, 1, 1, 12), (8, (1 << 64) - 1, 1, 12), (1 << 20, 1 << 10, 9, 100)])
def test_table_params_match_oracle(args):
    try:
        want = o.table_params_for_security(*args)
    except o.OracleError as e:
        with pytest.raises(z.Z0pError) as ze:
            z.table.for_security(*args)
        assert ze.value.code == e.code
        return
    assert bytes(z.table.for_security(*args)) == bytes(want)


def test_table_bound_and_params_new_match_oracle():
    for t, n, r in [(4, 8, 0), (4, 8, 8), (1 << 63, 1 << 62, 0), (2, 2, 0), (17, 1 << 20, 300)]:
        assert z.table.membership_bound(t, n, r) == o.table_membership_bound(t, n, r)
    for args in [(8, 8, 8, 4, 30), (3, 8, 8, 4, 1), (8, 1, 8, 4, 1), (8, 8, 0, 4, 1), (8, 8, 8, 1, 1),
                 (1 << 40, 4, 8, 4, 1), (8, 2, 8, 4, P + 5)]:
        try:
            want = o.table_params_new(*args)
        except o.OracleError as e:
            with pytest.raises(z.Z0pError) as ze:
                z.table.params_new(*args)
            assert ze.value.code == e.code, args
            continue
        assert bytes(z.table.params_new(*args)) == bytes(want), args


def test_table_verifier_matches_oracle():
    T = [10, 20, 30, 40]

    def agree(op, zp, label, proof, table=T):
        ok_o = o.table_verify(op, table, o.Transcript(label), proof)
        try:
            z.table.verify(zp, table, proof, z.Transcript(label))
            ok_z = 0
        except z.Z0pError as e:
            ok_z = e.code
        assert ok_z == ok_o, (label, ok_o, ok_z)
        return ok_o

    op, zp = o.table_params_new(8, 8, 8, 4, 30), z.table.params_new(8, 8, 8, 4, 30)
    pr = o.table_prove(op, T, o.Transcript(b"bench/table"))
    assert agree(op, zp, b"bench/table", pr) == 0
    assert agree(op, zp, b"bench/table", pr[:-1]) == 8
    assert agree(op, zp, b"bench/table", pr, [10, 21, 30, 40]) != 0
    assert agree(op, zp, b"bench/table", pr, [10]) == 8
    for off in range(5, len(pr), 89):
        bad = bytearray(pr)
        bad[off] ^= 0x08
        agree(op, zp, b"bench/table", bytes(bad))
    cp, cz = op.replace(cap_height=2, grinding_bits=4), zp.replace(cap_height=2, grinding_bits=4)
    assert agree(cp, cz, b"cap", o.table_prove(cp, T, o.Transcript(b"cap"))) == 0
    np_, nz = o.table_params_new(8, 8, 8, 4, 99), z.table.params_new(8, 8, 8, 4, 99)
    assert agree(np_, nz, b"non", o.table_prove(np_, T, o.Transcript(b"non"))) != 0
    kp, kz = o.table_params_new(8, 8, 8, 8, 40), z.table.params_new(8, 8, 8, 8, 40)
    zk = o.table_prove_zk(kp, T, 8, o.ChaCha20(bytes([55] * 32), bytes([66] * 12)), o.Transcript(b"zk"))
    assert agree(kp, kz, b"zk", zk) == 0
    down = o.table_prove_coeffs(op, T, [30] + [0] * 7, 21, o.Transcript(b"down"))
    assert agree(op, zp, b"down", down) != 0


@pytest.mark.parametrize("args", [(8, 1, 1, 5, 12), (1 << 10, 1, 1, 7, 100), (1 << 16, 3, 4, 9, 100), (2, 1, 1, 1, 100),
                                  (3, 1, 1, 1, 12), (1 << 24, 1, 1, 1, 100)])
def test_fold_params_match_oracle(args):
    try:
        want = o.fold_params_for_security(*args)
    except o.OracleError as e:
        with pytest.raises(z.Z0pError) as ze:
            z.fold.for_security(*args)
        assert ze.value.code == e.code
        return
    assert bytes(z.fold.for_security(*args)) == bytes(want)
    for a2 in [(8, 8, 8, 4, 1, 1, 5), (3, 8, 8, 4, 1, 1, 1), (8, 2, 8, 4, 1, 1, 1), (8, 8, 0, 4, 1, 1, 1),
               (8, 8, 8, 1, 1, 1, 1), (1 << 40, 4, 8, 4, 1, 1, 1), (8, 4, 8, 4, P + 1, 2, P)]:
        try:
            w2 = o.fold_params_new(*a2)
        except o.OracleError as e:
            with pytest.raises(z.Z0pError) as ze:
                z.fold.params_new(*a2)
            assert ze.value.code == e.code, a2
            continue
        assert bytes(z.fold.params_new(*a2)) == bytes(w2), a2


def test_fold_verifier_matches_oracle():
    def prove(params, traces, label):
        t = o.Transcript(label)
        acc = o.FoldAccumulator(params)
        for ca, cb in traces:
            acc.add_segment(ca, cb, t)
        return acc, acc.finalize(t)

    def agree(op, zp, label, proof):
        ok_o = o.fold_verify(op, o.Transcript(label), proof)
        try:
            z.fold.verify(zp, proof, z.Transcript(label))
            ok_z = 0
        except z.Z0pError as e:
            ok_z = e.code
        assert ok_z == ok_o, (label, ok_o, ok_z)
        return ok_o

    traces = o.chained_traces(8, 4)
    out_b = traces[-1][1][-1]
    op, zp = o.fold_params_new(8, 8, 8, 4, 1, 1, out_b), z.fold.params_new(8, 8, 8, 4, 1, 1, out_b)
    acc, pr = prove(op, traces, b"bench/fold")
    assert len(pr) == 16845 and agree(op, zp, b"bench/fold", pr) == 0
    assert agree(op, zp, b"bench/fold", pr[:-1]) == 8
    assert agree(op, zp, b"fold-B", pr) != 0
    assert agree(op.replace(out_b=out_b + 1), zp.replace(out_b=out_b + 1), b"bench/fold", pr) != 0
    for off in range(5, len(pr), 173):
        bad = bytearray(pr)
        bad[off] ^= 0x20
        agree(op, zp, b"bench/fold", bytes(bad))
    first = o.chained_traces(8, 1)[0]
    second = o.chained_traces(8, 1, 99, first[1][-1])[0]
    fp, fz = o.fold_params_new(8, 8, 8, 4, 1, 1, second[1][-1]), z.fold.params_new(8, 8, 8, 4, 1, 1, second[1][-1])
    facc = o.FoldAccumulator(fp)
    t = o.Transcript(b"forge")
    facc.add_segment(*first, t)
    facc.add_segment(*second, t)
    assert agree(fp, fz, b"forge", facc.finalize_forged_out_a(0, 99, b"forge")) != 0
    cp, cz = op.replace(cap_height=2, grinding_bits=3), zp.replace(cap_height=2, grinding_bits=3)
    _, cpr = prove(cp, traces[:3], b"cap")
    cp3, cz3 = cp.replace(out_b=traces[2][1][-1]), cz.replace(out_b=traces[2][1][-1])
    agree(cp3, cz3, b"cap", cpr)
    _, cpr = prove(cp3, traces[:3], b"cap")
    assert agree(cp3, cz3, b"cap", cpr) == 0
