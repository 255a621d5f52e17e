"""The reference's adversarial suites (tests/security.rs, tests/redteam.rs,
tests/hardening.rs) against the drop-in verifier: honest proofs come from
the oracle (byte-identical to the GPU prover's), every forgery is built by
tampering one decoded field (tests/wire.py) or one byte, and each must be
rejected by the product's z0p_*_verify over the C ABI, with no crash. Each
test prints its vector with attempt and rejection counts, as the reference
suites do.

Run with Z0P_HOST_LIBRARY pointing at a sanitizer build of the host half
(`make -C z0p_amd/csrc sanitize`) these suites drive the verifier under
ASan/UBSan; tests/test_sanitizers.py does that in a subprocess.
"""
import os

import pytest

import oracle_ffi as o
import wire
from helpers import P, splitmix64

import z0p_amd as z


def report(vector, attempts, rejected):
    print("adversarial %-28s attempts=%-5d rejected=%-5d ACCEPTED=%d" % (vector, attempts, rejected,
                                                                         attempts - rejected))


def splitmix(state):
    """tests/security.rs splitmix: returns (new_state, output)."""
    return splitmix64(state)


def b_last(a0, b0, n):
    return o.fib_b_last(a0, b0, n)


def air_pair(n, blowup, nq, final_size, a0, b0, bl, grinding=0, cap=0):
    """(product params, oracle params) for the same AIR statement."""
    zp = z.air.params_new(n, blowup, nq, final_size, a0, b0, bl)
    zp.grinding_bits, zp.cap_height = grinding, cap
    op = o.air_params_new(n, blowup, nq, final_size, a0, b0, bl)
    op.grinding_bits, op.cap_height = grinding, cap
    return zp, op


def air_accepts(zp, proof, label):
    try:
        z.air.verify(zp, proof, z.Transcript(label))
        return True
    except z.Z0pError:
        return False


# ---- tests/security.rs ---------------------------------------------------
SEC_LABEL = b"security/air"


@pytest.fixture(scope="module")
def security():
    """security.rs:28-44: N = 32, B8 Q8 F16, 8 grinding bits, cap height 2."""
    zp, op = air_pair(32, 8, 8, 16, 1, 1, b_last(1, 1, 32), grinding=8, cap=2)
    proof = o.air_prove(op, o.Transcript(SEC_LABEL))
    return zp, op, proof


def test_honest_proof_accepts(security):
    zp, _, proof = security
    assert air_accepts(zp, proof, SEC_LABEL)


def test_tampered_fri_values_rejected(security):
    zp, _, proof = security
    base = wire.AirProof.from_bytes(proof)
    attempts = rejected = 0
    for q in range(len(base.fri_proof.queries)):
        for f in range(len(base.fri_proof.queries[q])):
            for v in range(4):
                bad = wire.AirProof.from_bytes(proof)
                fw = bad.fri_proof.queries[q][f]
                fw.values[v] = wire.ext_add_one(fw.values[v])
                attempts += 1
                rejected += not air_accepts(zp, bad.to_bytes(), SEC_LABEL)
    report("tamper-fri-values", attempts, rejected)
    assert attempts > 0 and rejected == attempts


def test_tampered_final_poly_rejected(security):
    zp, _, proof = security
    n = len(wire.AirProof.from_bytes(proof).fri_proof.final_poly)
    attempts = rejected = 0
    for i in range(n):
        bad = wire.AirProof.from_bytes(proof)
        bad.fri_proof.final_poly[i] = wire.ext_add_one(bad.fri_proof.final_poly[i])
        attempts += 1
        rejected += not air_accepts(zp, bad.to_bytes(), SEC_LABEL)
    report("tamper-final-poly", attempts, rejected)
    assert attempts > 0 and rejected == attempts


def test_tampered_trace_values_rejected(security):
    zp, _, proof = security
    n = len(wire.AirProof.from_bytes(proof).openings)
    attempts = rejected = 0
    for i in range(n):
        for which in ("a", "b", "a_next", "b_next"):
            bad = wire.AirProof.from_bytes(proof)
            setattr(bad.openings[i], which, (getattr(bad.openings[i], which) + 1) % P)
            attempts += 1
            rejected += not air_accepts(zp, bad.to_bytes(), SEC_LABEL)
    report("tamper-trace-values", attempts, rejected)
    assert attempts > 0 and rejected == attempts


def test_tampered_paths_rejected(security):
    zp, _, proof = security
    forgeries = []
    bad = wire.AirProof.from_bytes(proof)
    bad.fri_proof.queries[0][0].path.pop()
    forgeries.append(bad)
    bad = wire.AirProof.from_bytes(proof)
    bad.openings[0].path.pop()
    forgeries.append(bad)
    bad = wire.AirProof.from_bytes(proof)
    d, left = bad.fri_proof.queries[0][0].path[0]
    bad.fri_proof.queries[0][0].path[0] = (bytes([d[0] ^ 1]) + d[1:], left)
    forgeries.append(bad)
    bad = wire.AirProof.from_bytes(proof)
    d, left = bad.openings[0].path[0]
    bad.openings[0].path[0] = (bytes([d[0] ^ 1]) + d[1:], left)
    forgeries.append(bad)
    bad = wire.AirProof.from_bytes(proof)
    d, left = bad.openings[0].path[0]
    bad.openings[0].path[0] = (d, not left)
    forgeries.append(bad)
    rejected = sum(not air_accepts(zp, f.to_bytes(), SEC_LABEL) for f in forgeries)
    report("tamper-merkle-paths", len(forgeries), rejected)
    assert rejected == len(forgeries)


def test_tampered_commitments_rejected(security):
    zp, _, proof = security
    bad1 = wire.AirProof.from_bytes(proof)
    c = bad1.fri_commitment.caps[0][0]
    bad1.fri_commitment.caps[0][0] = bytes([c[0] ^ 1]) + c[1:]
    bad2 = wire.AirProof.from_bytes(proof)
    c = bad2.column_cap[0]
    bad2.column_cap[0] = bytes([c[0] ^ 1]) + c[1:]
    rejected = sum(not air_accepts(zp, b.to_bytes(), SEC_LABEL) for b in (bad1, bad2))
    report("tamper-commitments", 2, rejected)
    assert rejected == 2


def test_wrong_domain_and_statement_rejected(security):
    zp, _, proof = security
    rejected = not air_accepts(zp, proof, b"security/WRONG")
    other, _ = air_pair(32, 8, 8, 16, 1, 1, (b_last(1, 1, 32) + 1) % P, grinding=8, cap=2)
    rejected += not air_accepts(other, proof, SEC_LABEL)
    report("wrong-domain-statement", 2, rejected)
    assert rejected == 2


def test_forged_pow_nonce_rejected(security):
    zp, _, proof = security
    found = 0
    for delta in range(1, 6):
        bad = wire.AirProof.from_bytes(proof)
        bad.fri_proof.pow_nonce = (bad.fri_proof.pow_nonce + delta) & ((1 << 64) - 1)
        found += not air_accepts(zp, bad.to_bytes(), SEC_LABEL)
    report("forged-pow-nonce", 5, found)
    assert found >= 1
    t = z.Transcript(b"security/pow")
    t.absorb(b"stmt", b"hello")
    assert not t.verify_grind(b"pow", 0, 257)


def _fri_coeffs16(mult, add):
    return [((i * mult) + add) & 0xFFFFFFFF for i in range(16)]


def fri_accepts(cm, pf, zp, label):
    try:
        z.fri.verify(cm, pf, zp, z.Transcript(label))
        return True
    except z.Z0pError:
        return False


def test_rate_gate_forgeries_rejected():
    # security.rs:226-262
    zp, op = z.fri.params_new(4, 8, 4), o.fri_params_new(4, 8, 4)
    coeffs = [c % P for c in _fri_coeffs16(0x9E3779B9, 17)]
    prover = o.Transcript(b"security/rate")
    st = o.FriState(o.FRI_KIND_BASE, coeffs, op, prover)
    cm = st.commitment()
    pf, _ = st.open(prover)
    assert fri_accepts(cm, pf, zp, b"security/rate")
    high = wire.Commitment.from_bytes(cm)
    high.degree_bound = high.lde_len - 1
    raised = not fri_accepts(high.to_bytes(), pf, zp, b"security/rate")

    zl, ol = z.fri.params_new(4, 8, 16), o.fri_params_new(4, 8, 16)
    seed, evals = 0x2B2B2B2B2B2B2B2B, []
    for _ in range(64):
        _, seed = splitmix(seed)
        evals.append(seed % P)
    prover = o.Transcript(b"security/lowbound")
    st = o.FriState(o.FRI_KIND_EVALS, evals, ol, prover, degree_bound=7)
    cm = st.commitment()
    pf, _ = st.open(prover)
    lowered = not fri_accepts(cm, pf, zl, b"security/lowbound")
    report("rate-gate-forgeries", 2, raised + lowered)
    assert raised and lowered


def test_byte_fuzz_proof_rejected(security):
    zp, _, proof = security
    trials, state, rejected, accepted_at = 200, 0xF0221E57CA71E500, 0, []
    for _ in range(trials):
        state, r = splitmix(state)
        pos = r % len(proof)
        state, r = splitmix(state)
        bad = bytearray(proof)
        bad[pos] ^= 1 << (r % 8)
        if air_accepts(zp, bytes(bad), SEC_LABEL):
            accepted_at.append(pos)
        else:
            rejected += 1
    report("byte-fuzz-proof", trials, rejected)
    assert rejected == trials, accepted_at


def test_byte_fuzz_commitment_rejected(security):
    # security.rs:295-330: a decodable mutated commitment replaces the
    # proof's own, which the verifier must then reject.
    zp, _, proof = security
    base = wire.AirProof.from_bytes(proof)
    cbytes = base.fri_commitment.to_bytes()
    trials, state, rejected, accepted_at = 100, 0xC00117F055000000, 0, []
    for _ in range(trials):
        state, r = splitmix(state)
        pos = r % len(cbytes)
        state, r = splitmix(state)
        mutated = bytearray(cbytes)
        mutated[pos] ^= 1 << (r % 8)
        try:
            decoded = wire.Commitment.from_bytes(bytes(mutated))
        except wire.WireError:
            rejected += 1
            continue
        bad = wire.AirProof.from_bytes(proof)
        bad.fri_commitment = decoded
        if air_accepts(zp, bad.to_bytes(), SEC_LABEL):
            accepted_at.append(pos)
        else:
            rejected += 1
    report("byte-fuzz-commitment", trials, rejected)
    assert rejected == trials, accepted_at


def test_truncated_and_empty_inputs_rejected(security):
    zp, _, proof = security
    fp = z.fri.params_new(4, 8, 4)
    attempts = rejected = 0
    rejected += not air_accepts(zp, b"", SEC_LABEL)
    attempts += 1
    rejected += not fri_accepts(b"", b"", fp, b"security/empty")
    attempts += 1
    for cut in range(1, 9):
        rejected += not air_accepts(zp, proof[:-cut], SEC_LABEL)
        attempts += 1
    rejected += not air_accepts(zp, proof + b"\xAA", SEC_LABEL)
    attempts += 1
    report("truncated-empty-inputs", attempts, rejected)
    assert rejected == attempts


def _tree(leaves):
    """MerkleTree::new over byte leaves (merkle.rs:40-76) from the product's
    hash_leaf / hash_node: layers, leaf level first."""
    layer = [z.merkle.hash_leaf(l) for l in leaves]
    while len(layer) & (len(layer) - 1):
        layer.append(layer[-1])
    layers = [layer]
    while len(layer) > 1:
        layer = [z.merkle.hash_node(layer[2 * i], layer[2 * i + 1]) for i in range(len(layer) // 2)]
        layers.append(layer)
    return layers


def _prove(layers, index):
    """MerkleTree::prove (merkle.rs:139-158)."""
    path = []
    for layer in layers[:-1]:
        sib = index ^ 1
        path.append((layer[sib], index % 2 == 0))
        index //= 2
    return path


def test_merkle_hashes_match_oracle():
    for data in [b"", b"\x00" * 32, bytes(range(200)), b"x" * 1500]:
        assert z.merkle.hash_leaf(data) == o.blake3(b"\x00" + data)
    l, r = bytes(range(32)), bytes(range(32, 64))
    assert z.merkle.hash_node(l, r) == o.node(l, r) == o.blake3(b"\x01" + l + r)


def test_merkle_adversarial_vectors():
    leaves = [bytes([i]) * 32 for i in range(8)]
    layers = _tree(leaves)
    root = layers[-1][0]
    attempts = rejected = 0
    for i, leaf in enumerate(leaves):
        path = _prove(layers, i)
        assert z.merkle.verify(root, leaf, i, path)
        bad = bytes([leaf[0] ^ 1]) + leaf[1:]
        checks = [
            z.merkle.verify(root, bad, i, path),
            z.merkle.verify(root, leaf, i + 8, path),
            z.merkle.verify(root, leaf, i ^ 1, path),
            z.merkle.verify(root, leaf, i, path[:-1]),
            z.merkle.verify(root, leaf, i, path + [(bytes(32), True)]),
        ]
        attempts += len(checks)
        rejected += sum(not c for c in checks)
    report("merkle-adversarial", attempts, rejected)
    assert rejected == attempts


# ---- tests/redteam.rs ----------------------------------------------------
RT_LABEL = b"redteam/air"


@pytest.fixture(scope="module")
def redteam_pair():
    """redteam.rs:29-45: the true statement and b_last + 1 over one trace."""
    bl = b_last(1, 1, 32)
    z1, o1 = air_pair(32, 8, 8, 16, 1, 1, bl)
    z2, o2 = air_pair(32, 8, 8, 16, 1, 1, (bl + 1) % P)
    p1 = o.air_prove(o1, o.Transcript(RT_LABEL))
    p2 = o.air_prove(o2, o.Transcript(RT_LABEL))
    return z1, p1, z2, p2


def test_false_statements_never_verify():
    attempts = rejected = 0
    zr = z.range.params_new(8, 8, 8, 8, 4, 300)
    orr = o.range_params_new(8, 8, 8, 8, 4, 300)
    rp = o.range_prove(orr, o.Transcript(b"redteam/range-false"))
    with pytest.raises(z.Z0pError):
        z.range.verify(zr, rp, z.Transcript(b"redteam/range-false"))
    attempts += 1
    rejected += 1
    table = [10, 20, 30, 40]
    zt = z.table.params_new(8, 8, 8, 4, 999)
    ot = o.table_params_new(8, 8, 8, 4, 999)
    tp = o.table_prove(ot, table, o.Transcript(b"redteam/table-false"))
    with pytest.raises(z.Z0pError):
        z.table.verify(zt, table, tp, z.Transcript(b"redteam/table-false"))
    attempts += 1
    rejected += 1
    report("false-statements", attempts, rejected)


def test_transplant_attacks_rejected(redteam_pair):
    z1, p1, z2, p2 = redteam_pair
    assert air_accepts(z1, p1, RT_LABEL)
    assert not air_accepts(z2, p2, RT_LABEL)
    d1, d2 = wire.AirProof.from_bytes(p1), wire.AirProof.from_bytes(p2)
    assert d1.column_cap == d2.column_cap  # same trace: the caps agree
    bl3 = b_last(2, 3, 32)
    _, o3 = air_pair(32, 8, 8, 16, 2, 3, bl3)
    d3 = wire.AirProof.from_bytes(o.air_prove(o3, o.Transcript(RT_LABEL)))
    assert d1.column_cap != d3.column_cap
    forgeries = []
    bad = wire.AirProof.from_bytes(p1)
    bad.openings[0] = d2.openings[0]
    forgeries.append(bad)
    bad = wire.AirProof.from_bytes(p1)
    bad.column_cap = d3.column_cap
    forgeries.append(bad)
    bad = wire.AirProof.from_bytes(p1)
    bad.fri_proof = d2.fri_proof
    forgeries.append(bad)
    bad = wire.AirProof.from_bytes(p1)
    bad.fri_commitment = d2.fri_commitment
    forgeries.append(bad)
    rejected = sum(not air_accepts(z1, f.to_bytes(), RT_LABEL) for f in forgeries)
    report("transplant-attacks", len(forgeries), rejected)
    assert rejected == len(forgeries)


def test_fri_surgery_rejected():
    zp, op = z.fri.params_new(4, 8, 4), o.fri_params_new(4, 8, 4)
    pa_c = [c % P for c in _fri_coeffs16(0x9E3779B9, 17)]
    pb_c = [c % P for c in _fri_coeffs16(0x12345678, 99)]
    t = o.Transcript(b"redteam/fri-a")
    sa = o.FriState(o.FRI_KIND_BASE, pa_c, op, t)
    ca = sa.commitment()
    pa, _ = sa.open(t)
    t = o.Transcript(b"redteam/fri-b")
    sb = o.FriState(o.FRI_KIND_BASE, pb_c, op, t)
    cb = sb.commitment()
    pb, _ = sb.open(t)
    assert fri_accepts(ca, pa, zp, b"redteam/fri-a")
    A, B = wire.FriProof.from_bytes(pa), wire.FriProof.from_bytes(pb)
    swapped = wire.FriProof.from_bytes(pa)
    swapped.final_poly = B.final_poly
    cut = wire.FriProof.from_bytes(pa)
    cut.queries.pop()
    dup = wire.FriProof.from_bytes(pa)
    dup.queries.append(A.queries[0])
    cases = [(cb, pa), (ca, swapped.to_bytes()), (ca, cut.to_bytes()), (ca, dup.to_bytes())]
    rejected = sum(not fri_accepts(c, p, zp, b"redteam/fri-a") for c, p in cases)
    report("fri-surgery", len(cases), rejected)
    assert rejected == len(cases)


def test_fold_tamper_rejected():
    segs = o.chained_traces(8, 2)
    out_b = segs[-1][1][-1]
    zp = z.fold.params_new(8, 8, 8, 4, 1, 1, out_b)
    op = o.fold_params_new(8, 8, 8, 4, 1, 1, out_b)
    t = o.Transcript(b"redteam/fold")
    acc = o.FoldAccumulator(op)
    for ca, cb in segs:
        acc.add_segment(ca, cb, t)
    proof = acc.finalize(t)

    def accepts(pb):
        try:
            z.fold.verify(zp, pb, z.Transcript(b"redteam/fold"))
            return True
        except z.Z0pError:
            return False

    assert accepts(proof)
    forgeries = []
    bad = wire.IvcProof.from_bytes(proof)
    bad.steps[0].claims[2] = (bad.steps[0].claims[2] + 1) % P
    forgeries.append(bad)
    bad = wire.IvcProof.from_bytes(proof)
    bad.steps[1].openings[0].cols[0] = (bad.steps[1].openings[0].cols[0] + 1) % P
    forgeries.append(bad)
    bad = wire.IvcProof.from_bytes(proof)
    c = bad.steps[0].column_cap[0]
    bad.steps[0].column_cap[0] = bytes([c[0] ^ 1]) + c[1:]
    forgeries.append(bad)
    bad = wire.IvcProof.from_bytes(proof)
    bad.steps.pop()
    forgeries.append(bad)
    rejected = sum(not accepts(f.to_bytes()) for f in forgeries)
    report("fold-tamper", len(forgeries), rejected)
    assert rejected == len(forgeries)


def test_leaf_node_confusion_rejected():
    leaves = [bytes([i]) * 32 for i in range(4)]
    layers = _tree(leaves)
    root = layers[-1][0]
    path0 = _prove(layers, 0)
    prelude = b"\x01" + leaves[0] + leaves[1]
    node_digest = z.merkle.hash_node(z.merkle.hash_leaf(leaves[0]), z.merkle.hash_leaf(leaves[1]))
    checks = [z.merkle.verify(node_digest, prelude, 0, [])]
    checks += [z.merkle.verify(root, prelude, i, path0) for i in range(4)]
    rejected = sum(not c for c in checks)
    report("leaf-node-confusion", len(checks), rejected)
    assert rejected == len(checks)


def _all_verifiers():
    """Every product verifier on its honest reference instance."""
    bl = b_last(1, 1, 32)
    za, oa = air_pair(32, 4, 8, 8, 1, 1, bl)
    fz, fo = z.fri.params_new(4, 8, 4), o.fri_params_new(4, 8, 4)
    t = o.Transcript(b"hostile/fri")
    st = o.FriState(o.FRI_KIND_BASE, [c % P for c in _fri_coeffs16(0x9E3779B9, 17)], fo, t)
    fcm = st.commitment()
    zr = z.range.params_new(32, 8, 4, 8, 8, 5)
    zt = z.table.params_new(8, 8, 8, 4, 30)
    segs = o.chained_traces(8, 2)
    zf = z.fold.params_new(8, 8, 8, 4, 1, 1, segs[-1][1][-1])
    table = [10, 20, 30, 40]
    return {
        "air": lambda b: z.air.verify(za, b, z.Transcript(b"hostile")),
        "fri-proof": lambda b: z.fri.verify(fcm, b, fz, z.Transcript(b"hostile/fri")),
        "fri-commitment": lambda b: z.fri.verify(b, b"", fz, z.Transcript(b"hostile/fri")),
        "range": lambda b: z.range.verify(zr, b, z.Transcript(b"hostile")),
        "table": lambda b: z.table.verify(zt, table, b, z.Transcript(b"hostile")),
        "fold": lambda b: z.fold.verify(zf, b, z.Transcript(b"hostile")),
    }


def test_garbage_never_panics_or_verifies():
    verifiers = _all_verifiers()
    state, trials, accepted = 0xBAD5EED512345678, 500, []
    for _ in range(trials):
        state, r = splitmix(state)
        n = r % 320
        data = bytearray(n)
        for i in range(n):
            state, r = splitmix(state)
            data[i] = r % 256
        for name, fn in verifiers.items():
            try:
                fn(bytes(data))
                accepted.append(name)
            except z.Z0pError:
                pass
    report("garbage-hostile", trials * len(verifiers), trials * len(verifiers) - len(accepted))
    assert not accepted


def test_count_limits_fail_closed():
    verifiers = _all_verifiers()
    cases = [
        ("fri-proof", b"\x06" + (0xFFFFFFFF).to_bytes(4, "little")),
        ("air", b"\x06" + (257).to_bytes(4, "little")),
        ("fold", b"\x06" + ((1 << 20) + 1).to_bytes(4, "little")),
        ("fri-commitment", b"\x7f" * 64),
    ]
    rejected = 0
    for name, data in cases:
        with pytest.raises(z.Z0pError) as e:
            verifiers[name](data)
        rejected += e.value.code == 8  # MalformedInput
    report("count-limits", len(cases), rejected)
    assert rejected == len(cases)


def test_transcript_abuse_diverges():
    attempts = rejected = 0
    t1, t2 = z.Transcript(b"redteam/order"), z.Transcript(b"redteam/order")
    t1.absorb(b"a", b"first")
    t1.absorb(b"b", b"second")
    t2.absorb(b"b", b"second")
    t2.absorb(b"a", b"first")
    rejected += t1.challenge_bb(b"x") != t2.challenge_bb(b"x")
    attempts += 1
    t = z.Transcript(b"redteam/replay")
    first = t.challenge_bb(b"same")
    rejected += t.challenge_bb(b"same") != first
    attempts += 1
    t = z.Transcript(b"redteam/range")
    rejected += all((lambda i: i is not None and i < 64)(t.challenge_index(b"q", 64)) for _ in range(200))
    attempts += 1
    bl = b_last(1, 1, 32)
    _, og = air_pair(32, 8, 8, 16, 1, 1, bl, grinding=8, cap=2)
    z0, _ = air_pair(32, 8, 8, 16, 1, 1, bl, grinding=0, cap=2)
    proof = o.air_prove(og, o.Transcript(b"redteam/downgrade"))
    rejected += not air_accepts(z0, proof, b"redteam/downgrade")
    attempts += 1
    report("transcript-abuse", attempts, rejected)
    assert rejected == attempts


def test_rate_boundary_exact_verify_side():
    # redteam.rs:374-416 (the prove-time refusal of bound 32 runs on the
    # GPU in tests/test_gpu_boundary.py).
    zp, op = z.fri.params_new(4, 8, 4), o.fri_params_new(4, 8, 4)
    coeffs = [c % P for c in _fri_coeffs16(0x9E3779B9, 5)]
    lde = o.evaluate_on_coset(coeffs, 4)
    assert len(lde) == 64
    t = o.Transcript(b"redteam/edge")
    st = o.FriState(o.FRI_KIND_EVALS, lde, op, t, degree_bound=31)
    cm = st.commitment()
    pf, _ = st.open(t)
    edge_ok = fri_accepts(cm, pf, zp, b"redteam/edge")
    with pytest.raises(o.OracleError):
        o.FriState(o.FRI_KIND_EVALS, lde, op, o.Transcript(b"redteam/edge"), degree_bound=32)
    high = wire.Commitment.from_bytes(cm)
    high.degree_bound = 32
    raised = not fri_accepts(high.to_bytes(), pf, zp, b"redteam/edge")
    report("rate-boundary-exact", 3, edge_ok + 1 + raised)
    assert edge_ok and raised


def test_zk_roundtrips_and_tamper():
    rng = o.ChaCha20(bytes([77]) * 32, bytes([88]) * 12)
    attempts = rejected = 0
    za, oa = air_pair(16, 8, 8, 8, 1, 1, b_last(1, 1, 16))
    ap = o.air_prove_zk(oa, 128, rng, o.Transcript(b"redteam/zk-air"))
    rejected += air_accepts(za, ap, b"redteam/zk-air")
    attempts += 1
    bad = wire.AirProof.from_bytes(ap)
    bad.openings[0].a = (bad.openings[0].a + 1) % P
    rejected += not air_accepts(za, bad.to_bytes(), b"redteam/zk-air")
    attempts += 1

    def range_ok(zr, proof, label):
        try:
            z.range.verify(zr, proof, z.Transcript(label))
            return True
        except z.Z0pError:
            return False

    def table_ok(zt, table, proof, label):
        try:
            z.table.verify(zt, table, proof, z.Transcript(label))
            return True
        except z.Z0pError:
            return False

    zr, orr = z.range.params_new(8, 16, 8, 8, 8, 42), o.range_params_new(8, 16, 8, 8, 8, 42)
    rp = o.range_prove_zk(orr, 8, rng, o.Transcript(b"redteam/zk-range"))
    rejected += range_ok(zr, rp, b"redteam/zk-range")
    attempts += 1
    table = [10, 20, 30, 40]
    zo, oo = z.range.params_new(8, 16, 8, 8, 8, 70000), o.range_params_new(8, 16, 8, 8, 8, 70000)
    rp = o.range_prove_zk(oo, 8, rng, o.Transcript(b"redteam/zk-range-oor"))
    rejected += not range_ok(zo, rp, b"redteam/zk-range-oor")
    attempts += 1
    zn, on = z.table.params_new(8, 8, 8, 8, 99), o.table_params_new(8, 8, 8, 8, 99)
    tp = o.table_prove_zk(on, table, 8, rng, o.Transcript(b"redteam/zk-table-nm"))
    rejected += not table_ok(zn, table, tp, b"redteam/zk-table-nm")
    attempts += 1
    zt, ot = z.table.params_new(8, 8, 8, 8, 30), o.table_params_new(8, 8, 8, 8, 30)
    tp = o.table_prove_zk(ot, table, 8, rng, o.Transcript(b"redteam/zk-table"))
    rejected += table_ok(zt, table, tp, b"redteam/zk-table")
    attempts += 1
    report("zk-roundtrips-tamper", attempts, rejected)
    assert rejected == attempts


# ---- tests/hardening.rs --------------------------------------------------
HARD = b"hardening"


def _valid_air():
    """hardening.rs:16-19: N = 32, B4 Q8 F8."""
    return air_pair(32, 4, 8, 8, 1, 1, b_last(1, 1, 32))


def test_air_invalid_params_error_without_panic():
    zp, op = _valid_air()
    proof = o.air_prove(op, o.Transcript(HARD))
    assert air_accepts(zp, proof, HARD)
    invalid = [dict(n_steps=0), dict(n_steps=3), dict(blowup=1), dict(blowup=3), dict(blowup=(1 << 64) - 1),
               dict(num_queries=0), dict(num_queries=(1 << 64) - 1), dict(final_size=0), dict(final_size=3),
               dict(grinding_bits=64)]
    for change in invalid:
        bad, _ = _valid_air()
        for k, v in change.items():
            setattr(bad, k, v)
        assert not air_accepts(bad, proof, HARD), change


def test_range_table_fold_invalid_params_error_without_panic():
    zr, orr = z.range.params_new(32, 8, 4, 8, 8, 5), o.range_params_new(32, 8, 4, 8, 8, 5)
    rp = o.range_prove(orr, o.Transcript(HARD))
    z.range.verify(zr, rp, z.Transcript(HARD))
    for change in [dict(n_rows=0), dict(n_rows=3), dict(n_bits=0), dict(n_bits=33), dict(blowup=1),
                   dict(num_queries=0), dict(final_size=0)]:
        bad = z.range.params_new(32, 8, 4, 8, 8, 5)
        for k, v in change.items():
            setattr(bad, k, v)
        with pytest.raises(z.Z0pError):
            z.range.verify(bad, rp, z.Transcript(HARD))
    table = [10, 20, 30, 40]
    zt, ot = z.table.params_new(8, 8, 8, 4, 30), o.table_params_new(8, 8, 8, 4, 30)
    tp = o.table_prove(ot, table, o.Transcript(HARD))
    z.table.verify(zt, table, tp, z.Transcript(HARD))
    for change in [dict(n_rows=0), dict(n_rows=3), dict(blowup=1), dict(num_queries=0),
                   dict(num_queries=(1 << 64) - 1), dict(final_size=3)]:
        bad = z.table.params_new(8, 8, 8, 4, 30)
        for k, v in change.items():
            setattr(bad, k, v)
        with pytest.raises(z.Z0pError):
            z.table.verify(bad, table, tp, z.Transcript(HARD))
    ca, cb = o.chained_traces(32, 1)[0]
    zf = z.fold.params_new(32, 4, 8, 8, ca[0], cb[0], cb[-1])
    of = o.fold_params_new(32, 4, 8, 8, ca[0], cb[0], cb[-1])
    t = o.Transcript(HARD)
    acc = o.FoldAccumulator(of)
    acc.add_segment(ca, cb, t)
    fp = acc.finalize(t)
    z.fold.verify(zf, fp, z.Transcript(HARD))
    for change in [dict(seg_steps=0), dict(seg_steps=3), dict(blowup=1), dict(blowup=3), dict(num_queries=0),
                   dict(final_size=0)]:
        bad = z.fold.params_new(32, 4, 8, 8, ca[0], cb[0], cb[-1])
        for k, v in change.items():
            setattr(bad, k, v)
        with pytest.raises(z.Z0pError):
            z.fold.verify(bad, fp, z.Transcript(HARD))


def test_air_decoder_fuzz_every_bit_rejected():
    """hardening.rs:223-250, strengthened: every single-bit flip of the
    proof (decode, then verify) is rejected, and so are all truncations and
    the extended and prefixed encodings."""
    zp, op = _valid_air()
    proof = o.air_prove(op, o.Transcript(HARD))
    accepted = []
    for i in range(len(proof)):
        for bit in range(8):
            bad = bytearray(proof)
            bad[i] ^= 1 << bit
            if air_accepts(zp, bytes(bad), HARD):
                accepted.append((i, bit))
    for cut in range(len(proof)):
        assert not air_accepts(zp, proof[:cut], HARD)
    assert not air_accepts(zp, proof + bytes(17), HARD)
    assert not air_accepts(zp, b"\xff" + proof, HARD)
    report("air-decoder-fuzz", 8 * len(proof), 8 * len(proof) - len(accepted))
    assert not accepted


def test_one_point_final_layers_are_rejected():
    for ctor in [lambda: z.fri.params_new(4, 8, 1),
                 lambda: z.air.params_new(16, 16, 8, 1, 1, 1, b_last(1, 1, 16)),
                 lambda: z.range.params_new(8, 8, 8, 8, 1, 5),
                 lambda: z.table.params_new(8, 8, 8, 1, 5),
                 lambda: z.fold.params_new(8, 8, 8, 1, 1, 1, 1)]:
        with pytest.raises(z.Z0pError):
            ctor()
    tiny = z.table.for_security(2, 2, 999, 12)
    assert tiny.final_size >= 2
    ot = o.table_params_for_security(2, 2, 999, 12)
    assert bytes(ot) == bytes(tiny)
    tp = o.table_prove(ot, [1, 2], o.Transcript(HARD))
    with pytest.raises(z.Z0pError):
        z.table.verify(tiny, [1, 2], tp, z.Transcript(HARD))


def test_oversized_final_layers_fail_closed():
    with pytest.raises(z.Z0pError):
        z.fri.params_new(4, 8, 2048)
    z.fri.params_new(4, 8, 1024)
    assert z.table.membership_bound((1 << 64) - 1, 2, 0) == (1 << 64) - 1 - 2


def test_next_field_is_in_range_and_uniform_ish():
    rng = z.ChaCha20(bytes([7]) * 32, bytes([9]) * 12)
    buckets = [0] * 8
    draws = 40000
    for _ in range(draws):
        x = rng.next_field()
        assert x < P
        buckets[min(x * 8 // P, 7)] += 1
    for c in buckets:
        assert abs(c - draws // 8) < draws // 8 // 4


def test_maximal_count_headers_fail_fast_in_bounded_memory():
    """codec.rs:10-13: no vector is sized from an untrusted count. Headers
    that announce the largest legal counts over an empty body must fail as
    MalformedInput quickly, without the allocation the count would need
    (checked in a fresh process: peak RSS grows by under 32 MiB)."""
    import subprocess
    import sys
    import textwrap
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = textwrap.dedent("""
        import resource, struct, sys, time
        sys.path.insert(0, %r)
        sys.path.insert(0, %r)
        import oracle_ffi as o
        import wire
        import z0p_amd as z
        bl = o.fib_b_last(1, 1, 32)
        zp = z.air.params_new(32, 4, 8, 8, 1, 1, bl)
        op = o.air_params_new(32, 4, 8, 8, 1, 1, bl)
        proof = wire.AirProof.from_bytes(o.air_prove(op, o.Transcript(b"hdr")))
        w = wire.Writer()
        w.byte(6)
        w.cap(proof.column_cap)
        proof.fri_commitment.write_body(w)
        proof.fri_proof.write_body(w)
        air_hdr = bytes(w.out) + struct.pack("<I", 1 << 20)
        fz = z.fri.params_new(4, 8, 4)
        cm = proof.fri_commitment.to_bytes()
        zf = z.fold.params_new(8, 8, 8, 4, 1, 1, 1)
        zr = z.range.params_new(32, 8, 4, 8, 8, 5)
        zt = z.table.params_new(8, 8, 8, 4, 30)
        cases = [
            lambda: z.air.verify(zp, air_hdr, z.Transcript(b"hdr")),
            lambda: z.fri.verify(cm, b"\\x06" + struct.pack("<I", 1 << 27), fz, z.Transcript(b"hdr")),
            lambda: z.fri.verify(cm, b"\\x06" + struct.pack("<I", 0) + bytes(8) + struct.pack("<I", 1 << 20), fz,
                                 z.Transcript(b"hdr")),
            lambda: z.fold.verify(zf, b"\\x06" + struct.pack("<I", 1 << 20), z.Transcript(b"hdr")),
            lambda: z.range.verify(zr, b"\\x06" + struct.pack("<I", 256), z.Transcript(b"hdr")),
            lambda: z.table.verify(zt, [10, 20, 30, 40], b"\\x06" + struct.pack("<I", 256), z.Transcript(b"hdr")),
        ]
        base = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss
        for i, fn in enumerate(cases):
            t0 = time.perf_counter()
            try:
                fn()
                print("case", i, "accepted"); sys.exit(1)
            except z.Z0pError as e:
                assert e.code == 8, (i, e.code)
            assert time.perf_counter() - t0 < 1.0, i
        grew = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss - base
        print("rss growth KiB", grew)
        sys.exit(0 if grew < 32 * 1024 else 2)
    """) % (root, os.path.join(root, "tests"))
    r = subprocess.run([sys.executable, "-c", script], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout, r.stderr[-2000:])
