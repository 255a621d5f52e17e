"""The FRI boundary beyond commit-and-open in one call, against the oracle:
the query indices fri::open and fri::verify return (fri.rs:645-666, 686),
fri::commit and fri::open as separate calls over a device-resident prover
state (fri.rs:276-294, 408-666), and Transcript::grind on the device for
any label (transcript.rs:182-351)."""
import ctypes

import pytest

import oracle_ffi as o
from helpers import P, fri_bench_coeffs, splitmix64, splitmix_fields

import z0p_amd as z
from z0p_amd import _ffi

pytestmark = pytest.mark.gpu


def _pair(blowup, nq, final, grind=0, cap=0):
    return (z.fri.params_new(blowup, nq, final, grind, cap), o.fri_params_new(blowup, nq, final, grind, cap))


def _rand(n, seed):
    out, st = [], seed
    for _ in range(n):
        st, x = splitmix64(st)
        out.append(x % P)
    return out


def _low_degree_evals(n_coeffs, blowup, seed, ext=False):
    """LDE values of a random polynomial on the 31-coset (natural order)."""
    if ext:
        return o.evaluate_on_coset_ext(_rand(4 * n_coeffs, seed), blowup)
    return o.evaluate_on_coset(_rand(n_coeffs, seed), blowup)


CASES = [
    # (name, kind, values, degree_bound, params: "prod" (for_security) or Params::new fields)
    ("coeffs-prod-1024", o.FRI_KIND_BASE, fri_bench_coeffs(1024), 0, "prod"),
    ("coeffs-grind8", o.FRI_KIND_BASE, _rand(256, 3), 0, (4, 12, 16, 8, 1)),
    ("ext", o.FRI_KIND_EXT, _rand(4 * 64, 5), 0, (8, 10, 8, 6, 2)),
    ("evals", o.FRI_KIND_EVALS, None, 63, (4, 9, 4, 0, 0)),
    ("ext-evals", o.FRI_KIND_EXT_EVALS, None, 127, (8, 11, 16, 5, 1)),
]


def _materialise(kind, values, bound):
    if values is not None:
        return values
    if kind == o.FRI_KIND_EVALS:
        return _low_degree_evals(bound + 1, 4, 11)
    return _low_degree_evals(bound + 1, 8, 13, ext=True)


def _params(spec, vals):
    if spec == "prod":
        return z.fri.for_security(len(vals)), o.fri_params_for_security(len(vals))
    return _pair(*spec)


_Z_COMBINED = {
    o.FRI_KIND_BASE: lambda v, b, zp, zt, ctx: z.fri.commit_open(v, zp, zt, ctx, return_indices=True),
    o.FRI_KIND_EXT: lambda v, b, zp, zt, ctx: z.fri.commit_open_ext(v, zp, zt, ctx, return_indices=True),
    o.FRI_KIND_EVALS: lambda v, b, zp, zt, ctx: z.fri.commit_open_evals(v, b, zp, zt, ctx, return_indices=True),
    o.FRI_KIND_EXT_EVALS: lambda v, b, zp, zt, ctx: z.fri.commit_open_ext_evals(v, b, zp, zt, ctx,
                                                                               return_indices=True),
}
_Z_COMMIT = {
    o.FRI_KIND_BASE: lambda v, b, zp, zt, ctx: z.fri.commit(v, zp, zt, ctx),
    o.FRI_KIND_EXT: lambda v, b, zp, zt, ctx: z.fri.commit_ext(v, zp, zt, ctx),
    o.FRI_KIND_EVALS: lambda v, b, zp, zt, ctx: z.fri.commit_evals(v, b, zp, zt, ctx),
    o.FRI_KIND_EXT_EVALS: lambda v, b, zp, zt, ctx: z.fri.commit_ext_evals(v, b, zp, zt, ctx),
}


@pytest.mark.parametrize("name,kind,values,bound,spec", CASES, ids=[c[0] for c in CASES])
def test_commit_open_indices_equal_oracle(ctx, name, kind, values, bound, spec):
    """fri::open's (Proof, Vec<usize>) through the combined entry points: the
    bytes and the indices equal the oracle's, and fri::verify (product and
    oracle) returns the same indices."""
    vals = _materialise(kind, values, bound)
    zp, op = _params(spec, vals)
    print("case", name, zp.shape(), flush=True)
    zt, ot = z.Transcript(b"idx/" + name.encode()), o.Transcript(b"idx/" + name.encode())
    cm, pf, idx = _Z_COMBINED[kind](vals, bound, zp, zt, ctx)
    ocm, opf, oidx = o.fri_commit_open_kind(kind, vals, op, ot, bound, return_indices=True)
    assert (cm, pf) == (ocm, opf)
    assert idx == oidx and len(idx) == zp.num_queries
    assert zt.challenge_bytes(b"after") == ot.challenge_bytes(b"after")
    assert z.fri.verify(cm, pf, zp, z.Transcript(b"idx/" + name.encode())) == idx
    got = []
    assert o.fri_verify(cm, pf, op, o.Transcript(b"idx/" + name.encode()), got) == 0
    assert got == idx


@pytest.mark.parametrize("name,kind,values,bound,spec", CASES, ids=[c[0] for c in CASES])
def test_commit_then_open_equals_oracle(ctx, name, kind, values, bound, spec):
    """fri::commit, a caller absorb, then fri::open (twice, from clones of
    the transcript): equal to the oracle doing the same, and with no absorb
    between, equal to the combined call byte for byte."""
    vals = _materialise(kind, values, bound)
    zp, op = _params(spec, vals)
    label = b"split/" + name.encode()
    print("case", name, flush=True)
    zt, ot = z.Transcript(label), o.Transcript(label)
    st = _Z_COMMIT[kind](vals, bound, zp, zt, ctx)
    ost = o.FriState(kind, vals, op, ot, bound)
    assert st.commitment() == ost.commitment()
    zt.absorb(b"outer/claim", b"between commit and open")
    ot.absorb(b"outer/claim", b"between commit and open")
    zt2 = zt.clone()
    pf, idx = st.open(zt)
    opf, oidx = ost.open(ot)
    assert (pf, idx) == (opf, oidx)
    assert zt.challenge_bytes(b"after") == ot.challenge_bytes(b"after")
    # The state stays usable: a second open from the same position repeats.
    assert st.open(zt2) == (pf, idx)
    st.close()
    # Without the absorb the split calls are the combined call.
    zt3 = z.Transcript(label)
    st = _Z_COMMIT[kind](vals, bound, zp, zt3, ctx)
    pf3, idx3 = st.open(zt3)
    cm, pf_all, idx_all = _Z_COMBINED[kind](vals, bound, zp, z.Transcript(label), ctx)
    assert (st.commitment(), pf3, idx3) == (cm, pf_all, idx_all)
    st.close()


def test_split_commit_open_at_production_size(ctx):
    """The 2^22-coefficient commitment (2^26-point LDE) split into commit and
    open equals the combined call, whose bytes the golden suite pins."""
    n = 1 << 22
    coeffs = splitmix_fields(n, 0)
    zp = z.fri.for_security(n)
    cm, pf, idx = z.fri.commit_open(coeffs, zp, z.Transcript(b"bench/fri-prod-big"), ctx, return_indices=True)
    zt = z.Transcript(b"bench/fri-prod-big")
    state = z.fri.commit(coeffs, zp, zt, ctx)
    assert state.commitment() == cm
    assert state.open(zt) == (pf, idx)
    state.close()
    assert len(cm) + len(pf) == 181390
    assert z.fri.verify(cm, pf, zp, z.Transcript(b"bench/fri-prod-big")) == idx


def test_fri_state_errors(ctx):
    """The split commit fails as the combined call does (the reference's
    error for an empty polynomial), and a too-small index buffer is an
    argument error checked before any work."""
    zp, op = _pair(4, 8, 4)
    with pytest.raises(o.OracleError) as oe:
        o.FriState(o.FRI_KIND_BASE, [], op, o.Transcript(b"e"))
    with pytest.raises(z.Z0pError) as e:
        z.fri.commit([], zp, z.Transcript(b"e"), ctx)
    assert e.value.code == oe.value.code
    cm, pf = _ffi.u8p(), _ffi.u8p()
    cn, pn = ctypes.c_size_t(), ctypes.c_size_t()
    arr, ptr = _ffi.u32_buf(_rand(16, 1))
    small = (ctypes.c_uint64 * 4)()
    rc = _ffi.lib().z0p_fri_commit_open(ctx.handle, ptr, 16, ctypes.byref(zp), z.Transcript(b"e").handle,
                                        ctypes.byref(cm), ctypes.byref(cn), ctypes.byref(pf), ctypes.byref(pn),
                                        small, 4)
    assert rc == 102
    del arr


LABELS = [b"fri-pow", b"", b"p", b"sixteen-bytes-xx", b"seventeen-bytes-x", b"twenty-byte-label-xx",
          b"twenty-three-bytes-labl", b"twenty-four-bytes-label!", b"L" * 81, b"q" * 120, b"z" * 976]


@pytest.mark.parametrize("label", LABELS, ids=[str(len(l)) for l in LABELS])
def test_transcript_grind_equals_oracle(ctx, label):
    """Transcript::grind on the device = the oracle's sequential scan for
    labels whose nonce lands in every tail position (one block, straddling
    two blocks, after full head blocks), and the transcript continues equal."""
    for bits in (0, 1, 8, 13):
        zt, ot = z.Transcript(b"grind/%d" % len(label)), o.Transcript(b"grind/%d" % len(label))
        zt.absorb(b"stmt", b"x" * bits)
        ot.absorb(b"stmt", b"x" * bits)
        print("label", len(label), "bits", bits, flush=True)
        got = zt.grind(label, bits, ctx)
        want = ot.grind(label, bits)
        assert got == want, (len(label), bits)
        assert zt.challenge_bytes(b"after") == ot.challenge_bytes(b"after")


def test_transcript_grind_limits(ctx):
    # difficulty above 32 fails closed: nonce 0, absorbed (transcript.rs:207-213)
    zt, ot = z.Transcript(b"lim"), o.Transcript(b"lim")
    assert zt.grind(b"pow", 33, ctx) == 0
    assert ot.grind(b"pow", 33) == 0
    assert zt.challenge_bytes(b"after") == ot.challenge_bytes(b"after")
    # a label whose head and nonce exceed one chunk is refused, untouched
    zt = z.Transcript(b"lim")
    before = zt.clone().challenge_bytes(b"probe")
    with pytest.raises(z.Z0pError) as e:
        zt.grind(b"y" * 977, 4, ctx)
    assert e.value.code == 8
    assert zt.challenge_bytes(b"probe") == before


def test_transcript_grind_20_bits_matches_proof_nonce(ctx):
    """The production 20-bit grind: the exported call reproduces the nonce
    the oracle finds, and verify_grind accepts it on a fresh transcript."""
    zt, ot = z.Transcript(b"g20"), o.Transcript(b"g20")
    got = zt.grind(b"fri-pow", 20, ctx)
    assert got == ot.grind(b"fri-pow", 20)
    vt = z.Transcript(b"g20")
    assert vt.verify_grind(b"fri-pow", got, 20)
    assert vt.challenge_bytes(b"after") == zt.challenge_bytes(b"after")
