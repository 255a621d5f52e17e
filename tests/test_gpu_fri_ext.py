"""The other FRI commitments on the GPU: fri::commit_ext (Ext4 coefficients,
four coordinate LDEs interleaved into layer 0), fri::commit_evals and
fri::commit_ext_evals (caller LDE values, bit-reversed into layer 0). Every
commitment and proof must be byte-identical to the oracle's, advance the
transcript as the oracle does, and verify exactly when the oracle's does."""
import numpy as np
import pytest

import oracle_ffi as o
from helpers import P

import z0p_amd as z

pytestmark = pytest.mark.gpu


def _params(m, blowup=4, nq=12, grinding=0, cap=0):
    final = 4 if (m.bit_length() - 1) % 2 == 0 else 8
    return (o.fri_params_new(blowup, nq, final, grinding, cap),
            z.fri.params_new(blowup, nq, final, grinding, cap))


def _check(got, want, zp, op, label):
    assert got == want
    ok_o = o.fri_verify(want[0], want[1], op, o.Transcript(label)) == 0
    try:
        z.fri.verify(got[0], got[1], zp, z.Transcript(label))
        ok_z = True
    except z.Z0pError:
        ok_z = False
    assert ok_o == ok_z
    return ok_o


@pytest.mark.parametrize("n,blowup", [(16, 4), (1 << 10, 8), (1 << 12, 16), (1 << 16, 16)])
def test_commit_ext_byte_identical(ctx, n, blowup):
    rng = np.random.default_rng(n)
    coeffs = rng.integers(0, P, 4 * n, dtype=np.uint32)
    op, zp = _params(n * blowup, blowup=blowup, grinding=6, cap=2)
    print("commit_ext n", n, "blowup", blowup, flush=True)
    ot, zt = o.Transcript(b"ext"), z.Transcript(b"ext")
    want = o.fri_commit_open_kind(o.FRI_KIND_EXT, coeffs, op, ot)
    got = z.fri.commit_open_ext(coeffs, zp, zt, ctx)
    assert _check(got, want, zp, op, b"ext")
    assert zt.challenge_bytes(b"after") == ot.challenge_bytes(b"after")


def test_commit_ext_production_shape(ctx):
    # fri::Params::for_security over 2^14 Ext4 coefficients (2^18 LDE).
    n = 1 << 14
    coeffs = np.random.default_rng(9).integers(0, P, 4 * n, dtype=np.uint32)
    op, zp = o.fri_params_for_security(n), z.fri.for_security(n)
    want = o.fri_commit_open_kind(o.FRI_KIND_EXT, coeffs, op, o.Transcript(b"prod"))
    got = z.fri.commit_open_ext(coeffs, zp, z.Transcript(b"prod"), ctx)
    assert _check(got, want, zp, op, b"prod")


@pytest.mark.parametrize("m", [64, 1 << 13, 1 << 14, 1 << 20])
def test_commit_evals_low_degree_and_random(ctx, m):
    # Values of a degree < m/4 polynomial verify at bound m/4 - 1; random
    # values are proved the same way and rejected (fri.rs:1880-1910).
    rng = np.random.default_rng(m)
    coeffs = rng.integers(0, P, m // 4, dtype=np.uint32)
    low = np.array(o.evaluate_on_coset([int(c) for c in coeffs], 4), dtype=np.uint32)
    rand = rng.integers(0, P, m, dtype=np.uint32)
    op, zp = _params(m, grinding=4, cap=1)
    print("commit_evals m", m, flush=True)
    for vals, honest in [(low, True), (rand, False)]:
        want = o.fri_commit_open_kind(o.FRI_KIND_EVALS, vals, op, o.Transcript(b"ev"), degree_bound=m // 4 - 1)
        got = z.fri.commit_open_evals(vals, m // 4 - 1, zp, z.Transcript(b"ev"), ctx)
        assert _check(got, want, zp, op, b"ev") == honest


@pytest.mark.parametrize("m", [64, 1 << 12, 1 << 16])
def test_commit_ext_evals_byte_identical(ctx, m):
    rng = np.random.default_rng(m + 1)
    base = rng.integers(0, P, 4 * (m // 8), dtype=np.uint32)
    low = np.array(o.evaluate_on_coset_ext([int(v) for v in base], 8), dtype=np.uint32)
    rand = rng.integers(0, P, 4 * m, dtype=np.uint32)
    op, zp = _params(m, blowup=8, cap=3)
    print("commit_ext_evals m", m, flush=True)
    for vals, honest in [(low, True), (rand, False)]:
        ot, zt = o.Transcript(b"xe"), z.Transcript(b"xe")
        want = o.fri_commit_open_kind(o.FRI_KIND_EXT_EVALS, vals, op, ot, degree_bound=m // 8 - 1)
        got = z.fri.commit_open_ext_evals(vals, m // 8 - 1, zp, zt, ctx)
        assert _check(got, want, zp, op, b"xe") == honest
        assert zt.challenge_bytes(b"after") == ot.challenge_bytes(b"after")


def test_commit_variants_error_codes(ctx):
    op, zp = _params(64)
    cases = [
        (o.FRI_KIND_EXT, [], 0, lambda: z.fri.commit_open_ext([], zp, z.Transcript(b"e"), ctx)),
        (o.FRI_KIND_EXT, [1] * (4 * 12), 0, lambda: z.fri.commit_open_ext([1] * 48, zp, z.Transcript(b"e"), ctx)),
        (o.FRI_KIND_EVALS, [1] * 64, 64, lambda: z.fri.commit_open_evals([1] * 64, 64, zp, z.Transcript(b"e"), ctx)),
        (o.FRI_KIND_EVALS, [1] * 48, 3, lambda: z.fri.commit_open_evals([1] * 48, 3, zp, z.Transcript(b"e"), ctx)),
        (o.FRI_KIND_EXT_EVALS, [1] * 256, 40,
         lambda: z.fri.commit_open_ext_evals([1] * 256, 40, zp, z.Transcript(b"e"), ctx)),
    ]
    for kind, vals, db, call in cases:
        with pytest.raises(o.OracleError) as oe:
            o.fri_commit_open_kind(kind, vals, op, o.Transcript(b"e"), degree_bound=db)
        with pytest.raises(z.Z0pError) as ze:
            call()
        assert ze.value.code == oe.value.code, (kind, len(vals), db)
