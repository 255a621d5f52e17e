"""Pins the oracle's other FRI commitments: fri::commit_ext (Ext4
coefficients, fri.rs:444-461) with evaluate_on_coset_ext (ntt.rs:394-420),
fri::commit_evals (fri.rs:473-500) and fri::commit_ext_evals (fri.rs:510-549).
Checks: Horner evaluation at the coset points, byte identity with the
already-pinned base commitment where the inputs coincide, and the reference's
own acceptance and rejection tests (fri.rs:1680-1697, 1880-1910)."""
import pytest

import oracle_ffi as o
from helpers import P, splitmix64


def _words(count, seed):
    out, st = [], seed
    for _ in range(count):
        st, r = splitmix64(st)
        out.append(r % P)
    return out


def _ext_mul_base(acc4, x):
    return [(a * x) % P for a in acc4]


def test_evaluate_on_coset_ext_is_horner():
    n, blowup = 8, 4
    coeffs = _words(4 * n, 11)
    got = o.evaluate_on_coset_ext(coeffs, blowup)
    m = n * blowup
    w = o.two_adic_generator(m.bit_length() - 1)
    for i in range(m):
        x = 31 * pow(w, i, P) % P
        acc = [0, 0, 0, 0]
        for k in reversed(range(n)):
            acc = [(a + c) % P for a, c in zip(_ext_mul_base(acc, x), coeffs[4 * k:4 * k + 4])]
        assert got[4 * i:4 * i + 4] == acc, i


@pytest.mark.parametrize("n", [16, 256, 4096])
def test_commit_ext_of_base_coefficients_equals_commit(n):
    # from_base coefficients give commit_inner exactly commit's input
    # (same LDE, poly_len n, bound n - 1), hence the same bytes.
    base = _words(n, n)
    ext = []
    for v in base:
        ext += [v, 0, 0, 0]
    fp = o.fri_params_for_security(n)
    want = o.fri_commit_open(base, fp, o.Transcript(b"ext"))
    got = o.fri_commit_open_kind(o.FRI_KIND_EXT, ext, fp, o.Transcript(b"ext"))
    assert got == want


def test_commit_evals_equals_ext_evals_of_base():
    evals = _words(1024, 3)
    ext = []
    for v in evals:
        ext += [v, 0, 0, 0]
    fp = o.fri_params_new(4, 8, 4)
    a = o.fri_commit_open_kind(o.FRI_KIND_EVALS, evals, fp, o.Transcript(b"ev"), degree_bound=255)
    b = o.fri_commit_open_kind(o.FRI_KIND_EXT_EVALS, ext, fp, o.Transcript(b"ev"), degree_bound=255)
    assert a == b


def test_commit_ext_roundtrip_and_errors():
    n = 64
    coeffs = _words(4 * n, 5)
    fp = o.fri_params_new(8, 8, 8)
    cm, pf = o.fri_commit_open_kind(o.FRI_KIND_EXT, coeffs, fp, o.Transcript(b"x"))
    assert o.fri_verify(cm, pf, fp, o.Transcript(b"x")) == 0
    with pytest.raises(o.OracleError) as e:
        o.fri_commit_open_kind(o.FRI_KIND_EXT, [], fp, o.Transcript(b"x"))
    assert e.value.code == 1  # EmptyInput
    with pytest.raises(o.OracleError) as e:
        o.fri_commit_open_kind(o.FRI_KIND_EXT, coeffs[:4 * 48], fp, o.Transcript(b"x"))
    assert e.value.code == 3  # InvalidDomainSize


def test_commit_evals_roundtrip():
    # fri.rs:1880-1893: a degree-15 polynomial's coset values verify at bound 15.
    fp = o.fri_params_new(4, 8, 4)
    lde = o.evaluate_on_coset(_words(16, 0x5150), 4)
    cm, pf = o.fri_commit_open_kind(o.FRI_KIND_EVALS, list(lde), fp, o.Transcript(b"fri-evals"), degree_bound=15)
    assert o.fri_verify(cm, pf, fp, o.Transcript(b"fri-evals")) == 0


def test_commit_evals_high_degree_rejected():
    # fri.rs:1895-1910 and 1680-1697: random values fail a tight bound.
    evals = _words(64, 0xDEADBEEF00112233)
    fp = o.fri_params_new(4, 8, 4)
    cm, pf = o.fri_commit_open_kind(o.FRI_KIND_EVALS, evals, fp, o.Transcript(b"fri-highdeg"), degree_bound=15)
    assert o.fri_verify(cm, pf, fp, o.Transcript(b"fri-highdeg")) != 0
    fp16 = o.fri_params_new(4, 8, 16)
    cm, pf = o.fri_commit_open_kind(o.FRI_KIND_EVALS, evals, fp16, o.Transcript(b"fri-lowbound"), degree_bound=7)
    assert o.fri_verify(cm, pf, fp16, o.Transcript(b"fri-lowbound")) != 0
    with pytest.raises(o.OracleError) as e:
        o.fri_commit_open_kind(o.FRI_KIND_EVALS, evals, fp, o.Transcript(b"x"), degree_bound=64)
    assert e.value.code == 8  # MalformedInput: rate >= 1
