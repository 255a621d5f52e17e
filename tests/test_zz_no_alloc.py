"""The allocation-free grind entry point against the allocating one, through
the C ABI: z0p_transcript_grind_noalloc must return the nonce that
z0p_transcript_grind returns for the same transcript and leave the
transcript in the same state, call after call on one GPU, with a smaller
input following a larger one (the shared device array keeps the earlier
call's contents) and from a second context on the same device."""
import ctypes

import pytest

import z0p_amd as z
from z0p_amd import _ffi

pytestmark = pytest.mark.gpu


def _grind(fn, ctx, t, label, bits):
    nonce = ctypes.c_uint64()
    rc = fn(ctx.handle, t.handle, label, bits, ctypes.byref(nonce))
    return rc, nonce.value


def _pair(ctx, label, bits, domain):
    """(nonce, next challenge) from the allocating and the no-alloc entry on
    twin transcripts."""
    L = _ffi.lib()
    t_ref, t_new = z.Transcript(domain), z.Transcript(domain)
    for t in (t_ref, t_new):
        t.absorb(b"stmt", domain * 3)
    rc_ref, n_ref = _grind(L.z0p_transcript_grind, ctx, t_ref, label, bits)
    rc_new, n_new = _grind(L.z0p_transcript_grind_noalloc, ctx, t_new, label, bits)
    assert rc_ref == 0 and rc_new == 0, (rc_ref, rc_new)
    return (n_ref, t_ref.challenge_bytes(b"after")), (n_new, t_new.challenge_bytes(b"after"))


def test_noalloc_grind_matches_allocating_entry():
    ctx = z.Context(0)
    try:
        # A larger input first (a 120-byte label: two head blocks before the
        # nonce's block, 16 bits), then smaller ones on the same array.
        cases = [(b"q" * 120, 16, b"na/0"), (b"fri-pow", 9, b"na/1"), (b"seventeen-bytes-x", 12, b"na/2"),
                 (b"", 0, b"na/3")]
        for label, bits, domain in cases:
            print("noalloc grind label", len(label), "bits", bits, flush=True)
            ref, new = _pair(ctx, label, bits, domain)
            assert new == ref, (len(label), bits)
        # The production 20-bit grind.
        ref, new = _pair(ctx, b"fri-pow", 20, b"na/20")
        assert new == ref
    finally:
        ctx.close()


def test_noalloc_grind_second_context_and_limits():
    a, b = z.Context(0), z.Context(0)
    try:
        # Two contexts on one device share the array; calls in sequence agree.
        ref, new = _pair(a, b"pow", 14, b"na/ctx")
        assert new == ref
        ref2, new2 = _pair(b, b"pow", 14, b"na/ctx")
        assert new2 == ref2 == ref
        L = _ffi.lib()
        # difficulty above 32: nonce 0, absorbed, as the allocating entry does
        ta, tb = z.Transcript(b"lim"), z.Transcript(b"lim")
        assert _grind(L.z0p_transcript_grind, a, ta, b"p", 40) == (0, 0)
        assert _grind(L.z0p_transcript_grind_noalloc, a, tb, b"p", 40) == (0, 0)
        assert ta.challenge_bytes(b"x") == tb.challenge_bytes(b"x")
        # a label beyond one chunk is refused with MalformedInput
        assert _grind(L.z0p_transcript_grind_noalloc, a, z.Transcript(b"lim"), b"y" * 977, 4)[0] == 8
        assert _grind(L.z0p_transcript_grind_noalloc, a, z.Transcript(b"lim"), b"ok", 3)[0] == 0
    finally:
        a.close()
        b.close()
