"""Failure handling of the pre-enqueued (latency-mode) schedule, forced from
the host: a phase gate that gives up at once (zero polls) stands for a host
that never answers. Every kernel still runs to completion: the gates hand the
later phases zeros, so the openings read live = 0 and skip, and the other
phases compute in-bounds values the host discards. The proof must fail with
the device status, and the next proof on the same context must be
byte-identical to the oracle."""
import pytest

import oracle_ffi as o

import z0p_amd as z
from z0p_amd import _ffi

pytestmark = pytest.mark.gpu

DEFAULT_SPIN = 1 << 22


def _params(n):
    bl = o.fib_b_last(1, 1, n)
    return z.air.for_security(n, 1, 1, bl), o.air_params_for_security(n, 1, 1, bl)


@pytest.mark.parametrize("n", [1 << 10, 1 << 16])
def test_gate_timeout_fails_cleanly_then_recovers(ctx, n):
    zp, op = _params(n)
    print("gate timeout, n", n, flush=True)
    _ffi.check(_ffi.lib().z0p_test_set_gate_spin(ctx.handle, 0))
    try:
        with pytest.raises(z.Z0pError) as e:
            z.air.prove(zp, z.Transcript(b"gate"), ctx)
        assert e.value.code == 100
    finally:
        _ffi.check(_ffi.lib().z0p_test_set_gate_spin(ctx.handle, DEFAULT_SPIN))
    print("recovery proof", flush=True)
    got = z.air.prove(zp, z.Transcript(b"gate"), ctx)
    assert got == o.air_prove(op, o.Transcript(b"gate"))


def test_fri_gate_timeout_then_recovers(ctx):
    coeffs = list(range(1, 257))
    zp, op = z.fri.params_new(4, 8, 4, 8, 1), o.fri_params_new(4, 8, 4, 8, 1)
    _ffi.check(_ffi.lib().z0p_test_set_gate_spin(ctx.handle, 0))
    try:
        with pytest.raises(z.Z0pError) as e:
            z.fri.commit_open(coeffs, zp, z.Transcript(b"fgate"), ctx)
        assert e.value.code == 100
    finally:
        _ffi.check(_ffi.lib().z0p_test_set_gate_spin(ctx.handle, DEFAULT_SPIN))
    assert z.fri.commit_open(coeffs, zp, z.Transcript(b"fgate"), ctx) == \
        o.fri_commit_open(coeffs, op, o.Transcript(b"fgate"))
