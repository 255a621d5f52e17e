"""Fibonacci AIR prover/verifier (reference: src/air.rs)."""
import ctypes

from ._ffi import AirParamsC, Z0pError, buf, check, lib, take, u8p, view

Params = AirParamsC


def _params_repr(self):
    return "air.Params(n_steps=%d, B%dQ%dG%dC%dF%d, a0=%d, b0=%d, b_last=%d)" % (
        self.n_steps, self.blowup, self.num_queries, self.grinding_bits, self.cap_height, self.final_size,
        self.a0, self.b0, self.b_last)


def _shape(self):
    return "B%dQ%dG%dC%dF%d" % (self.blowup, self.num_queries, self.grinding_bits, self.cap_height, self.final_size)


Params.__repr__ = _params_repr
Params.shape = _shape


def params_new(n_steps, blowup, num_queries, final_size, a0, b0, b_last) -> Params:
    """air::Params::new (air.rs:63-101)."""
    p = Params()
    check(lib().z0p_air_params_new(n_steps, blowup, num_queries, final_size, a0, b0, b_last, ctypes.byref(p)))
    return p


def for_security(n_steps, a0, b0, b_last, target_bits=100) -> Params:
    """air::Params::for_security (air.rs:129-162)."""
    p = Params()
    check(lib().z0p_air_params_for_security(n_steps, a0, b0, b_last, target_bits, ctypes.byref(p)))
    return p


Params.new = staticmethod(params_new)
Params.for_security = staticmethod(for_security)


def b_last(a0, b0, n_steps) -> int:
    """build_trace(a0, b0, n).1[n - 1] (air.rs:341-354)."""
    return lib().z0p_air_b_last(a0, b0, n_steps)


def prove(params: Params, transcript, ctx=None) -> bytes:
    """air::prove + Proof::to_bytes (air.rs:365-372, 262-281) on the GPU."""
    from . import default_context
    ctx = ctx or default_context()
    out = u8p()
    n = ctypes.c_size_t()
    check(lib().z0p_air_prove(ctx.handle, ctypes.byref(params), transcript.handle, ctypes.byref(out),
                              ctypes.byref(n)))
    return take(out, n.value)


def _u32_column(values):
    import numpy as np
    arr = np.ascontiguousarray(np.asarray(values, dtype=np.uint32))
    return arr, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def prove_with_trace(params: Params, col_a, col_b, transcript, ctx=None) -> bytes:
    """air::prove_with_trace (air.rs:385-406) on the GPU: explicit columns of
    n_steps canonical field values each, both interpolated and extended."""
    from . import default_context
    ctx = ctx or default_context()
    a, pa = _u32_column(col_a)
    b, pb = _u32_column(col_b)
    out = u8p()
    n = ctypes.c_size_t()
    check(lib().z0p_air_prove_with_trace(ctx.handle, ctypes.byref(params), pa, a.size, pb, b.size,
                                         transcript.handle, ctypes.byref(out), ctypes.byref(n)))
    return take(out, n.value)


def prove_zk(params: Params, n_rand: int, rng, transcript, ctx=None) -> bytes:
    """air::prove_zk (air.rs:439-464) on the GPU: both columns masked with
    ``n_rand`` maskers drawn from ``rng`` (a ``z0p_amd.ChaCha20``, advanced)."""
    from . import default_context
    ctx = ctx or default_context()
    out = u8p()
    n = ctypes.c_size_t()
    check(lib().z0p_air_prove_zk(ctx.handle, ctypes.byref(params), n_rand, rng.handle, transcript.handle,
                                 ctypes.byref(out), ctypes.byref(n)))
    return take(out, n.value)


def prove_batch(params_list, transcripts, ctx=None, zero_copy=False):
    """``air.prove`` for every (params, transcript) pair in one batch: the
    instances share every kernel launch and are admitted as device memory
    allows. Returns a list of proof bytes, or of Z0pError for instances that
    failed; each transcript is advanced as its single call would advance it.
    With ``zero_copy`` the proofs are read-only memoryviews of the buffers the
    C ABI hands over (no copy into ``bytes``).
    """
    from . import default_context
    ctx = ctx or default_context()
    count = len(params_list)
    if len(transcripts) != count:
        raise ValueError("one transcript per instance")
    ps = (Params * max(count, 1))(*params_list)
    ts = (ctypes.c_void_p * max(count, 1))(*[t.handle for t in transcripts])
    outs = (u8p * max(count, 1))()
    lens = (ctypes.c_size_t * max(count, 1))()
    sts = (ctypes.c_int * max(count, 1))()
    check(lib().z0p_air_prove_batch(ctx.handle, ps, ts, count, outs, lens, sts))
    result = []
    for i in range(count):
        result.append((view if zero_copy else take)(outs[i], lens[i]) if sts[i] == 0 else Z0pError(sts[i]))
    return result


def verify(params: Params, proof: bytes, transcript) -> None:
    """Proof::from_bytes + air::verify (air.rs:289-327, 607-752); raises Z0pError."""
    check(lib().z0p_air_verify(ctypes.byref(params), transcript.handle, buf(proof), len(proof)))


def _batch_args(params_list, proofs, transcripts):
    count = len(params_list)
    if len(proofs) != count or len(transcripts) != count:
        raise ValueError("one proof and one transcript per instance")
    ps = (Params * max(count, 1))(*params_list)
    ts = (ctypes.c_void_p * max(count, 1))(*[t.handle for t in transcripts])
    # Pointers into the bytes objects themselves (immutable, kept alive by
    # `keep`): no copy of the proofs.
    keep = [bytes(b) if b else b"\0" for b in proofs]
    pp = (u8p * max(count, 1))(*[ctypes.cast(ctypes.c_char_p(k), u8p) for k in keep])
    lens = (ctypes.c_size_t * max(count, 1))(*[len(b) for b in proofs])
    sts = (ctypes.c_int * max(count, 1))()
    return count, ps, ts, keep, pp, lens, sts


def verify_batch_device(params_list, proofs, transcripts, ctx=None, threads=0):
    """air::verify for every instance, the Merkle path hashing batched on the
    GPU of ctx and the rest on host threads; returns a list of None
    (accepted) or Z0pError per instance, as verify_batch does."""
    from . import default_context
    ctx = ctx or default_context()
    count, ps, ts, keep, pp, lens, sts = _batch_args(params_list, proofs, transcripts)
    check(lib().z0p_air_verify_batch_device(ctx.handle, ps, ts, pp, lens, count, threads, sts))
    return [None if sts[i] == 0 else Z0pError(sts[i]) for i in range(count)]


def verify_batch(params_list, proofs, transcripts, threads=0):
    """air::verify for every (params, proof, transcript) on host threads;
    returns a list of None (accepted) or Z0pError per instance."""
    count, ps, ts, keep, pp, lens, sts = _batch_args(params_list, proofs, transcripts)
    check(lib().z0p_air_verify_batch(ps, ts, pp, lens, count, threads, sts))
    return [None if sts[i] == 0 else Z0pError(sts[i]) for i in range(count)]


__all__ = ["Params", "params_new", "for_security", "b_last", "prove", "prove_with_trace", "prove_zk", "prove_batch",
           "verify", "verify_batch", "verify_batch_device", "Z0pError"]
