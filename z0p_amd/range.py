"""Range-check AIR prover/verifier (reference: src/range.rs): a public value
fits in ``n_bits`` bits, over a value column and its bit columns."""
import ctypes

from ._ffi import RangeParamsC, buf, check, lib, take, u8p

Params = RangeParamsC


def _params_repr(self):
    return "range.Params(n_rows=%d, n_bits=%d, B%dQ%dG%dC%dF%d, value=%d)" % (
        self.n_rows, self.n_bits, self.blowup, self.num_queries, self.grinding_bits, self.cap_height,
        self.final_size, self.value)


def _shape(self):
    return "B%dQ%dG%dC%dF%d" % (self.blowup, self.num_queries, self.grinding_bits, self.cap_height, self.final_size)


def _replace(self, **fields):
    """A copy with the given fields changed (the builders of range.rs:101-114)."""
    q = Params()
    ctypes.pointer(q)[0] = self
    for k, v in fields.items():
        setattr(q, k, v)
    return q


Params.__repr__ = _params_repr
Params.shape = _shape
Params.replace = _replace


def params_new(n_rows, n_bits, blowup, num_queries, final_size, value) -> Params:
    """range::Params::new (range.rs:59-98)."""
    p = Params()
    check(lib().z0p_range_params_new(n_rows, n_bits, blowup, num_queries, final_size, value, ctypes.byref(p)))
    return p


def for_security(n_rows, n_bits, value, target_bits=100) -> Params:
    """range::Params::for_security (range.rs:124-158)."""
    p = Params()
    check(lib().z0p_range_params_for_security(n_rows, n_bits, value, target_bits, ctypes.byref(p)))
    return p


Params.new = staticmethod(params_new)
Params.for_security = staticmethod(for_security)


def _ctx(ctx):
    from . import default_context
    return ctx or default_context()


def prove(params: Params, transcript, ctx=None) -> bytes:
    """range::prove + Proof::to_bytes (range.rs:334-342, 240-260) on the GPU."""
    out, n = u8p(), ctypes.c_size_t()
    check(lib().z0p_range_prove(_ctx(ctx).handle, ctypes.byref(params), transcript.handle, ctypes.byref(out),
                                ctypes.byref(n)))
    return take(out, n.value)


def prove_with_columns(params: Params, v_col, bit_cols, transcript, ctx=None) -> bytes:
    """range::prove_with_columns (range.rs:357-382) on the GPU: the value
    column and n_bits bit columns of n_rows canonical values each."""
    import numpy as np
    cols = [np.asarray(v_col, dtype=np.uint32)] + [np.asarray(c, dtype=np.uint32) for c in bit_cols]
    length = cols[0].size
    if any(c.size != length for c in cols):
        from ._ffi import Z0pError
        raise Z0pError(8)  # MalformedInput, as the reference's shape check
    flat = np.ascontiguousarray(np.concatenate(cols)) if cols else np.zeros(1, dtype=np.uint32)
    out, n = u8p(), ctypes.c_size_t()
    check(lib().z0p_range_prove_with_columns(_ctx(ctx).handle, ctypes.byref(params),
                                             flat.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(cols),
                                             length, transcript.handle, ctypes.byref(out), ctypes.byref(n)))
    return take(out, n.value)


def prove_zk(params: Params, n_rand: int, rng, transcript, ctx=None) -> bytes:
    """range::prove_zk (range.rs:401-428) on the GPU: every column masked with
    ``n_rand`` maskers drawn from ``rng`` (a ``z0p_amd.ChaCha20``, advanced)."""
    out, n = u8p(), ctypes.c_size_t()
    check(lib().z0p_range_prove_zk(_ctx(ctx).handle, ctypes.byref(params), n_rand, rng.handle, transcript.handle,
                                   ctypes.byref(out), ctypes.byref(n)))
    return take(out, n.value)


def verify(params: Params, proof: bytes, transcript) -> None:
    """Proof::from_bytes + range::verify (range.rs:280-306, 521-721); raises Z0pError."""
    check(lib().z0p_range_verify(ctypes.byref(params), transcript.handle, buf(proof), len(proof)))
