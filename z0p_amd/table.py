"""Table-membership AIR prover/verifier (reference: src/table.rs): a public
value belongs to a public table, over one committed value column."""
import ctypes

from ._ffi import TableParamsC, buf, check, lib, take, u8p

Params = TableParamsC


def _params_repr(self):
    return "table.Params(n_rows=%d, B%dQ%dG%dC%dF%d, value=%d)" % (
        self.n_rows, self.blowup, self.num_queries, self.grinding_bits, self.cap_height, self.final_size,
        self.value)


def _shape(self):
    return "B%dQ%dG%dC%dF%d" % (self.blowup, self.num_queries, self.grinding_bits, self.cap_height, self.final_size)


def _replace(self, **fields):
    """A copy with the given fields changed (the builders of table.rs:89-101)."""
    q = Params()
    ctypes.pointer(q)[0] = self
    for k, v in fields.items():
        setattr(q, k, v)
    return q


Params.__repr__ = _params_repr
Params.shape = _shape
Params.replace = _replace


def params_new(n_rows, blowup, num_queries, final_size, value) -> Params:
    """table::Params::new (table.rs:53-87)."""
    p = Params()
    check(lib().z0p_table_params_new(n_rows, blowup, num_queries, final_size, value, ctypes.byref(p)))
    return p


def for_security(n_rows, table_len, value, target_bits=100) -> Params:
    """table::Params::for_security (table.rs:114-168)."""
    p = Params()
    check(lib().z0p_table_params_for_security(n_rows, table_len, value, target_bits, ctypes.byref(p)))
    return p


Params.new = staticmethod(params_new)
Params.for_security = staticmethod(for_security)


def membership_bound(table_len, n_rows, n_rand=0) -> int:
    """table::membership_bound (table.rs:357-363)."""
    return lib().z0p_table_membership_bound(table_len, n_rows, n_rand)


def _table(table):
    import numpy as np
    arr = np.ascontiguousarray(np.asarray(table, dtype=np.uint64) % 0x78000001).astype(np.uint32)
    if arr.size == 0:
        arr = np.zeros(1, dtype=np.uint32)
        return arr, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), 0
    return arr, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), arr.size


def _ctx(ctx):
    from . import default_context
    return ctx or default_context()


def prove(params: Params, table, transcript, ctx=None) -> bytes:
    """table::prove + Proof::to_bytes (table.rs:305-318, 244-262) on the GPU."""
    arr, ptr, n_t = _table(table)
    out, n = u8p(), ctypes.c_size_t()
    check(lib().z0p_table_prove(_ctx(ctx).handle, ctypes.byref(params), ptr, n_t, transcript.handle,
                                ctypes.byref(out), ctypes.byref(n)))
    return take(out, n.value)


def prove_zk(params: Params, table, n_rand: int, rng, transcript, ctx=None) -> bytes:
    """table::prove_zk (table.rs:329-355) on the GPU: the column masked with
    ``n_rand`` maskers drawn from ``rng`` (a ``z0p_amd.ChaCha20``, advanced)."""
    arr, ptr, n_t = _table(table)
    out, n = u8p(), ctypes.c_size_t()
    check(lib().z0p_table_prove_zk(_ctx(ctx).handle, ctypes.byref(params), ptr, n_t, n_rand, rng.handle,
                                   transcript.handle, ctypes.byref(out), ctypes.byref(n)))
    return take(out, n.value)


def verify(params: Params, table, proof: bytes, transcript) -> None:
    """Proof::from_bytes + table::verify (table.rs:265-288, 430-538); raises Z0pError."""
    arr, ptr, n_t = _table(table)
    check(lib().z0p_table_verify(ctypes.byref(params), ptr, n_t, transcript.handle, buf(proof), len(proof)))
