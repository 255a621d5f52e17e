"""Hash-based folding / IVC accumulation over Fibonacci segments (reference:
src/fold.rs): every segment is committed on the GPU as it is added; one
batched FRI over the gamma-combination of their compositions finalizes."""
import ctypes

from ._ffi import FoldParamsC, buf, check, lib, take, u8p

Params = FoldParamsC


def _params_repr(self):
    return "fold.Params(seg_steps=%d, B%dQ%dG%dC%dF%d, in=(%d, %d), out_b=%d)" % (
        self.seg_steps, self.blowup, self.num_queries, self.grinding_bits, self.cap_height, self.final_size,
        self.in_a, self.in_b, self.out_b)


def _shape(self):
    return "B%dQ%dG%dC%dF%d" % (self.blowup, self.num_queries, self.grinding_bits, self.cap_height, self.final_size)


def _replace(self, **fields):
    """A copy with the given fields changed (the builders of fold.rs:113-125)."""
    q = Params()
    ctypes.pointer(q)[0] = self
    for k, v in fields.items():
        setattr(q, k, v)
    return q


Params.__repr__ = _params_repr
Params.shape = _shape
Params.replace = _replace


def params_new(seg_steps, blowup, num_queries, final_size, in_a, in_b, out_b) -> Params:
    """fold::Params::new (fold.rs:72-111)."""
    p = Params()
    check(lib().z0p_fold_params_new(seg_steps, blowup, num_queries, final_size, in_a, in_b, out_b, ctypes.byref(p)))
    return p


def for_security(seg_steps, in_a, in_b, out_b, target_bits=100) -> Params:
    """fold::Params::for_security (fold.rs:136-170)."""
    p = Params()
    check(lib().z0p_fold_params_for_security(seg_steps, in_a, in_b, out_b, target_bits, ctypes.byref(p)))
    return p


Params.new = staticmethod(params_new)
Params.for_security = staticmethod(for_security)


def _col(values):
    import numpy as np
    arr = np.ascontiguousarray(np.asarray(values, dtype=np.uint32))
    if arr.size == 0:
        arr = np.zeros(1, dtype=np.uint32)
        return arr, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), 0
    return arr, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), arr.size


class Accumulator:
    """fold::Accumulator (fold.rs:361-570) on the GPU. Segments stay in device
    memory on the context until ``finalize``, which consumes them."""

    def __init__(self, params: Params, ctx=None):
        from . import default_context
        self._ctx = ctx or default_context()
        h = ctypes.c_void_p()
        check(lib().z0p_fold_acc_new(self._ctx.handle, ctypes.byref(params), ctypes.byref(h)))
        self._h = h.value

    def close(self):
        if getattr(self, "_h", None):
            lib().z0p_fold_acc_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __len__(self):
        return lib().z0p_fold_acc_len(self._h)

    def is_empty(self):
        return len(self) == 0

    def add_segment(self, col_a, col_b, transcript):
        """Accumulator::add_segment (fold.rs:407-492)."""
        a, pa, na = _col(col_a)
        b, pb, nb = _col(col_b)
        check(lib().z0p_fold_acc_add_segment(self._h, pa, na, pb, nb, transcript.handle))

    def finalize(self, transcript) -> bytes:
        """Accumulator::finalize + IvcProof::to_bytes (fold.rs:494-570, 281-310)."""
        out, n = u8p(), ctypes.c_size_t()
        check(lib().z0p_fold_acc_finalize(self._h, transcript.handle, ctypes.byref(out), ctypes.byref(n)))
        return take(out, n.value)


def verify(params: Params, proof: bytes, transcript) -> None:
    """IvcProof::from_bytes + fold::verify (fold.rs:312-359, 572-667); raises Z0pError."""
    check(lib().z0p_fold_verify(ctypes.byref(params), transcript.handle, buf(proof), len(proof)))
