"""FRI polynomial commitment (reference: src/fri.rs)."""
import ctypes

from ._ffi import FriParamsC, Z0pError, buf, check, lib, take, u32_buf, u8p

Params = FriParamsC


def _shape(self):
    return "B%dQ%dG%dC%dF%d" % (self.blowup, self.num_queries, self.grinding_bits, self.cap_height, self.final_size)


Params.shape = _shape


def params_new(blowup, num_queries, final_size, grinding_bits=0, cap_height=0) -> Params:
    """fri::Params::new (fri.rs:141-158) plus the builder setters."""
    p = Params()
    check(lib().z0p_fri_params_new(blowup, num_queries, final_size, ctypes.byref(p)))
    p.grinding_bits = grinding_bits
    p.cap_height = cap_height
    return p


def for_security(poly_len, target_bits=100) -> Params:
    """fri::Params::for_security (fri.rs:232-255)."""
    p = Params()
    check(lib().z0p_fri_params_for_security(poly_len, target_bits, ctypes.byref(p)))
    return p


Params.new = staticmethod(params_new)
Params.for_security = staticmethod(for_security)


def _indices_buf(params):
    n = int(params.num_queries)
    return (ctypes.c_uint64 * max(n, 1))(), n


def commit_open(coeffs, params: Params, transcript, ctx=None, return_indices=False):
    """fri::commit + public_commitment + open (fri.rs:408-433, 863-870,
    645-666) on the GPU; returns (Commitment::to_bytes, Proof::to_bytes), and
    the sampled query indices as a third item with return_indices."""
    return _commit_values(lib().z0p_fri_commit_open, coeffs, params, transcript, ctx, 1, None, return_indices)


def _commit_values(fn, values, params, transcript, ctx, per, degree_bound=None, return_indices=False):
    from . import default_context
    ctx = ctx or default_context()
    arr, ptr = u32_buf(values)
    if arr.size % per:
        raise ValueError("Ext4 values take 4 words each")
    cm, pf = u8p(), u8p()
    cn, pn = ctypes.c_size_t(), ctypes.c_size_t()
    idx, nq = _indices_buf(params)
    args = [ctx.handle, ptr, arr.size // per] + ([] if degree_bound is None else [degree_bound])
    check(fn(*args, ctypes.byref(params), transcript.handle, ctypes.byref(cm), ctypes.byref(cn), ctypes.byref(pf),
             ctypes.byref(pn), idx, nq))
    out = (take(cm, cn.value), take(pf, pn.value))
    return out + (list(idx[:nq]),) if return_indices else out


def commit_open_ext(coeffs4, params: Params, transcript, ctx=None, return_indices=False):
    """fri::commit_ext (fri.rs:444-461) + public_commitment + open on the GPU;
    Ext4 coefficients as 4 words each, coordinate 0 first."""
    return _commit_values(lib().z0p_fri_commit_open_ext, coeffs4, params, transcript, ctx, 4, None, return_indices)


def commit_open_evals(evals, degree_bound, params: Params, transcript, ctx=None, return_indices=False):
    """fri::commit_evals (fri.rs:473-500) + public_commitment + open on the
    GPU: base-field LDE values in natural order."""
    return _commit_values(lib().z0p_fri_commit_open_evals, evals, params, transcript, ctx, 1, degree_bound,
                          return_indices)


def commit_open_ext_evals(evals4, degree_bound, params: Params, transcript, ctx=None, return_indices=False):
    """fri::commit_ext_evals (fri.rs:510-549) + public_commitment + open on the
    GPU: Ext4 LDE values (4 words each) in natural order."""
    return _commit_values(lib().z0p_fri_commit_open_ext_evals, evals4, params, transcript, ctx, 4, degree_bound,
                          return_indices)


class ProverState:
    """fri::ProverState (fri.rs:276-294): committed layers and trees held in
    device memory, opened later with ``open`` (fri.rs:645-666)."""

    def __init__(self, handle, params, ctx):
        self._h = handle
        self._params = params
        self._ctx = ctx  # keeps the context alive while the state exists

    def commitment(self) -> bytes:
        """fri::public_commitment (fri.rs:863-870) + Commitment::to_bytes."""
        cm, cn = u8p(), ctypes.c_size_t()
        check(lib().z0p_fri_state_commitment(self._h, ctypes.byref(cm), ctypes.byref(cn)))
        return take(cm, cn.value)

    def open(self, transcript):
        """fri::open: (Proof::to_bytes, sampled query indices)."""
        pf, pn = u8p(), ctypes.c_size_t()
        idx, nq = _indices_buf(self._params)
        check(lib().z0p_fri_open(self._h, transcript.handle, ctypes.byref(pf), ctypes.byref(pn), idx, nq))
        return take(pf, pn.value), list(idx[:nq])

    def close(self):
        if getattr(self, "_h", None):
            lib().z0p_fri_state_free(self._h)
            self._h = None

    def __del__(self):
        self.close()


def _commit_state(fn, values, params, transcript, ctx, per, degree_bound=None):
    from . import default_context
    ctx = ctx or default_context()
    arr, ptr = u32_buf(values)
    if arr.size % per:
        raise ValueError("Ext4 values take 4 words each")
    h = ctypes.c_void_p()
    args = [ctx.handle, ptr, arr.size // per] + ([] if degree_bound is None else [degree_bound])
    check(fn(*args, ctypes.byref(params), transcript.handle, ctypes.byref(h)))
    return ProverState(h.value, params, ctx)


def commit(coeffs, params: Params, transcript, ctx=None) -> ProverState:
    """fri::commit (fri.rs:408-433) on the GPU."""
    return _commit_state(lib().z0p_fri_commit, coeffs, params, transcript, ctx, 1)


def commit_ext(coeffs4, params: Params, transcript, ctx=None) -> ProverState:
    """fri::commit_ext (fri.rs:444-461) on the GPU."""
    return _commit_state(lib().z0p_fri_commit_ext, coeffs4, params, transcript, ctx, 4)


def commit_evals(evals, degree_bound, params: Params, transcript, ctx=None) -> ProverState:
    """fri::commit_evals (fri.rs:473-500) on the GPU."""
    return _commit_state(lib().z0p_fri_commit_evals, evals, params, transcript, ctx, 1, degree_bound)


def commit_ext_evals(evals4, degree_bound, params: Params, transcript, ctx=None) -> ProverState:
    """fri::commit_ext_evals (fri.rs:510-549) on the GPU."""
    return _commit_state(lib().z0p_fri_commit_ext_evals, evals4, params, transcript, ctx, 4, degree_bound)


def verify(commitment: bytes, proof: bytes, params: Params, transcript):
    """Commitment/Proof::from_bytes + fri::verify (fri.rs:686-859); returns the
    sampled query indices, raises Z0pError on rejection."""
    idx, nq = _indices_buf(params)
    check(lib().z0p_fri_verify(buf(commitment), len(commitment), buf(proof), len(proof), ctypes.byref(params),
                               transcript.handle, idx, nq))
    return list(idx[:nq])


__all__ = ["Params", "params_new", "for_security", "commit_open", "commit_open_ext", "commit_open_evals",
           "commit_open_ext_evals", "ProverState", "commit", "commit_ext", "commit_evals", "commit_ext_evals",
           "verify", "Z0pError"]
