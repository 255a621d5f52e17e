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


def commit_open(coeffs, params: Params, transcript, ctx=None):
    """fri::commit + public_commitment + open (fri.rs:408-433, 863-870,
    645-666) on the GPU; returns (Commitment::to_bytes, Proof::to_bytes)."""
    from . import default_context
    ctx = ctx or default_context()
    arr, ptr = u32_buf(coeffs)
    cm, pf = u8p(), u8p()
    cn, pn = ctypes.c_size_t(), ctypes.c_size_t()
    check(lib().z0p_fri_commit_open(ctx.handle, ptr, arr.size, ctypes.byref(params), transcript.handle,
                                    ctypes.byref(cm), ctypes.byref(cn), ctypes.byref(pf), ctypes.byref(pn)))
    return take(cm, cn.value), take(pf, pn.value)


def _commit_values(fn, values, params, transcript, ctx, per, degree_bound=None):
    from . import default_context
    ctx = ctx or default_context()
    arr, ptr = u32_buf(values)
    if arr.size % per:
        raise ValueError("Ext4 values take 4 words each")
    cm, pf = u8p(), u8p()
    cn, pn = ctypes.c_size_t(), ctypes.c_size_t()
    args = [ctx.handle, ptr, arr.size // per] + ([] if degree_bound is None else [degree_bound])
    check(fn(*args, ctypes.byref(params), transcript.handle, ctypes.byref(cm), ctypes.byref(cn), ctypes.byref(pf),
             ctypes.byref(pn)))
    return take(cm, cn.value), take(pf, pn.value)


def commit_open_ext(coeffs4, params: Params, transcript, ctx=None):
    """fri::commit_ext (fri.rs:444-461) + public_commitment + open on the GPU;
    Ext4 coefficients as 4 words each, coordinate 0 first."""
    return _commit_values(lib().z0p_fri_commit_open_ext, coeffs4, params, transcript, ctx, 4)


def commit_open_evals(evals, degree_bound, params: Params, transcript, ctx=None):
    """fri::commit_evals (fri.rs:473-500) + public_commitment + open on the
    GPU: base-field LDE values in natural order."""
    return _commit_values(lib().z0p_fri_commit_open_evals, evals, params, transcript, ctx, 1, degree_bound)


def commit_open_ext_evals(evals4, degree_bound, params: Params, transcript, ctx=None):
    """fri::commit_ext_evals (fri.rs:510-549) + public_commitment + open on the
    GPU: Ext4 LDE values (4 words each) in natural order."""
    return _commit_values(lib().z0p_fri_commit_open_ext_evals, evals4, params, transcript, ctx, 4, degree_bound)


def verify(commitment: bytes, proof: bytes, params: Params, transcript) -> None:
    """Commitment/Proof::from_bytes + fri::verify (fri.rs:686-859); raises Z0pError."""
    check(lib().z0p_fri_verify(buf(commitment), len(commitment), buf(proof), len(proof), ctypes.byref(params),
                               transcript.handle))


__all__ = ["Params", "params_new", "for_security", "commit_open", "commit_open_ext", "commit_open_evals",
           "commit_open_ext_evals", "verify", "Z0pError"]
