"""ctypes bindings to the oracle (the CPU restatement of the reference
prover under ``oracle/``). Test infrastructure: only tests, ``smoke()`` and the
benchmark's CPU-baseline leg load it, always as the checker."""
import ctypes
import os
import pathlib

ROOT = pathlib.Path(__file__).resolve().parent.parent
LIB_PATH = ROOT / "oracle" / "build" / "liboracle.so"

u8p = ctypes.POINTER(ctypes.c_uint8)
u32p = ctypes.POINTER(ctypes.c_uint32)
u64p = ctypes.POINTER(ctypes.c_uint64)


class AirParams(ctypes.Structure):
    """Layout shared with ``z0p_air_params`` (include/z0p.h)."""

    _fields_ = [
        ("n_steps", ctypes.c_uint64),
        ("blowup", ctypes.c_uint64),
        ("num_queries", ctypes.c_uint64),
        ("final_size", ctypes.c_uint64),
        ("grinding_bits", ctypes.c_uint32),
        ("reserved0", ctypes.c_uint32),
        ("cap_height", ctypes.c_uint64),
        ("a0", ctypes.c_uint32),
        ("b0", ctypes.c_uint32),
        ("b_last", ctypes.c_uint32),
        ("reserved1", ctypes.c_uint32),
    ]

    def shape(self):
        return "B%dQ%dG%dC%dF%d" % (self.blowup, self.num_queries, self.grinding_bits, self.cap_height,
                                     self.final_size)


class FriParams(ctypes.Structure):
    """Layout shared with ``z0p_fri_params`` (include/z0p.h)."""

    _fields_ = [
        ("blowup", ctypes.c_uint64),
        ("num_queries", ctypes.c_uint64),
        ("final_size", ctypes.c_uint64),
        ("grinding_bits", ctypes.c_uint32),
        ("reserved0", ctypes.c_uint32),
        ("cap_height", ctypes.c_uint64),
    ]

    def shape(self):
        return "B%dQ%dG%dC%dF%d" % (self.blowup, self.num_queries, self.grinding_bits, self.cap_height,
                                     self.final_size)


class TableParams(ctypes.Structure):
    """Layout shared with ``z0p_table_params`` (include/z0p.h)."""

    _fields_ = [
        ("n_rows", ctypes.c_uint64),
        ("blowup", ctypes.c_uint64),
        ("num_queries", ctypes.c_uint64),
        ("final_size", ctypes.c_uint64),
        ("grinding_bits", ctypes.c_uint32),
        ("reserved0", ctypes.c_uint32),
        ("cap_height", ctypes.c_uint64),
        ("value", ctypes.c_uint32),
        ("reserved1", ctypes.c_uint32),
    ]

    def replace(self, **kw):
        q = TableParams()
        ctypes.pointer(q)[0] = self
        for k, v in kw.items():
            setattr(q, k, v)
        return q


class FoldParams(ctypes.Structure):
    """Layout shared with ``z0p_fold_params`` (include/z0p.h)."""

    _fields_ = [
        ("seg_steps", ctypes.c_uint64),
        ("blowup", ctypes.c_uint64),
        ("num_queries", ctypes.c_uint64),
        ("final_size", ctypes.c_uint64),
        ("grinding_bits", ctypes.c_uint32),
        ("reserved0", ctypes.c_uint32),
        ("cap_height", ctypes.c_uint64),
        ("in_a", ctypes.c_uint32),
        ("in_b", ctypes.c_uint32),
        ("out_b", ctypes.c_uint32),
        ("reserved1", ctypes.c_uint32),
    ]

    def replace(self, **kw):
        q = FoldParams()
        ctypes.pointer(q)[0] = self
        for k, v in kw.items():
            setattr(q, k, v)
        return q


class RangeParams(ctypes.Structure):
    """Layout shared with ``z0p_range_params`` (include/z0p.h)."""

    _fields_ = [
        ("n_rows", ctypes.c_uint64),
        ("n_bits", ctypes.c_uint64),
        ("blowup", ctypes.c_uint64),
        ("num_queries", ctypes.c_uint64),
        ("final_size", ctypes.c_uint64),
        ("grinding_bits", ctypes.c_uint32),
        ("reserved0", ctypes.c_uint32),
        ("cap_height", ctypes.c_uint64),
        ("value", ctypes.c_uint32),
        ("reserved1", ctypes.c_uint32),
    ]

    def shape(self):
        return "B%dQ%dG%dC%dF%d" % (self.blowup, self.num_queries, self.grinding_bits, self.cap_height,
                                     self.final_size)

    def replace(self, **kw):
        q = RangeParams()
        ctypes.pointer(q)[0] = self
        for k, v in kw.items():
            setattr(q, k, v)
        return q


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError("oracle library missing: build with `make -C oracle` (or __graft_entry__.build())")
    L = ctypes.CDLL(str(LIB_PATH))
    sz = ctypes.c_size_t
    vp = ctypes.c_void_p
    sig = {
        "oracle_set_threads": (None, [sz]),
        "oracle_get_threads": (sz, []),
        "oracle_free": (None, [vp]),
        "oracle_blake3": (None, [u8p, sz, u8p]),
        "oracle_blake3_small": (None, [u8p, sz, u8p]),
        "oracle_blake3_leaf64": (None, [u8p, u8p]),
        "oracle_blake3_node": (None, [u8p, u8p, u8p]),
        "oracle_blake3_split": (None, [u8p, sz, ctypes.POINTER(sz), sz, u8p]),
        "oracle_field_mul": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32]),
        "oracle_field_inv": (ctypes.c_uint32, [ctypes.c_uint32]),
        "oracle_field_pow": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint64]),
        "oracle_field_halve": (ctypes.c_uint32, [ctypes.c_uint32]),
        "oracle_two_adic_generator": (ctypes.c_int, [ctypes.c_uint32, u32p]),
        "oracle_ext4_mul": (None, [u32p, u32p, u32p]),
        "oracle_ext4_inv": (None, [u32p, u32p]),
        "oracle_ext4_pow": (None, [u32p, ctypes.c_uint64, ctypes.c_uint64, u32p]),
        "oracle_ntt_forward": (ctypes.c_int, [u32p, sz]),
        "oracle_ntt_inverse": (ctypes.c_int, [u32p, sz]),
        "oracle_ntt_evaluate": (ctypes.c_int, [u32p, sz, ctypes.c_uint32, u32p]),
        "oracle_evaluate_on_coset": (ctypes.c_int, [u32p, sz, sz, u32p]),
        "oracle_merkle_flat64": (ctypes.c_int, [u8p, sz, u8p]),
        "oracle_fold_values": (None, [u32p, ctypes.c_uint32, u32p, ctypes.c_uint32, u32p]),
        "oracle_point_at": (ctypes.c_int, [sz, sz, sz, u32p]),
        "oracle_linked_lde": (ctypes.c_int, [u32p, sz, sz, ctypes.c_uint32, ctypes.c_uint32, u32p]),
        "oracle_compose": (ctypes.c_int, [u32p, u32p, sz, sz, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                          u32p, ctypes.c_int, u32p]),
        "oracle_transcript_new": (vp, [u8p, sz]),
        "oracle_transcript_clone": (vp, [vp]),
        "oracle_transcript_free": (None, [vp]),
        "oracle_transcript_absorb": (None, [vp, ctypes.c_char_p, u8p, sz]),
        "oracle_transcript_challenge_bytes": (None, [vp, ctypes.c_char_p, u8p]),
        "oracle_transcript_challenge_ext": (None, [vp, ctypes.c_char_p, u32p]),
        "oracle_transcript_challenge_index": (ctypes.c_int, [vp, ctypes.c_char_p, sz, ctypes.POINTER(sz)]),
        "oracle_transcript_grind": (ctypes.c_uint64, [vp, ctypes.c_char_p, ctypes.c_uint32]),
        "oracle_transcript_verify_grind": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint32]),
        "oracle_transcript_history_digest": (None, [vp, u8p]),
        "oracle_grind_head": (ctypes.c_uint64, [u8p, sz, ctypes.c_uint32]),
        "oracle_air_params_for_security": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                                          ctypes.c_uint32, ctypes.c_uint32,
                                                          ctypes.POINTER(AirParams)]),
        "oracle_air_params_new": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                                 ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                                 ctypes.POINTER(AirParams)]),
        "oracle_fri_params_for_security": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint32,
                                                          ctypes.POINTER(FriParams)]),
        "oracle_fib_b_last": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64]),
        "oracle_build_trace": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, u32p, u32p]),
        "oracle_air_prove": (ctypes.c_int, [ctypes.POINTER(AirParams), vp, ctypes.POINTER(u8p), ctypes.POINTER(sz)]),
        "oracle_air_prove_with_trace": (ctypes.c_int, [ctypes.POINTER(AirParams), u32p, u32p, vp,
                                                       ctypes.POINTER(u8p), ctypes.POINTER(sz)]),
        "oracle_air_prove_zk": (ctypes.c_int, [ctypes.POINTER(AirParams), sz, vp, vp, ctypes.POINTER(u8p),
                                               ctypes.POINTER(sz)]),
        "oracle_chacha20_new": (vp, [u8p, u8p]),
        "oracle_chacha20_free": (None, [vp]),
        "oracle_chacha20_fill": (None, [vp, u8p, sz]),
        "oracle_chacha20_next_u32": (ctypes.c_uint32, [vp]),
        "oracle_chacha20_next_u64": (ctypes.c_uint64, [vp]),
        "oracle_chacha20_next_field": (ctypes.c_uint32, [vp]),
        "oracle_chacha20_block": (None, [u8p, ctypes.c_uint32, u8p, u8p]),
        "oracle_chacha20_quarter": (None, [u32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "oracle_mask_coeffs": (ctypes.c_int, [u32p, sz, sz, sz, vp, u32p, sz, ctypes.POINTER(sz)]),
        "oracle_air_verify": (ctypes.c_int, [ctypes.POINTER(AirParams), vp, u8p, sz]),
        "oracle_fri_commit_open": (ctypes.c_int, [u32p, sz, ctypes.POINTER(FriParams), vp, ctypes.POINTER(u8p),
                                                  ctypes.POINTER(sz), ctypes.POINTER(u8p), ctypes.POINTER(sz), u64p]),
        "oracle_fri_verify": (ctypes.c_int, [u8p, sz, u8p, sz, ctypes.POINTER(FriParams), vp, u64p]),
        "oracle_fri_commit_open_kind": (ctypes.c_int, [ctypes.c_int, u32p, sz, sz, ctypes.POINTER(FriParams), vp,
                                                       ctypes.POINTER(u8p), ctypes.POINTER(sz), ctypes.POINTER(u8p),
                                                       ctypes.POINTER(sz), u64p]),
        "oracle_fri_commit_state": (ctypes.c_int, [ctypes.c_int, u32p, sz, sz, ctypes.POINTER(FriParams), vp,
                                                   ctypes.POINTER(vp)]),
        "oracle_fri_state_commitment": (ctypes.c_int, [vp, ctypes.POINTER(u8p), ctypes.POINTER(sz)]),
        "oracle_fri_state_open": (ctypes.c_int, [vp, vp, ctypes.POINTER(u8p), ctypes.POINTER(sz), u64p]),
        "oracle_fri_state_free": (None, [vp]),
        "oracle_evaluate_on_coset_ext": (ctypes.c_int, [u32p, sz, sz, u32p]),
        "oracle_range_params_new": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                                   ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(RangeParams)]),
        "oracle_range_params_for_security": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                                                            ctypes.c_uint32, ctypes.POINTER(RangeParams)]),
        "oracle_range_prove": (ctypes.c_int, [ctypes.POINTER(RangeParams), vp, ctypes.POINTER(u8p),
                                              ctypes.POINTER(sz)]),
        "oracle_range_prove_with_columns": (ctypes.c_int, [ctypes.POINTER(RangeParams), u32p, sz, sz, vp,
                                                           ctypes.POINTER(u8p), ctypes.POINTER(sz)]),
        "oracle_range_prove_coeffs": (ctypes.c_int, [ctypes.POINTER(RangeParams), u32p, sz, sz, sz, vp,
                                                     ctypes.POINTER(u8p), ctypes.POINTER(sz)]),
        "oracle_range_prove_zk": (ctypes.c_int, [ctypes.POINTER(RangeParams), sz, vp, vp, ctypes.POINTER(u8p),
                                                 ctypes.POINTER(sz)]),
        "oracle_range_verify": (ctypes.c_int, [ctypes.POINTER(RangeParams), vp, u8p, sz]),
        "oracle_range_composition_at": (ctypes.c_int, [ctypes.c_uint32, u32p, sz, ctypes.c_uint32, ctypes.c_uint32,
                                                       ctypes.c_uint64, u32p, u32p]),
        "oracle_table_params_new": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                                   ctypes.c_uint32, ctypes.POINTER(TableParams)]),
        "oracle_table_params_for_security": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                                                            ctypes.c_uint32, ctypes.POINTER(TableParams)]),
        "oracle_table_membership_bound": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]),
        "oracle_table_prove": (ctypes.c_int, [ctypes.POINTER(TableParams), u32p, sz, vp, ctypes.POINTER(u8p),
                                              ctypes.POINTER(sz)]),
        "oracle_table_prove_coeffs": (ctypes.c_int, [ctypes.POINTER(TableParams), u32p, sz, u32p, sz, sz, vp,
                                                     ctypes.POINTER(u8p), ctypes.POINTER(sz)]),
        "oracle_table_prove_zk": (ctypes.c_int, [ctypes.POINTER(TableParams), u32p, sz, sz, vp, vp,
                                                 ctypes.POINTER(u8p), ctypes.POINTER(sz)]),
        "oracle_table_verify": (ctypes.c_int, [ctypes.POINTER(TableParams), u32p, sz, vp, u8p, sz]),
        "oracle_fold_params_new": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                                  ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                                  ctypes.POINTER(FoldParams)]),
        "oracle_fold_params_for_security": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                                           ctypes.c_uint32, ctypes.c_uint32,
                                                           ctypes.POINTER(FoldParams)]),
        "oracle_fold_acc_new": (vp, [ctypes.POINTER(FoldParams)]),
        "oracle_fold_acc_free": (None, [vp]),
        "oracle_fold_acc_len": (sz, [vp]),
        "oracle_fold_acc_set_params": (None, [vp, ctypes.POINTER(FoldParams)]),
        "oracle_fold_acc_add_segment": (ctypes.c_int, [vp, u32p, sz, u32p, sz, vp]),
        "oracle_fold_acc_finalize": (ctypes.c_int, [vp, vp, ctypes.POINTER(u8p), ctypes.POINTER(sz)]),
        "oracle_fold_acc_finalize_forged_out_a": (ctypes.c_int, [vp, sz, ctypes.c_uint32, u8p, sz,
                                                                 ctypes.POINTER(u8p), ctypes.POINTER(sz)]),
        "oracle_fold_verify": (ctypes.c_int, [ctypes.POINTER(FoldParams), vp, u8p, sz]),
        "oracle_stage_timers_clear": (None, []),
        "oracle_stage_timers": (sz, [ctypes.c_char_p, sz]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _buf(data: bytes):
    return (ctypes.c_uint8 * max(len(data), 1)).from_buffer_copy(data if data else b"\0")


def _take(ptr, n):
    out = ctypes.string_at(ptr, n)
    lib().oracle_free(ptr)
    return out


def u32_array(values):
    arr = (ctypes.c_uint32 * max(len(values), 1))()
    for i, v in enumerate(values):
        arr[i] = v
    return arr


def set_threads(n):
    lib().oracle_set_threads(n)


def blake3(data: bytes) -> bytes:
    out = (ctypes.c_uint8 * 32)()
    lib().oracle_blake3(_buf(data), len(data), out)
    return bytes(out)


def blake3_split(data: bytes, splits) -> bytes:
    out = (ctypes.c_uint8 * 32)()
    arr = (ctypes.c_size_t * max(len(splits), 1))(*splits)
    lib().oracle_blake3_split(_buf(data), len(data), arr, len(splits), out)
    return bytes(out)


def blake3_small(data: bytes) -> bytes:
    out = (ctypes.c_uint8 * 32)()
    lib().oracle_blake3_small(_buf(data), len(data), out)
    return bytes(out)


def leaf64(block: bytes) -> bytes:
    assert len(block) == 64
    out = (ctypes.c_uint8 * 32)()
    lib().oracle_blake3_leaf64(_buf(block), out)
    return bytes(out)


def node(left: bytes, right: bytes) -> bytes:
    out = (ctypes.c_uint8 * 32)()
    lib().oracle_blake3_node(_buf(left), _buf(right), out)
    return bytes(out)


class Transcript:
    """Oracle transcript (transcript.rs:54-171)."""

    def __init__(self, label: bytes, _handle=None):
        self.h = _handle if _handle is not None else lib().oracle_transcript_new(_buf(label), len(label))

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_transcript_free(self.h)
            self.h = None

    def clone(self):
        return Transcript(b"", lib().oracle_transcript_clone(self.h))

    def absorb(self, label: bytes, data: bytes):
        lib().oracle_transcript_absorb(self.h, label, _buf(data), len(data))

    def challenge_bytes(self, label: bytes) -> bytes:
        out = (ctypes.c_uint8 * 32)()
        lib().oracle_transcript_challenge_bytes(self.h, label, out)
        return bytes(out)

    def challenge_ext(self, label: bytes):
        out = (ctypes.c_uint32 * 4)()
        lib().oracle_transcript_challenge_ext(self.h, label, out)
        return list(out)

    def challenge_index(self, label: bytes, domain: int):
        out = ctypes.c_size_t()
        rc = lib().oracle_transcript_challenge_index(self.h, label, domain, ctypes.byref(out))
        return None if rc else out.value

    def grind(self, label: bytes, bits: int) -> int:
        return lib().oracle_transcript_grind(self.h, label, bits)

    def verify_grind(self, label: bytes, nonce: int, bits: int) -> bool:
        return bool(lib().oracle_transcript_verify_grind(self.h, label, nonce, bits))

    def history_digest(self) -> bytes:
        out = (ctypes.c_uint8 * 32)()
        lib().oracle_transcript_history_digest(self.h, out)
        return bytes(out)


def air_params_for_security(n, a0, b0, b_last, bits=100):
    p = AirParams()
    rc = lib().oracle_air_params_for_security(n, a0, b0, b_last, bits, ctypes.byref(p))
    if rc:
        raise ValueError("oracle params error %d" % rc)
    return p


def air_params_new(n, blowup, nq, final_size, a0, b0, b_last):
    p = AirParams()
    rc = lib().oracle_air_params_new(n, blowup, nq, final_size, a0, b0, b_last, ctypes.byref(p))
    if rc:
        raise ValueError("oracle params error %d" % rc)
    return p


def fri_params_for_security(poly_len, bits=100):
    p = FriParams()
    rc = lib().oracle_fri_params_for_security(poly_len, bits, ctypes.byref(p))
    if rc:
        raise ValueError("oracle params error %d" % rc)
    return p


def fib_b_last(a0, b0, n):
    return lib().oracle_fib_b_last(a0, b0, n)


def air_prove(params: AirParams, transcript: Transcript) -> bytes:
    out = u8p()
    n = ctypes.c_size_t()
    rc = lib().oracle_air_prove(ctypes.byref(params), transcript.h, ctypes.byref(out), ctypes.byref(n))
    if rc:
        raise ValueError("oracle prove error %d" % rc)
    return _take(out, n.value)


def air_prove_with_trace(params: AirParams, col_a, col_b, transcript: Transcript) -> bytes:
    out = u8p()
    n = ctypes.c_size_t()
    rc = lib().oracle_air_prove_with_trace(ctypes.byref(params), u32_array(col_a), u32_array(col_b), transcript.h,
                                           ctypes.byref(out), ctypes.byref(n))
    if rc:
        raise ValueError("oracle prove error %d" % rc)
    return _take(out, n.value)


class OracleError(ValueError):
    """An oracle entry point returned a nonzero status (src/error.rs order)."""

    def __init__(self, code):
        super().__init__("oracle error %d" % code)
        self.code = code


class ChaCha20:
    """Oracle ChaCha20 stream (rng.rs:81-177)."""

    def __init__(self, key: bytes, nonce: bytes):
        assert len(key) == 32 and len(nonce) == 12
        self.h = lib().oracle_chacha20_new(_buf(key), _buf(nonce))

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_chacha20_free(self.h)
            self.h = None

    def fill_bytes(self, n: int) -> bytes:
        out = (ctypes.c_uint8 * max(n, 1))()
        lib().oracle_chacha20_fill(self.h, out, n)
        return bytes(out)[:n]

    def next_u32(self) -> int:
        return lib().oracle_chacha20_next_u32(self.h)

    def next_u64(self) -> int:
        return lib().oracle_chacha20_next_u64(self.h)

    def next_field(self) -> int:
        return lib().oracle_chacha20_next_field(self.h)


def chacha20_block(key: bytes, counter: int, nonce: bytes) -> bytes:
    out = (ctypes.c_uint8 * 64)()
    lib().oracle_chacha20_block(_buf(key), counter, _buf(nonce), out)
    return bytes(out)


def chacha20_quarter(state, a, b, c, d):
    arr = u32_array(state)
    lib().oracle_chacha20_quarter(arr, a, b, c, d)
    return list(arr)


def mask_coeffs(coeffs, n, n_rand, rng: ChaCha20):
    cap = max(1, 2 * (n + n_rand))
    out = (ctypes.c_uint32 * cap)()
    used = ctypes.c_size_t()
    rc = lib().oracle_mask_coeffs(u32_array(coeffs), len(coeffs), n, n_rand, rng.h, out, cap, ctypes.byref(used))
    if rc:
        raise OracleError(rc)
    return list(out)[:used.value]


def air_prove_zk(params: AirParams, n_rand: int, rng: ChaCha20, transcript: Transcript) -> bytes:
    out = u8p()
    n = ctypes.c_size_t()
    rc = lib().oracle_air_prove_zk(ctypes.byref(params), n_rand, rng.h, transcript.h, ctypes.byref(out),
                                   ctypes.byref(n))
    if rc:
        raise OracleError(rc)
    return _take(out, n.value)


def air_verify(params: AirParams, transcript: Transcript, proof: bytes) -> int:
    return lib().oracle_air_verify(ctypes.byref(params), transcript.h, _buf(proof), len(proof))


def _idx(params):
    n = int(params.num_queries)
    return (ctypes.c_uint64 * max(n, 1))(), n


def fri_commit_open(coeffs, params: FriParams, transcript: Transcript, return_indices=False):
    """fri::commit + public_commitment + open: (commitment, proof) and, with
    return_indices, the indices fri::open returns (fri.rs:645-666)."""
    cm, pf = u8p(), u8p()
    cn, pn = ctypes.c_size_t(), ctypes.c_size_t()
    idx, nq = _idx(params)
    rc = lib().oracle_fri_commit_open(u32_array(coeffs), len(coeffs), ctypes.byref(params), transcript.h,
                                      ctypes.byref(cm), ctypes.byref(cn), ctypes.byref(pf), ctypes.byref(pn), idx)
    if rc:
        raise ValueError("oracle fri error %d" % rc)
    out = (_take(cm, cn.value), _take(pf, pn.value))
    return out + (list(idx[:nq]),) if return_indices else out


FRI_KIND_BASE = -1


class FriState:
    """fri::ProverState from fri::commit* (kinds: FRI_KIND_BASE coefficients,
    FRI_KIND_EXT, FRI_KIND_EVALS, FRI_KIND_EXT_EVALS)."""

    def __init__(self, kind, values, params: FriParams, transcript: Transcript, degree_bound=0):
        words = [int(v) for v in values]
        count = len(words) if kind in (FRI_KIND_BASE, FRI_KIND_EVALS) else len(words) // 4
        h = ctypes.c_void_p()
        rc = lib().oracle_fri_commit_state(kind, u32_array(words), count, degree_bound, ctypes.byref(params),
                                           transcript.h, ctypes.byref(h))
        if rc:
            raise OracleError(rc)
        self.h, self.params = h.value, params

    def commitment(self) -> bytes:
        cm, cn = u8p(), ctypes.c_size_t()
        rc = lib().oracle_fri_state_commitment(self.h, ctypes.byref(cm), ctypes.byref(cn))
        if rc:
            raise OracleError(rc)
        return _take(cm, cn.value)

    def open(self, transcript: Transcript):
        pf, pn = u8p(), ctypes.c_size_t()
        idx, nq = _idx(self.params)
        rc = lib().oracle_fri_state_open(self.h, transcript.h, ctypes.byref(pf), ctypes.byref(pn), idx)
        if rc:
            raise OracleError(rc)
        return _take(pf, pn.value), list(idx[:nq])

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_fri_state_free(self.h)
            self.h = None


def fri_params_new(blowup, num_queries, final_size, grinding_bits=0, cap_height=0) -> FriParams:
    """fri::Params::new (fri.rs:141-158) plus its two builders; validated
    when used, as the reference validates at every entry point."""
    p = FriParams()
    p.blowup, p.num_queries, p.final_size = blowup, num_queries, final_size
    p.grinding_bits, p.cap_height = grinding_bits, cap_height
    return p


FRI_KIND_EXT, FRI_KIND_EVALS, FRI_KIND_EXT_EVALS = 0, 1, 2


def fri_commit_open_kind(kind, values, params: FriParams, transcript: Transcript, degree_bound=0,
                         return_indices=False):
    """fri::commit_ext (kind 0, Ext4 coefficients as 4 words each),
    commit_evals (1, base LDE values) or commit_ext_evals (2, Ext4 LDE
    values), then public_commitment + open; returns (commitment, proof)
    and, with return_indices, the sampled indices."""
    words = [int(v) for v in values]
    count = len(words) if kind in (FRI_KIND_BASE, FRI_KIND_EVALS) else len(words) // 4
    cm, pf = u8p(), u8p()
    cn, pn = ctypes.c_size_t(), ctypes.c_size_t()
    idx, nq = _idx(params)
    rc = lib().oracle_fri_commit_open_kind(kind, u32_array(words), count, degree_bound, ctypes.byref(params),
                                           transcript.h, ctypes.byref(cm), ctypes.byref(cn), ctypes.byref(pf),
                                           ctypes.byref(pn), idx)
    if rc:
        raise OracleError(rc)
    out = (_take(cm, cn.value), _take(pf, pn.value))
    return out + (list(idx[:nq]),) if return_indices else out


def evaluate_on_coset_ext(coeffs4, blowup):
    n = len(coeffs4) // 4
    out = (ctypes.c_uint32 * (4 * n * blowup))()
    rc = lib().oracle_evaluate_on_coset_ext(u32_array([int(v) for v in coeffs4]), n, blowup, out)
    if rc:
        raise OracleError(rc)
    return list(out)


def fri_verify(commitment: bytes, proof: bytes, params: FriParams, transcript: Transcript, indices=None) -> int:
    """Status of fri::verify; on success `indices` (a list, when given)
    receives the indices it returns (fri.rs:686)."""
    idx, nq = _idx(params)
    rc = lib().oracle_fri_verify(_buf(commitment), len(commitment), _buf(proof), len(proof), ctypes.byref(params),
                                 transcript.h, idx)
    if rc == 0 and indices is not None:
        indices[:] = list(idx[:nq])
    return rc


def stage_timers():
    n = lib().oracle_stage_timers(None, 0)
    buf = ctypes.create_string_buffer(n)
    lib().oracle_stage_timers(buf, n)
    out = {}
    for part in buf.value.decode().split(";"):
        if part:
            k, v = part.split("=")
            out[k] = float(v)
    return out


def clear_stage_timers():
    lib().oracle_stage_timers_clear()


# ---- numpy conveniences over the primitive entry points ----
def _np():
    import numpy as np
    return np


def _u32ptr(arr):
    return arr.ctypes.data_as(u32p)


def ntt_inverse(values):
    np = _np()
    a = np.ascontiguousarray(np.asarray(values, dtype=np.uint32)).copy()
    rc = lib().oracle_ntt_inverse(_u32ptr(a), a.size)
    if rc:
        raise ValueError("oracle ntt error %d" % rc)
    return a


def ntt_forward(values):
    np = _np()
    a = np.ascontiguousarray(np.asarray(values, dtype=np.uint32)).copy()
    rc = lib().oracle_ntt_forward(_u32ptr(a), a.size)
    if rc:
        raise ValueError("oracle ntt error %d" % rc)
    return a


def ntt_evaluate(coeffs, point):
    np = _np()
    a = np.ascontiguousarray(np.asarray(coeffs, dtype=np.uint32))
    out = ctypes.c_uint32()
    rc = lib().oracle_ntt_evaluate(_u32ptr(a), a.size, point, ctypes.byref(out))
    if rc:
        raise ValueError("oracle error %d" % rc)
    return out.value


def evaluate_on_coset(coeffs, blowup):
    np = _np()
    a = np.ascontiguousarray(np.asarray(coeffs, dtype=np.uint32))
    out = np.zeros(a.size * blowup, dtype=np.uint32)
    rc = lib().oracle_evaluate_on_coset(_u32ptr(a), a.size, blowup, _u32ptr(out))
    if rc:
        raise ValueError("oracle lde error %d" % rc)
    return out


def merkle_flat64(flat: bytes):
    """All levels leaf to root (32 bytes per digest), concatenated."""
    leaves = len(flat) // 64
    p2 = 1
    while p2 < leaves:
        p2 <<= 1
    out = (ctypes.c_uint8 * (32 * (2 * p2 - 1)))()
    rc = lib().oracle_merkle_flat64(_buf(flat), len(flat), out)
    if rc:
        raise ValueError("oracle merkle error %d" % rc)
    return bytes(out)


def linked_lde(lde_a, n, actual_b_last, a0):
    np = _np()
    a = np.ascontiguousarray(np.asarray(lde_a, dtype=np.uint32))
    out = np.zeros_like(a)
    rc = lib().oracle_linked_lde(_u32ptr(a), a.size, n, actual_b_last, a0, _u32ptr(out))
    if rc:
        raise ValueError("oracle linked error %d" % rc)
    return out


def compose(lde_a, lde_b, n, a0, b0, b_last, alphas, scalar=False):
    np = _np()
    a = np.ascontiguousarray(np.asarray(lde_a, dtype=np.uint32))
    b = np.ascontiguousarray(np.asarray(lde_b, dtype=np.uint32))
    al = np.ascontiguousarray(np.asarray(alphas, dtype=np.uint32).reshape(20))
    out = np.zeros(4 * a.size, dtype=np.uint32)
    rc = lib().oracle_compose(_u32ptr(a), _u32ptr(b), a.size, n, a0, b0, b_last, _u32ptr(al), 1 if scalar else 0,
                              _u32ptr(out))
    if rc:
        raise ValueError("oracle compose error %d" % rc)
    return out.reshape(-1, 4)


def fold_values(quad, inv_point, challenges, unit):
    np = _np()
    q = np.ascontiguousarray(np.asarray(quad, dtype=np.uint32).reshape(16))
    c = np.ascontiguousarray(np.asarray(challenges, dtype=np.uint32).reshape(12))
    out = np.zeros(4, dtype=np.uint32)
    lib().oracle_fold_values(_u32ptr(q), inv_point, _u32ptr(c), unit, _u32ptr(out))
    return out


def point_at(lde_len, depth, j):
    out = ctypes.c_uint32()
    rc = lib().oracle_point_at(lde_len, depth, j, ctypes.byref(out))
    if rc:
        raise ValueError("oracle error %d" % rc)
    return out.value


def field_inv(a):
    return lib().oracle_field_inv(a)


def field_mul(a, b):
    return lib().oracle_field_mul(a, b)


def field_pow(a, e):
    return lib().oracle_field_pow(a, e)


def two_adic_generator(bits):
    out = ctypes.c_uint32()
    rc = lib().oracle_two_adic_generator(bits, ctypes.byref(out))
    if rc:
        raise ValueError("oracle error %d" % rc)
    return out.value


def grind_head(head: bytes, bits: int) -> int:
    return lib().oracle_grind_head(_buf(head), len(head), bits)


def build_trace(a0, b0, n):
    np = _np()
    a = np.zeros(n, dtype=np.uint32)
    b = np.zeros(n, dtype=np.uint32)
    lib().oracle_build_trace(a0, b0, n, _u32ptr(a), _u32ptr(b))
    return a, b


# range.rs: the range-check AIR.
def range_params_new(n_rows, n_bits, blowup, nq, final_size, value, grinding_bits=0, cap_height=0) -> RangeParams:
    p = RangeParams()
    rc = lib().oracle_range_params_new(n_rows, n_bits, blowup, nq, final_size, value, ctypes.byref(p))
    if rc:
        raise OracleError(rc)
    p.grinding_bits, p.cap_height = grinding_bits, cap_height
    return p


def range_params_for_security(n_rows, n_bits, value, bits=100) -> RangeParams:
    p = RangeParams()
    rc = lib().oracle_range_params_for_security(n_rows, n_bits, value, bits, ctypes.byref(p))
    if rc:
        raise OracleError(rc)
    return p


def _range_out(call):
    out = u8p()
    n = ctypes.c_size_t()
    rc = call(ctypes.byref(out), ctypes.byref(n))
    if rc:
        raise OracleError(rc)
    return _take(out, n.value)


def range_prove(params: RangeParams, transcript: Transcript) -> bytes:
    return _range_out(lambda o, n: lib().oracle_range_prove(ctypes.byref(params), transcript.h, o, n))


def range_prove_with_columns(params: RangeParams, columns, transcript: Transcript) -> bytes:
    """columns: V then the bit columns, each a list of n_rows values."""
    length = len(columns[0]) if columns else 0
    flat = [v for c in columns for v in c]
    return _range_out(lambda o, n: lib().oracle_range_prove_with_columns(
        ctypes.byref(params), u32_array(flat), len(columns), length, transcript.h, o, n))


def range_prove_coeffs(params: RangeParams, columns, degree_bound, transcript: Transcript) -> bytes:
    length = len(columns[0]) if columns else 0
    flat = [v for c in columns for v in c]
    return _range_out(lambda o, n: lib().oracle_range_prove_coeffs(
        ctypes.byref(params), u32_array(flat), len(columns), length, degree_bound, transcript.h, o, n))


def range_prove_zk(params: RangeParams, n_rand: int, rng: ChaCha20, transcript: Transcript) -> bytes:
    return _range_out(lambda o, n: lib().oracle_range_prove_zk(ctypes.byref(params), n_rand, rng.h, transcript.h,
                                                               o, n))


def range_verify(params: RangeParams, transcript: Transcript, proof: bytes) -> int:
    return lib().oracle_range_verify(ctypes.byref(params), transcript.h, _buf(proof), len(proof))


def range_composition_at(v, bits, point, value, n, alphas):
    """alphas: n_bits + 2 Ext4 values as 4-tuples (decomposition, bits, boundary)."""
    out = (ctypes.c_uint32 * 4)()
    flat = [c for a in alphas for c in a]
    rc = lib().oracle_range_composition_at(v, u32_array(bits), len(bits), point, value, n, u32_array(flat), out)
    if rc:
        raise OracleError(rc)
    return tuple(out)


# table.rs: the table-membership AIR.
def table_params_new(n_rows, blowup, nq, final_size, value, grinding_bits=0, cap_height=0) -> TableParams:
    p = TableParams()
    rc = lib().oracle_table_params_new(n_rows, blowup, nq, final_size, value, ctypes.byref(p))
    if rc:
        raise OracleError(rc)
    p.grinding_bits, p.cap_height = grinding_bits, cap_height
    return p


def table_params_for_security(n_rows, table_len, value, bits=100) -> TableParams:
    p = TableParams()
    rc = lib().oracle_table_params_for_security(n_rows, table_len, value, bits, ctypes.byref(p))
    if rc:
        raise OracleError(rc)
    return p


def table_membership_bound(table_len, n, n_rand):
    return lib().oracle_table_membership_bound(table_len, n, n_rand)


def table_prove(params: TableParams, table, transcript: Transcript) -> bytes:
    return _range_out(lambda o, n: lib().oracle_table_prove(ctypes.byref(params), u32_array(table), len(table),
                                                            transcript.h, o, n))


def table_prove_coeffs(params: TableParams, table, coeffs, degree_bound, transcript: Transcript) -> bytes:
    return _range_out(lambda o, n: lib().oracle_table_prove_coeffs(
        ctypes.byref(params), u32_array(table), len(table), u32_array(coeffs), len(coeffs), degree_bound,
        transcript.h, o, n))


def table_prove_zk(params: TableParams, table, n_rand: int, rng: ChaCha20, transcript: Transcript) -> bytes:
    return _range_out(lambda o, n: lib().oracle_table_prove_zk(ctypes.byref(params), u32_array(table), len(table),
                                                               n_rand, rng.h, transcript.h, o, n))


def table_verify(params: TableParams, table, transcript: Transcript, proof: bytes) -> int:
    return lib().oracle_table_verify(ctypes.byref(params), u32_array(table), len(table), transcript.h, _buf(proof),
                                     len(proof))


# fold.rs: the IVC accumulator.
def fold_params_new(seg, blowup, nq, final_size, in_a, in_b, out_b, grinding_bits=0, cap_height=0) -> FoldParams:
    p = FoldParams()
    rc = lib().oracle_fold_params_new(seg, blowup, nq, final_size, in_a, in_b, out_b, ctypes.byref(p))
    if rc:
        raise OracleError(rc)
    p.grinding_bits, p.cap_height = grinding_bits, cap_height
    return p


def fold_params_for_security(seg, in_a, in_b, out_b, bits=100) -> FoldParams:
    p = FoldParams()
    rc = lib().oracle_fold_params_for_security(seg, in_a, in_b, out_b, bits, ctypes.byref(p))
    if rc:
        raise OracleError(rc)
    return p


class FoldAccumulator:
    """fold::Accumulator (fold.rs:361-570)."""

    def __init__(self, params: FoldParams):
        self.h = lib().oracle_fold_acc_new(ctypes.byref(params))

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_fold_acc_free(self.h)
            self.h = None

    def __len__(self):
        return lib().oracle_fold_acc_len(self.h)

    def set_params(self, params: FoldParams):
        lib().oracle_fold_acc_set_params(self.h, ctypes.byref(params))

    def add_segment(self, col_a, col_b, transcript: Transcript):
        rc = lib().oracle_fold_acc_add_segment(self.h, u32_array(col_a), len(col_a), u32_array(col_b), len(col_b),
                                               transcript.h)
        if rc:
            raise OracleError(rc)

    def finalize(self, transcript: Transcript) -> bytes:
        return _range_out(lambda o, n: lib().oracle_fold_acc_finalize(self.h, transcript.h, o, n))

    def finalize_forged_out_a(self, step, out_a, label: bytes) -> bytes:
        return _range_out(lambda o, n: lib().oracle_fold_acc_finalize_forged_out_a(self.h, step, out_a, _buf(label),
                                                                                  len(label), o, n))


def fold_verify(params: FoldParams, transcript: Transcript, proof: bytes) -> int:
    return lib().oracle_fold_verify(ctypes.byref(params), transcript.h, _buf(proof), len(proof))


def chained_traces(seg, count, a=1, b=1):
    """fold.rs:895-905: `count` segments of `seg` rows, each starting at the
    previous segment's last row."""
    out = []
    for _ in range(count):
        ca = (ctypes.c_uint32 * seg)()
        cb = (ctypes.c_uint32 * seg)()
        lib().oracle_build_trace(a, b, seg, ca, cb)
        ca, cb = list(ca), list(cb)
        a, b = ca[-1], cb[-1]
        out.append((ca, cb))
    return out
