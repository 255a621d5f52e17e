"""ctypes binding of libz0p_amd.so (the C ABI in include/z0p.h).

The library is built in-tree (``__graft_entry__.build()`` or
``make -C z0p_amd/csrc``); importing this package without it fails loudly, so
no code path can silently fall back to a CPU implementation.
"""
import ctypes
import os
import pathlib

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libz0p_amd.so"

u8p = ctypes.POINTER(ctypes.c_uint8)
u32p = ctypes.POINTER(ctypes.c_uint32)
u64p = ctypes.POINTER(ctypes.c_uint64)
u64 = ctypes.c_uint64
u32 = ctypes.c_uint32
sz = ctypes.c_size_t
vp = ctypes.c_void_p

ERROR_NAMES = {
    1: "EmptyInput",
    2: "IndexOutOfBounds",
    3: "InvalidDomainSize",
    4: "InvalidBlowup",
    5: "NoQueries",
    6: "BadMerklePath",
    7: "FriVerificationFailed",
    8: "MalformedInput",
    9: "RngUnavailable",
    100: "Device",
    101: "OutOfMemory",
    102: "InvalidArgument",
    103: "NoDevice",
}


class Z0pError(Exception):
    """A non-zero status from the library; ``code`` follows src/error.rs order."""

    def __init__(self, code):
        self.code = code
        self.name = ERROR_NAMES.get(code, "Unknown")
        super().__init__("%s (%d)" % (self.name, code))


class AirParamsC(ctypes.Structure):
    _fields_ = [
        ("n_steps", u64), ("blowup", u64), ("num_queries", u64), ("final_size", u64),
        ("grinding_bits", u32), ("reserved0", u32), ("cap_height", u64),
        ("a0", u32), ("b0", u32), ("b_last", u32), ("reserved1", u32),
    ]


class RangeParamsC(ctypes.Structure):
    _fields_ = [
        ("n_rows", u64), ("n_bits", u64), ("blowup", u64), ("num_queries", u64), ("final_size", u64),
        ("grinding_bits", u32), ("reserved0", u32), ("cap_height", u64), ("value", u32), ("reserved1", u32),
    ]


class TableParamsC(ctypes.Structure):
    _fields_ = [
        ("n_rows", u64), ("blowup", u64), ("num_queries", u64), ("final_size", u64),
        ("grinding_bits", u32), ("reserved0", u32), ("cap_height", u64), ("value", u32), ("reserved1", u32),
    ]


class FoldParamsC(ctypes.Structure):
    _fields_ = [
        ("seg_steps", u64), ("blowup", u64), ("num_queries", u64), ("final_size", u64),
        ("grinding_bits", u32), ("reserved0", u32), ("cap_height", u64),
        ("in_a", u32), ("in_b", u32), ("out_b", u32), ("reserved1", u32),
    ]


class FriParamsC(ctypes.Structure):
    _fields_ = [
        ("blowup", u64), ("num_queries", u64), ("final_size", u64),
        ("grinding_bits", u32), ("reserved0", u32), ("cap_height", u64),
    ]


class ScheduleC(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int32) for name in (
        "tail_max_log", "tail_group_max_log", "tail_spread_log", "levels_fuse", "small_max_log", "grind_blocks",
        "fold_fuse_max_log", "compose_fuse", "trace_fuse", "subtree_levels", "compose_subtree", "graph", "lde_split",
        "grind_pool", "fold_commit_max_log", "tree_top", "wave_streams")]

    def to_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}

    @classmethod
    def from_dict(cls, d):
        s = cls()
        for name, _ in cls._fields_:
            setattr(s, name, int(d[name]))
        return s


# name -> (restype, argtypes); every symbol declared in include/z0p.h and
# include/z0p_testing.h.
SIGNATURES = {
    "z0p_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(vp)]),
    "z0p_ctx_destroy": (None, [vp]),
    "z0p_ctx_set_profiling": (ctypes.c_int, [vp, ctypes.c_int]),
    "z0p_ctx_profile_only": (ctypes.c_int, [vp, ctypes.c_char_p]),
    "z0p_ctx_kernel_stats": (sz, [vp, ctypes.c_char_p, sz]),
    "z0p_ctx_reset_stats": (None, [vp]),
    "z0p_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "z0p_free": (None, [vp]),
    "z0p_transcript_new": (vp, [u8p, sz]),
    "z0p_transcript_clone": (vp, [vp]),
    "z0p_transcript_free": (None, [vp]),
    "z0p_transcript_absorb": (None, [vp, ctypes.c_char_p, u8p, sz]),
    "z0p_transcript_challenge_bytes": (None, [vp, ctypes.c_char_p, u8p]),
    "z0p_transcript_challenge_bb": (u32, [vp, ctypes.c_char_p]),
    "z0p_transcript_challenge_ext": (None, [vp, ctypes.c_char_p, u32p]),
    "z0p_transcript_challenge_index": (ctypes.c_int, [vp, ctypes.c_char_p, u64, ctypes.POINTER(u64)]),
    "z0p_transcript_grind": (ctypes.c_int, [vp, vp, ctypes.c_char_p, u32, ctypes.POINTER(u64)]),
    "z0p_transcript_grind_noalloc": (ctypes.c_int, [vp, vp, ctypes.c_char_p, u32, ctypes.POINTER(u64)]),
    "z0p_transcript_verify_grind": (ctypes.c_int, [vp, ctypes.c_char_p, u64, u32]),
    "z0p_air_params_new": (ctypes.c_int, [u64, u64, u64, u64, u32, u32, u32, ctypes.POINTER(AirParamsC)]),
    "z0p_air_params_for_security": (ctypes.c_int, [u64, u32, u32, u32, u32, ctypes.POINTER(AirParamsC)]),
    "z0p_fri_params_new": (ctypes.c_int, [u64, u64, u64, ctypes.POINTER(FriParamsC)]),
    "z0p_fri_params_for_security": (ctypes.c_int, [u64, u32, ctypes.POINTER(FriParamsC)]),
    "z0p_air_b_last": (u32, [u32, u32, u64]),
    "z0p_air_prove": (ctypes.c_int, [vp, ctypes.POINTER(AirParamsC), vp, ctypes.POINTER(u8p), ctypes.POINTER(sz)]),
    "z0p_air_verify": (ctypes.c_int, [ctypes.POINTER(AirParamsC), vp, u8p, sz]),
    "z0p_air_prove_with_trace": (ctypes.c_int, [vp, ctypes.POINTER(AirParamsC), u32p, sz, u32p, sz, vp,
                                                ctypes.POINTER(u8p), ctypes.POINTER(sz)]),
    "z0p_chacha20_new": (vp, [u8p, u8p]),
    "z0p_chacha20_from_os": (ctypes.c_int, [ctypes.POINTER(vp)]),
    "z0p_chacha20_free": (None, [vp]),
    "z0p_chacha20_fill_bytes": (None, [vp, u8p, sz]),
    "z0p_chacha20_next_u32": (u32, [vp]),
    "z0p_chacha20_next_u64": (u64, [vp]),
    "z0p_chacha20_next_field": (u32, [vp]),
    "z0p_air_prove_zk": (ctypes.c_int, [vp, ctypes.POINTER(AirParamsC), u64, vp, vp, ctypes.POINTER(u8p),
                                        ctypes.POINTER(sz)]),
    "z0p_range_params_new": (ctypes.c_int, [u64, u64, u64, u64, u64, u32, ctypes.POINTER(RangeParamsC)]),
    "z0p_range_params_for_security": (ctypes.c_int, [u64, u64, u32, u32, ctypes.POINTER(RangeParamsC)]),
    "z0p_range_prove": (ctypes.c_int, [vp, ctypes.POINTER(RangeParamsC), vp, ctypes.POINTER(u8p),
                                       ctypes.POINTER(sz)]),
    "z0p_range_prove_with_columns": (ctypes.c_int, [vp, ctypes.POINTER(RangeParamsC), u32p, sz, sz, vp,
                                                    ctypes.POINTER(u8p), ctypes.POINTER(sz)]),
    "z0p_range_prove_zk": (ctypes.c_int, [vp, ctypes.POINTER(RangeParamsC), u64, vp, vp, ctypes.POINTER(u8p),
                                          ctypes.POINTER(sz)]),
    "z0p_range_verify": (ctypes.c_int, [ctypes.POINTER(RangeParamsC), vp, u8p, sz]),
    "z0p_table_params_new": (ctypes.c_int, [u64, u64, u64, u64, u32, ctypes.POINTER(TableParamsC)]),
    "z0p_table_params_for_security": (ctypes.c_int, [u64, u64, u32, u32, ctypes.POINTER(TableParamsC)]),
    "z0p_table_membership_bound": (u64, [u64, u64, u64]),
    "z0p_table_prove": (ctypes.c_int, [vp, ctypes.POINTER(TableParamsC), u32p, sz, vp, ctypes.POINTER(u8p),
                                       ctypes.POINTER(sz)]),
    "z0p_table_prove_zk": (ctypes.c_int, [vp, ctypes.POINTER(TableParamsC), u32p, sz, u64, vp, vp,
                                          ctypes.POINTER(u8p), ctypes.POINTER(sz)]),
    "z0p_table_verify": (ctypes.c_int, [ctypes.POINTER(TableParamsC), u32p, sz, vp, u8p, sz]),
    "z0p_fold_params_new": (ctypes.c_int, [u64, u64, u64, u64, u32, u32, u32, ctypes.POINTER(FoldParamsC)]),
    "z0p_fold_params_for_security": (ctypes.c_int, [u64, u32, u32, u32, u32, ctypes.POINTER(FoldParamsC)]),
    "z0p_fold_acc_new": (ctypes.c_int, [vp, ctypes.POINTER(FoldParamsC), ctypes.POINTER(vp)]),
    "z0p_fold_acc_free": (None, [vp]),
    "z0p_fold_acc_len": (sz, [vp]),
    "z0p_fold_acc_add_segment": (ctypes.c_int, [vp, u32p, sz, u32p, sz, vp]),
    "z0p_fold_acc_finalize": (ctypes.c_int, [vp, vp, ctypes.POINTER(u8p), ctypes.POINTER(sz)]),
    "z0p_fold_verify": (ctypes.c_int, [ctypes.POINTER(FoldParamsC), vp, u8p, sz]),
    "z0p_fri_commit_open": (ctypes.c_int, [vp, u32p, sz, ctypes.POINTER(FriParamsC), vp, ctypes.POINTER(u8p),
                                           ctypes.POINTER(sz), ctypes.POINTER(u8p), ctypes.POINTER(sz), u64p, sz]),
    "z0p_fri_verify": (ctypes.c_int, [u8p, sz, u8p, sz, ctypes.POINTER(FriParamsC), vp, u64p, sz]),
    "z0p_merkle_hash_leaf": (None, [u8p, sz, u8p]),
    "z0p_merkle_hash_node": (None, [u8p, u8p, u8p]),
    "z0p_merkle_verify": (ctypes.c_int, [u8p, u8p, sz, u64, u8p, sz]),
    "z0p_merkle_verify_digest": (ctypes.c_int, [u8p, u8p, u64, u8p, sz]),
    "z0p_fri_commit_open_ext": (ctypes.c_int, [vp, u32p, sz, ctypes.POINTER(FriParamsC), vp, ctypes.POINTER(u8p),
                                               ctypes.POINTER(sz), ctypes.POINTER(u8p), ctypes.POINTER(sz), u64p,
                                               sz]),
    "z0p_fri_commit_open_evals": (ctypes.c_int, [vp, u32p, sz, u64, ctypes.POINTER(FriParamsC), vp,
                                                 ctypes.POINTER(u8p), ctypes.POINTER(sz), ctypes.POINTER(u8p),
                                                 ctypes.POINTER(sz), u64p, sz]),
    "z0p_fri_commit_open_ext_evals": (ctypes.c_int, [vp, u32p, sz, u64, ctypes.POINTER(FriParamsC), vp,
                                                     ctypes.POINTER(u8p), ctypes.POINTER(sz), ctypes.POINTER(u8p),
                                                     ctypes.POINTER(sz), u64p, sz]),
    "z0p_fri_commit": (ctypes.c_int, [vp, u32p, sz, ctypes.POINTER(FriParamsC), vp, ctypes.POINTER(vp)]),
    "z0p_fri_commit_ext": (ctypes.c_int, [vp, u32p, sz, ctypes.POINTER(FriParamsC), vp, ctypes.POINTER(vp)]),
    "z0p_fri_commit_evals": (ctypes.c_int, [vp, u32p, sz, u64, ctypes.POINTER(FriParamsC), vp, ctypes.POINTER(vp)]),
    "z0p_fri_commit_ext_evals": (ctypes.c_int, [vp, u32p, sz, u64, ctypes.POINTER(FriParamsC), vp,
                                                ctypes.POINTER(vp)]),
    "z0p_fri_state_commitment": (ctypes.c_int, [vp, ctypes.POINTER(u8p), ctypes.POINTER(sz)]),
    "z0p_fri_open": (ctypes.c_int, [vp, vp, ctypes.POINTER(u8p), ctypes.POINTER(sz), u64p, sz]),
    "z0p_fri_state_free": (None, [vp]),
    "z0p_ctx_set_hbm_budget": (ctypes.c_int, [vp, u64]),
    "z0p_ctx_set_waves": (ctypes.c_int, [vp, u32]),
    "z0p_air_verify_batch": (ctypes.c_int, [ctypes.POINTER(AirParamsC), ctypes.POINTER(vp), ctypes.POINTER(u8p),
                                            ctypes.POINTER(sz), sz, u32, ctypes.POINTER(ctypes.c_int)]),
    "z0p_air_verify_batch_device": (ctypes.c_int, [vp, ctypes.POINTER(AirParamsC), ctypes.POINTER(vp),
                                                   ctypes.POINTER(u8p), ctypes.POINTER(sz), sz, u32,
                                                   ctypes.POINTER(ctypes.c_int)]),
    "z0p_ctx_get_schedule": (ctypes.c_int, [vp, ctypes.POINTER(ScheduleC)]),
    "z0p_ctx_set_schedule": (ctypes.c_int, [vp, ctypes.POINTER(ScheduleC)]),
    "z0p_ctx_autotune": (ctypes.c_int, [vp, ctypes.POINTER(AirParamsC), sz, u32, ctypes.POINTER(ScheduleC),
                                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "z0p_ctx_autotune_log": (sz, [vp, ctypes.c_char_p, sz]),
    "z0p_air_prove_batch": (ctypes.c_int, [vp, ctypes.POINTER(AirParamsC), ctypes.POINTER(vp), sz,
                                           ctypes.POINTER(u8p), ctypes.POINTER(sz), ctypes.POINTER(ctypes.c_int)]),
    "z0p_server_create": (ctypes.c_int, [vp, u64, u32, ctypes.POINTER(vp)]),
    "z0p_server_create_lanes": (ctypes.c_int, [vp, ctypes.POINTER(ScheduleC), u64, u32, ctypes.POINTER(vp)]),
    "z0p_server_submit_air": (ctypes.c_int, [vp, ctypes.POINTER(AirParamsC), vp, ctypes.POINTER(u64)]),
    "z0p_server_poll": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(u64), ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(u8p), ctypes.POINTER(sz), ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double)]),
    "z0p_server_pending": (sz, [vp]),
    "z0p_server_destroy": (None, [vp]),
    # include/z0p_testing.h
    "z0p_test_intt": (ctypes.c_int, [vp, u32p, sz, u32p]),
    "z0p_test_lde": (ctypes.c_int, [vp, u32p, sz, sz, u32p]),
    "z0p_test_merkle64": (ctypes.c_int, [vp, u8p, sz, u8p]),
    "z0p_test_trace_tree": (ctypes.c_int, [vp, u32p, u32p, sz, sz, u8p]),
    "z0p_test_linked": (ctypes.c_int, [vp, u32p, sz, sz, u32, u32, u32p]),
    "z0p_test_compose": (ctypes.c_int, [vp, u32p, u32p, sz, sz, u32, u32, u32, u32p, u32p, u8p]),
    "z0p_test_fold": (ctypes.c_int, [vp, u32p, sz, sz, sz, u32p, u32p, u8p]),
    "z0p_test_grind": (ctypes.c_int, [vp, u8p, sz, u32, ctypes.POINTER(u64)]),
    "z0p_test_copy": (ctypes.c_int, [vp, sz, ctypes.POINTER(ctypes.c_float)]),
    "z0p_test_set_gate_spin": (ctypes.c_int, [vp, u32]),
}

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    host_only = os.environ.get("Z0P_HOST_LIBRARY")
    if host_only:
        # A host-only build (the verifier and host entries under a
        # sanitizer, `make -C z0p_amd/csrc sanitize`): it exports no device
        # entry, so calls that need one fail with AttributeError.
        L = ctypes.CDLL(host_only)
    else:
        if not LIB_PATH.exists():
            raise ImportError("libz0p_amd.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        if host_only and not hasattr(L, name):
            continue
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise Z0pError(rc)


def buf(data: bytes):
    return (ctypes.c_uint8 * max(len(data), 1)).from_buffer_copy(data if data else b"\0")


def u32_buf(values):
    import numpy as np
    arr = np.ascontiguousarray(np.asarray(values, dtype=np.uint32))
    return arr, arr.ctypes.data_as(u32p)


def take(ptr, n):
    out = ctypes.string_at(ptr, n)
    lib().z0p_free(ptr)
    return out


def view(ptr, n):
    """Zero-copy, read-only ``memoryview`` of a buffer the library handed
    over; the buffer is released (z0p_free) once the view and every view
    derived from it are gone. Compares equal to ``bytes`` of the same content."""
    import weakref
    addr = ctypes.cast(ptr, ctypes.c_void_p).value
    if n == 0 or not addr:
        lib().z0p_free(ptr)
        return memoryview(b"")
    arr = (ctypes.c_uint8 * n).from_address(addr)
    weakref.finalize(arr, lib().z0p_free, ctypes.c_void_p(addr))
    return memoryview(arr).cast("B").toreadonly()
