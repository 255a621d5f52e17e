"""z0p-mi355x: MI355X (gfx950) drop-in for the z0p Fibonacci AIR prover, the
range-check and table-membership AIRs, the IVC accumulator and the FRI
polynomial commitment.

Mirrors the reference crate's public API (src/air.rs, src/range.rs,
src/table.rs, src/fold.rs, src/fri.rs, src/transcript.rs): ``Transcript``, ``air.Params``/``air.prove``/
``air.verify``, ``range.Params``/``range.prove``/``range.verify``,
``table.Params``/``table.prove``/``table.verify``,
``fold.Accumulator``/``fold.verify`` and
``fri.Params``/``fri.commit_open``/``fri.verify``. Proofs
are the reference's canonical VERSION = 6 bytes. Proving runs on the GPU
through libz0p_amd.so; verification is host code, as in the reference.
"""
import ctypes

from ._ffi import Z0pError, buf, check, lib, take

__all__ = ["Z0pError", "Transcript", "ChaCha20", "Context", "Server", "air", "fri", "range", "table", "fold", "default_context"]


class Transcript:
    """Fiat-Shamir transcript (src/transcript.rs:54-171)."""

    def __init__(self, label: bytes = b"", _handle=None):
        L = lib()
        self._h = _handle if _handle is not None else L.z0p_transcript_new(buf(label), len(label))
        if not self._h:
            raise MemoryError("transcript allocation failed")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().z0p_transcript_free(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def clone(self):
        return Transcript(_handle=lib().z0p_transcript_clone(self._h))

    def absorb(self, label: bytes, data: bytes):
        lib().z0p_transcript_absorb(self._h, label, buf(data), len(data))

    def challenge_bytes(self, label: bytes) -> bytes:
        out = (ctypes.c_uint8 * 32)()
        lib().z0p_transcript_challenge_bytes(self._h, label, out)
        return bytes(out)

    def challenge_bb(self, label: bytes) -> int:
        """Transcript::challenge_bb (transcript.rs:124-129)."""
        return lib().z0p_transcript_challenge_bb(self._h, label)

    def challenge_ext(self, label: bytes):
        """Transcript::challenge_ext (transcript.rs:136-141): four canonical words."""
        out = (ctypes.c_uint32 * 4)()
        lib().z0p_transcript_challenge_ext(self._h, label, out)
        return tuple(out)

    def challenge_index(self, label: bytes, domain_size: int):
        """Transcript::challenge_index (transcript.rs:148-160): None unless the
        domain is a nonzero power of two."""
        out = ctypes.c_uint64()
        rc = lib().z0p_transcript_challenge_index(self._h, label, domain_size, ctypes.byref(out))
        if rc == 3:
            return None
        check(rc)
        return out.value

    def grind(self, label: bytes, difficulty_bits: int, ctx=None) -> int:
        """Transcript::grind (transcript.rs:206-228) on the GPU of ctx."""
        ctx = ctx or default_context()
        out = ctypes.c_uint64()
        check(lib().z0p_transcript_grind(ctx.handle, self._h, label, difficulty_bits, ctypes.byref(out)))
        return out.value

    def verify_grind(self, label: bytes, nonce: int, difficulty_bits: int) -> bool:
        """Transcript::verify_grind (transcript.rs:235-246)."""
        return lib().z0p_transcript_verify_grind(self._h, label, nonce, difficulty_bits) == 1


class ChaCha20:
    """ChaCha20 blinding stream (src/rng.rs:82-168): ``ChaCha20(key, nonce)``
    or ``ChaCha20.from_os()`` (rng::fresh_stream)."""

    def __init__(self, key: bytes = None, nonce: bytes = None, _handle=None):
        if _handle is None:
            if key is None or nonce is None or len(key) != 32 or len(nonce) != 12:
                raise ValueError("ChaCha20 needs a 32-byte key and a 12-byte nonce")
            _handle = lib().z0p_chacha20_new(buf(key), buf(nonce))
            if not _handle:
                raise MemoryError("rng allocation failed")
        self._h = _handle

    @classmethod
    def from_os(cls):
        h = ctypes.c_void_p()
        check(lib().z0p_chacha20_from_os(ctypes.byref(h)))
        return cls(_handle=h.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().z0p_chacha20_free(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def fill_bytes(self, n: int) -> bytes:
        out = (ctypes.c_uint8 * max(n, 1))()
        lib().z0p_chacha20_fill_bytes(self._h, out, n)
        return bytes(out)[:n]

    def next_u32(self) -> int:
        return lib().z0p_chacha20_next_u32(self._h)

    def next_u64(self) -> int:
        return lib().z0p_chacha20_next_u64(self._h)

    def next_field(self) -> int:
        return lib().z0p_chacha20_next_field(self._h)


class Context:
    """One HIP stream on one device plus its device memory (z0p_ctx)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        check(lib().z0p_ctx_create(device, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().z0p_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def handle(self):
        return self._h

    def set_profiling(self, enabled: bool, only: str = None):
        """Per-launch HIP events on the context stream; `only` restricts
        them to launches of one kernel name."""
        check(lib().z0p_ctx_profile_only(self._h, only.encode() if only else None))
        check(lib().z0p_ctx_set_profiling(self._h, 1 if enabled else 0))

    def kernel_stats(self):
        """{kernel: (launches, total_ms)} for launches since the last reset."""
        L = lib()
        n = L.z0p_ctx_kernel_stats(self._h, None, 0)
        s = ctypes.create_string_buffer(n)
        L.z0p_ctx_kernel_stats(self._h, s, n)
        out = {}
        for part in s.value.decode().split(";"):
            if part:
                name, rest = part.split("=")
                cnt, ms = rest.split(",")
                out[name] = (int(cnt), float(ms))
        return out

    def reset_stats(self):
        lib().z0p_ctx_reset_stats(self._h)

    def get_schedule(self) -> dict:
        """The kernel launch schedule (z0p_schedule) as a dict."""
        from ._ffi import ScheduleC
        s = ScheduleC()
        check(lib().z0p_ctx_get_schedule(self.handle, ctypes.byref(s)))
        return s.to_dict()

    def set_schedule(self, schedule: dict):
        from ._ffi import ScheduleC
        s = ScheduleC.from_dict(schedule)
        check(lib().z0p_ctx_set_schedule(self.handle, ctypes.byref(s)))

    def autotune(self, params_list, reps: int = 3):
        """Times the workload (one proof, or a batch) under candidate
        schedules and keeps the fastest on this context. Returns
        (schedule dict, best ms, starting ms, trial log)."""
        from ._ffi import AirParamsC, ScheduleC
        ps = (AirParamsC * len(params_list))(*params_list)
        out = ScheduleC()
        best, start = ctypes.c_double(), ctypes.c_double()
        check(lib().z0p_ctx_autotune(self.handle, ps, len(params_list), reps, ctypes.byref(out), ctypes.byref(best),
                                     ctypes.byref(start)))
        n = lib().z0p_ctx_autotune_log(self.handle, None, 0)
        buf = ctypes.create_string_buffer(n)
        lib().z0p_ctx_autotune_log(self.handle, buf, n)
        return out.to_dict(), best.value, start.value, buf.value.decode()

    def set_hbm_budget(self, nbytes: int):
        """Device-memory budget of batched proving on this context (0: auto)."""
        check(lib().z0p_ctx_set_hbm_budget(self._h, nbytes))

    def set_waves(self, min_active: int):
        """Batched and served proving alternate two waves of proofs (host
        transcript steps of one overlapping the other's kernels) once at
        least ``min_active`` are active; 0 turns waves off."""
        check(lib().z0p_ctx_set_waves(self._h, min_active))


class Server:
    """Asynchronous proving loop (z0p_server) that takes over a Context:
    requests arriving while others run join the running batch as device
    memory and slots free up. ``submit`` returns a request id; ``poll``
    returns ``(id, proof bytes or Z0pError, latency_ms, queue_ms)`` or None.
    ``large_schedule`` (a schedule dict) runs the lane of 2^16 steps and up
    under a schedule of its own; by default both lanes use the context's.
    """

    def __init__(self, ctx: "Context", hbm_budget: int = 0, max_active: int = 0, large_schedule: dict = None):
        from ._ffi import ScheduleC
        h = ctypes.c_void_p()
        ls = ctypes.byref(ScheduleC.from_dict(large_schedule)) if large_schedule else None
        check(lib().z0p_server_create_lanes(ctx.handle, ls, hbm_budget, max_active, ctypes.byref(h)))
        self._h = h
        self._ctx = ctx  # keep the context alive while the server runs

    def submit(self, params, transcript) -> int:
        rid = ctypes.c_uint64()
        check(lib().z0p_server_submit_air(self._h, ctypes.byref(params), transcript.handle, ctypes.byref(rid)))
        return rid.value

    def poll(self, timeout_ms: int = 0, zero_copy: bool = False):
        """With ``zero_copy`` the proof is a read-only memoryview of the
        buffer the C ABI hands over instead of a ``bytes`` copy."""
        from ._ffi import u8p, view
        rid = ctypes.c_uint64()
        st = ctypes.c_int()
        out = u8p()
        n = ctypes.c_size_t()
        lat = ctypes.c_double()
        q = ctypes.c_double()
        rc = lib().z0p_server_poll(self._h, timeout_ms, ctypes.byref(rid), ctypes.byref(st), ctypes.byref(out),
                                   ctypes.byref(n), ctypes.byref(lat), ctypes.byref(q))
        if rc == 1:
            return None
        check(rc)
        res = (view if zero_copy else take)(out, n.value) if st.value == 0 else Z0pError(st.value)
        return rid.value, res, lat.value, q.value

    def pending(self) -> int:
        return lib().z0p_server_pending(self._h)

    def close(self):
        if getattr(self, "_h", None):
            lib().z0p_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


_default = None


def default_context() -> Context:
    global _default
    if _default is None:
        _default = Context(0)
    return _default


from . import air, fold, fri, merkle, range, table  # noqa: E402
