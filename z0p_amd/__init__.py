"""z0p-mi355x: MI355X (gfx950) drop-in for the z0p Fibonacci AIR prover and
FRI polynomial commitment.

Mirrors the reference crate's public API (src/air.rs, src/fri.rs,
src/transcript.rs): ``Transcript``, ``air.Params``/``air.prove``/
``air.verify`` and ``fri.Params``/``fri.commit_open``/``fri.verify``. Proofs
are the reference's canonical VERSION = 6 bytes. Proving runs on the GPU
through libz0p_amd.so; verification is host code, as in the reference.
"""
import ctypes

from ._ffi import Z0pError, buf, check, lib, take

__all__ = ["Z0pError", "Transcript", "Context", "air", "fri", "default_context"]


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

    def set_profiling(self, enabled: bool):
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


_default = None


def default_context() -> Context:
    global _default
    if _default is None:
        _default = Context(0)
    return _default


from . import air, fri  # noqa: E402
