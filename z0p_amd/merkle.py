"""Generic byte-leaf Merkle helpers (merkle.rs:16-30, 165-216) over the C ABI.

Paths are lists of ``(sibling_digest, is_left)`` pairs, leaf level first, as
``MerkleTree::prove`` returns them; ``is_left`` means the running node is the
left child at that level.
"""
from ._ffi import buf, lib


def _path_bytes(path):
    out = bytearray()
    for sibling, is_left in path:
        if len(sibling) != 32:
            raise ValueError("path digests are 32 bytes")
        out += bytes(sibling) + (b"\x01" if is_left else b"\x00")
    return bytes(out)


def hash_leaf(data: bytes) -> bytes:
    """merkle::hash_leaf: BLAKE3(0x00 || data)."""
    import ctypes
    out = (ctypes.c_uint8 * 32)()
    lib().z0p_merkle_hash_leaf(buf(data), len(data), out)
    return bytes(out)


def hash_node(left: bytes, right: bytes) -> bytes:
    """merkle::hash_node: BLAKE3(0x01 || left || right)."""
    import ctypes
    if len(left) != 32 or len(right) != 32:
        raise ValueError("child digests are 32 bytes")
    out = (ctypes.c_uint8 * 32)()
    lib().z0p_merkle_hash_node(buf(left), buf(right), out)
    return bytes(out)


def verify(root: bytes, leaf: bytes, index: int, path) -> bool:
    """MerkleTree::verify: the path opens `leaf` at `index` under `root`."""
    if len(root) != 32 or index < 0 or index >= 1 << 64:
        return False
    pb = _path_bytes(path)
    return lib().z0p_merkle_verify(buf(root), buf(leaf), len(leaf), index, buf(pb), len(path)) == 1


def verify_digest(root: bytes, leaf_digest: bytes, index: int, path) -> bool:
    """MerkleTree::verify_digest: as verify, for an already-hashed leaf."""
    if len(root) != 32 or len(leaf_digest) != 32 or index < 0 or index >= 1 << 64:
        return False
    pb = _path_bytes(path)
    return lib().z0p_merkle_verify_digest(buf(root), buf(leaf_digest), index, buf(pb), len(path)) == 1
