"""Test-side VERSION = 6 wire codec (codec.rs:21-274, fri.rs:896-921 and
316-398, air.rs:250-334, fold.rs:267-359): decodes proofs and commitments
into plain Python structures so the adversarial suites can tamper with one
field and re-encode, as the reference suites mutate their Rust structs.
Decoding fails closed like the reference reader (count limits, canonical
field words, 0/1 direction bytes, no trailing bytes) with WireError.
"""
import struct
from dataclasses import dataclass, field
from typing import List, Tuple

VERSION = 6
P = 2013265921
MAX_QUERIES = 1 << 20
MAX_FOLDS = 64
MAX_PATH_STEPS = 128
MAX_CAP_DIGESTS = 256
MAX_ELEMS = 1 << 27
MAX_STEPS = 1 << 20

Path = List[Tuple[bytes, bool]]
Ext4 = Tuple[int, int, int, int]


class WireError(ValueError):
    pass


class Reader:
    def __init__(self, data: bytes):
        self.d = bytes(data)
        self.pos = 0

    def take(self, n):
        if n > len(self.d) - self.pos:
            raise WireError("truncated")
        out = self.d[self.pos:self.pos + n]
        self.pos += n
        return out

    def byte(self):
        return self.take(1)[0]

    def u32(self):
        return struct.unpack("<I", self.take(4))[0]

    def u64(self):
        return struct.unpack("<Q", self.take(8))[0]

    def field(self):
        v = self.u32()
        if v >= P:
            raise WireError("non-canonical field")
        return v

    def ext4(self) -> Ext4:
        return tuple(self.field() for _ in range(4))

    def count(self, limit):
        n = self.u32()
        if n > limit:
            raise WireError("count above limit")
        return n

    def digest(self):
        return self.take(32)

    def direction(self):
        b = self.byte()
        if b > 1:
            raise WireError("direction byte")
        return b == 1

    def path(self) -> Path:
        return [(self.digest(), self.direction()) for _ in range(self.count(MAX_PATH_STEPS))]

    def cap(self):
        return [self.digest() for _ in range(self.count(MAX_CAP_DIGESTS))]

    def version(self):
        if self.byte() != VERSION:
            raise WireError("version")

    def end(self):
        if self.pos != len(self.d):
            raise WireError("trailing bytes")


class Writer:
    def __init__(self):
        self.out = bytearray()

    def byte(self, b):
        self.out.append(b & 0xFF)

    def u32(self, x):
        self.out += struct.pack("<I", x & 0xFFFFFFFF)

    def u64(self, x):
        self.out += struct.pack("<Q", x & 0xFFFFFFFFFFFFFFFF)

    def field(self, v):
        self.u32(v)

    def ext4(self, e):
        for v in e:
            self.field(v)

    def path(self, p):
        self.u32(len(p))
        for d, is_left in p:
            self.out += d
            self.byte(1 if is_left else 0)

    def cap(self, c):
        self.u32(len(c))
        for d in c:
            self.out += d


@dataclass
class Commitment:
    lde_len: int = 0
    poly_len: int = 0
    degree_bound: int = 0
    caps: List[List[bytes]] = field(default_factory=list)

    @staticmethod
    def read_body(r: Reader):
        c = Commitment(r.u64(), r.u64(), r.u64())
        c.caps = [r.cap() for _ in range(r.count(MAX_FOLDS))]
        return c

    def write_body(self, w: Writer):
        w.u64(self.lde_len)
        w.u64(self.poly_len)
        w.u64(self.degree_bound)
        w.u32(len(self.caps))
        for c in self.caps:
            w.cap(c)

    @staticmethod
    def from_bytes(data: bytes):
        r = Reader(data)
        r.version()
        c = Commitment.read_body(r)
        r.end()
        return c

    def to_bytes(self):
        w = Writer()
        w.byte(VERSION)
        self.write_body(w)
        return bytes(w.out)


@dataclass
class FoldWitness:
    values: List[Ext4]
    path: Path


@dataclass
class FriProof:
    final_poly: List[Ext4] = field(default_factory=list)
    pow_nonce: int = 0
    queries: List[List[FoldWitness]] = field(default_factory=list)

    @staticmethod
    def read_body(r: Reader):
        p = FriProof()
        p.final_poly = [r.ext4() for _ in range(r.count(MAX_ELEMS))]
        p.pow_nonce = r.u64()
        for _ in range(r.count(MAX_QUERIES)):
            folds = []
            for _ in range(r.count(MAX_FOLDS)):
                vals = [r.ext4() for _ in range(4)]
                folds.append(FoldWitness(vals, r.path()))
            p.queries.append(folds)
        return p

    def write_body(self, w: Writer):
        w.u32(len(self.final_poly))
        for v in self.final_poly:
            w.ext4(v)
        w.u64(self.pow_nonce)
        w.u32(len(self.queries))
        for q in self.queries:
            w.u32(len(q))
            for f in q:
                for v in f.values:
                    w.ext4(v)
                w.path(f.path)

    @staticmethod
    def from_bytes(data: bytes):
        r = Reader(data)
        r.version()
        p = FriProof.read_body(r)
        r.end()
        return p

    def to_bytes(self):
        w = Writer()
        w.byte(VERSION)
        self.write_body(w)
        return bytes(w.out)


@dataclass
class TraceOpening:
    index: int
    a: int
    b: int
    a_next: int
    b_next: int
    path: Path
    shift_path: Path
    leaf: bytes
    shift_leaf: bytes


@dataclass
class AirProof:
    column_cap: List[bytes]
    fri_commitment: Commitment
    fri_proof: FriProof
    openings: List[TraceOpening]

    @staticmethod
    def from_bytes(data: bytes):
        r = Reader(data)
        r.version()
        cap = r.cap()
        cm = Commitment.read_body(r)
        fp = FriProof.read_body(r)
        ops = []
        for _ in range(r.count(MAX_QUERIES)):
            idx = r.u64()
            a, b, an, bn = r.field(), r.field(), r.field(), r.field()
            path, spath = r.path(), r.path()
            ops.append(TraceOpening(idx, a, b, an, bn, path, spath, r.take(64), r.take(64)))
        r.end()
        return AirProof(cap, cm, fp, ops)

    def to_bytes(self):
        w = Writer()
        w.byte(VERSION)
        w.cap(self.column_cap)
        self.fri_commitment.write_body(w)
        self.fri_proof.write_body(w)
        w.u32(len(self.openings))
        for o in self.openings:
            w.u64(o.index)
            for v in (o.a, o.b, o.a_next, o.b_next):
                w.field(v)
            w.path(o.path)
            w.path(o.shift_path)
            w.out += o.leaf
            w.out += o.shift_leaf
        return bytes(w.out)


@dataclass
class StepOpening:
    index: int
    cols: List[int]
    path: Path
    shift_path: Path


@dataclass
class Step:
    claims: List[int]  # in_a, in_b, out_a, out_b
    column_cap: List[bytes]
    openings: List[StepOpening]


@dataclass
class IvcProof:
    steps: List[Step]
    fri_commitment: Commitment
    fri_proof: FriProof

    @staticmethod
    def from_bytes(data: bytes):
        r = Reader(data)
        r.version()
        steps = []
        for _ in range(r.count(MAX_STEPS)):
            claims = [r.field() for _ in range(4)]
            cap = r.cap()
            ops = []
            for _ in range(r.count(MAX_QUERIES)):
                idx = r.u64()
                cols = [r.field() for _ in range(4)]
                ops.append(StepOpening(idx, cols, r.path(), r.path()))
            steps.append(Step(claims, cap, ops))
        cm = Commitment.read_body(r)
        fp = FriProof.read_body(r)
        r.end()
        return IvcProof(steps, cm, fp)

    def to_bytes(self):
        w = Writer()
        w.byte(VERSION)
        w.u32(len(self.steps))
        for s in self.steps:
            for v in s.claims:
                w.field(v)
            w.cap(s.column_cap)
            w.u32(len(s.openings))
            for o in s.openings:
                w.u64(o.index)
                for v in o.cols:
                    w.field(v)
                w.path(o.path)
                w.path(o.shift_path)
        self.fri_commitment.write_body(w)
        self.fri_proof.write_body(w)
        return bytes(w.out)


def ext_add_one(e: Ext4) -> Ext4:
    """e + Ext4::ONE (the constant coordinate comes first)."""
    return ((e[0] + 1) % P,) + tuple(e[1:])
