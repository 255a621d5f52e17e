"""Shared helpers for the parity tests: seeded inputs and layout permutations."""
import numpy as np

P = 2013265921
MASK64 = (1 << 64) - 1


def splitmix64(state):
    """(new_state, output): src/gf2.rs:425-431."""
    state = (state + 0x9E3779B97F4A7C15) & MASK64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return state, z ^ (z >> 31)


def splitmix_fields(n, seed):
    """c_i = Field::from_u64(splitmix64(&mut st)) for i < n as a uint32 numpy
    array (the same stream as n calls of splitmix64, vectorised: the state
    after i + 1 calls is seed + (i + 1) * golden)."""
    i = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z % np.uint64(P)).astype(np.uint32)


def seeded_claims(seed):
    """(a0, b0) for a synthetic Fibonacci instance: the reference's test idiom
    Field::from_u64(splitmix64(&mut st)) twice (ntt.rs:453-455)."""
    st, x = splitmix64(seed)
    st, y = splitmix64(st)
    return x % P, y % P


def rand_field(rng, n):
    return rng.integers(0, P, size=n, dtype=np.uint64).astype(np.uint32)


def brev(x, bits):
    r = 0
    for _ in range(bits):
        r = (r << 1) | (x & 1)
        x >>= 1
    return r


def brev_perm(bits):
    """perm[i] = brev_bits(i) as a numpy index array."""
    n = 1 << bits
    idx = np.arange(n, dtype=np.uint64)
    r = np.zeros(n, dtype=np.uint64)
    for b in range(bits):
        r |= ((idx >> np.uint64(b)) & np.uint64(1)) << np.uint64(bits - 1 - b)
    return r.astype(np.int64)


def fri_bench_coeffs(n):
    """Coefficients of the reference FRI benches: Field::new(i*0x9e3779b9 + 17)
    with wrapping u32 arithmetic (benches/speed_size.rs:612-614)."""
    return [((i * 0x9E3779B9 + 17) & 0xFFFFFFFF) % P for i in range(n)]
