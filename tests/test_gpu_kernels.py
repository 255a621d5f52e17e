"""Kernel-level parity on the GPU: each device stage against the oracle on
seeded inputs (bit-exact: all arithmetic is exact field/ARX integer work)."""
import ctypes

import numpy as np
import pytest

import oracle_ffi as o
from helpers import P, brev_perm, rand_field

from z0p_amd import _ffi

pytestmark = pytest.mark.gpu


def _ptr(a):
    return a.ctypes.data_as(_ffi.u32p)


def _check(rc):
    assert rc == 0, _ffi.ERROR_NAMES.get(rc, rc)


@pytest.mark.parametrize("logn", [1, 4, 10, 11, 12, 13, 15, 16, 17, 18, 20])
def test_intt_bit_reversed(ctx, logn):
    rng = np.random.default_rng(100 + logn)
    n = 1 << logn
    evals = rand_field(rng, n)
    out = np.zeros(n, dtype=np.uint32)
    print("intt logn", logn, flush=True)
    _check(_ffi.lib().z0p_test_intt(ctx.handle, _ptr(evals), n, _ptr(out)))
    coeffs = o.ntt_inverse(evals)
    np.testing.assert_array_equal(out, coeffs[brev_perm(logn)])


@pytest.mark.parametrize("logn,logb", [(1, 2), (3, 4), (10, 4), (11, 4), (12, 3), (13, 4), (14, 2), (15, 1), (16, 4), (17, 3), (19, 2)])
def test_lde_bit_reversed(ctx, logn, logb):
    rng = np.random.default_rng(200 + 16 * logn + logb)
    n, B = 1 << logn, 1 << logb
    coeffs = rand_field(rng, n)
    out = np.zeros(n * B, dtype=np.uint32)
    print("lde logn", logn, "logb", logb, flush=True)
    _check(_ffi.lib().z0p_test_lde(ctx.handle, _ptr(coeffs), n, B, _ptr(out)))
    lde = o.evaluate_on_coset(coeffs, B)
    np.testing.assert_array_equal(out, lde[brev_perm(logn + logb)])


@pytest.mark.parametrize("leaves", [1, 2, 8, 512, 1024, 8192, 1 << 15, 1 << 18])
def test_merkle64_all_levels(ctx, leaves):
    rng = np.random.default_rng(300 + leaves)
    flat = rng.integers(0, 256, size=64 * leaves, dtype=np.uint8).tobytes()
    out = (ctypes.c_uint8 * (32 * (2 * leaves - 1)))()
    print("merkle64 leaves", leaves, flush=True)
    _check(_ffi.lib().z0p_test_merkle64(ctx.handle, _ffi.buf(flat), leaves, out))
    assert bytes(out) == o.merkle_flat64(flat)


def _packed_rows(a, b):
    rows = np.empty(2 * a.size, dtype=np.uint32)
    rows[0::2] = a
    rows[1::2] = b
    return rows.tobytes()


@pytest.mark.parametrize("logn,logb", [(1, 2), (2, 3), (8, 4), (12, 4), (13, 2), (14, 3), (16, 4)])
def test_trace_tree_all_levels(ctx, logn, logb):
    rng = np.random.default_rng(400 + 16 * logn + logb)
    n, B = 1 << logn, 1 << logb
    m = n * B
    a = rand_field(rng, m)
    b = rand_field(rng, m)
    leaves = m // 8
    out = (ctypes.c_uint8 * (32 * (2 * leaves - 1)))()
    print("trace tree logn", logn, "logb", logb, flush=True)
    _check(_ffi.lib().z0p_test_trace_tree(ctx.handle, _ptr(a), _ptr(b), n, B, out))
    assert bytes(out) == o.merkle_flat64(_packed_rows(a, b))


@pytest.mark.parametrize("logn,logb", [(1, 2), (6, 4), (12, 4), (13, 3)])
def test_linked_column(ctx, logn, logb):
    rng = np.random.default_rng(500 + logn)
    n, B = 1 << logn, 1 << logb
    a = rand_field(rng, n * B)
    bl, a0 = int(rng.integers(0, P)), int(rng.integers(0, P))
    out = np.zeros_like(a)
    print("linked logn", logn, flush=True)
    _check(_ffi.lib().z0p_test_linked(ctx.handle, _ptr(a), n, B, bl, a0, _ptr(out)))
    np.testing.assert_array_equal(out, o.linked_lde(a, n, bl, a0))


@pytest.mark.parametrize("logn,logb", [(1, 2), (6, 4), (12, 4), (13, 3)])
def test_composition(ctx, logn, logb):
    rng = np.random.default_rng(600 + logn)
    n, B = 1 << logn, 1 << logb
    m = n * B
    a = rand_field(rng, m)
    b = rand_field(rng, m)
    a0, b0, bl = (int(x) for x in rng.integers(0, P, size=3))
    alphas = rand_field(rng, 20)
    out = np.zeros(4 * m, dtype=np.uint32)
    print("compose logn", logn, flush=True)
    _check(_ffi.lib().z0p_test_compose(ctx.handle, _ptr(a), _ptr(b), n, B, a0, b0, bl, _ptr(alphas), _ptr(out),
                                       None))
    want = o.compose(a, b, n, a0, b0, bl, alphas)
    np.testing.assert_array_equal(out.reshape(-1, 4), want)


@pytest.mark.parametrize("logn,logb", [(1, 2), (2, 3), (6, 4), (12, 4), (13, 3), (16, 4)])
def test_composition_with_layer0_leaves(ctx, logn, logb):
    # The fused kernel: values as above, plus FRI layer 0's quad leaves and
    # their parents, against the oracle tree over the bit-reversed layer.
    rng = np.random.default_rng(650 + logn)
    n, B = 1 << logn, 1 << logb
    m = n * B
    a = rand_field(rng, m)
    b = rand_field(rng, m)
    a0, b0, bl = (int(x) for x in rng.integers(0, P, size=3))
    alphas = rand_field(rng, 20)
    out = np.zeros(4 * m, dtype=np.uint32)
    lv = np.zeros(32 * (m // 4 + m // 8), dtype=np.uint8)
    print("compose+leaves logn", logn, "logb", logb, flush=True)
    _check(_ffi.lib().z0p_test_compose(ctx.handle, _ptr(a), _ptr(b), n, B, a0, b0, bl, _ptr(alphas), _ptr(out),
                                       lv.ctypes.data_as(_ffi.u8p)))
    want = np.asarray(o.compose(a, b, n, a0, b0, bl, alphas), dtype=np.uint32)
    np.testing.assert_array_equal(out.reshape(-1, 4), want)
    layer = want[brev_perm(logn + logb)]  # FRI layer 0: bit-reversed order
    tree = o.merkle_flat64(layer.astype("<u4").tobytes())
    assert lv.tobytes() == tree[:32 * (m // 4 + m // 8)]


@pytest.mark.parametrize("loglde,depth", [(4, 0), (8, 1), (14, 0), (14, 3), (18, 2)])
def test_fold(ctx, loglde, depth):
    rng = np.random.default_rng(700 + loglde + depth)
    lde = 1 << loglde
    length = lde >> (2 * depth)
    layer = rand_field(rng, 4 * length)
    ch = rand_field(rng, 12)
    out = np.zeros(length, dtype=np.uint32)
    print("fold lde", lde, "depth", depth, flush=True)
    _check(_ffi.lib().z0p_test_fold(ctx.handle, _ptr(layer), length, lde, depth, _ptr(ch), _ptr(out), None))
    unit = o.two_adic_generator(2)
    quads = layer.reshape(-1, 16)
    got = out.reshape(-1, 4)
    idx = list(range(min(64, quads.shape[0]))) + list(rng.integers(0, quads.shape[0], size=64))
    for i in idx:
        inv = o.field_inv(o.point_at(lde, depth, 4 * int(i)))
        np.testing.assert_array_equal(got[i], o.fold_values(quads[i], inv, ch, unit), err_msg="quad %d" % i)


@pytest.mark.parametrize("loglde,depth", [(5, 0), (8, 1), (14, 0), (16, 2), (20, 0)])
def test_fold_with_next_leaves(ctx, loglde, depth):
    # The fused kernel: every folded value against the oracle fold, and the
    # folded layer's quad leaves and parents against the oracle tree.
    rng = np.random.default_rng(750 + loglde + depth)
    lde = 1 << loglde
    length = lde >> (2 * depth)
    layer = rand_field(rng, 4 * length)
    ch = rand_field(rng, 12)
    out = np.zeros(length, dtype=np.uint32)
    leaves = length // 16
    lv = np.zeros(32 * (leaves + leaves // 2), dtype=np.uint8)
    print("fold+leaves lde", lde, "depth", depth, flush=True)
    _check(_ffi.lib().z0p_test_fold(ctx.handle, _ptr(layer), length, lde, depth, _ptr(ch), _ptr(out),
                                    lv.ctypes.data_as(_ffi.u8p)))
    _check(_ffi.lib().z0p_test_fold(ctx.handle, _ptr(layer), length, lde, depth, _ptr(ch), _ptr(plain := np.zeros(
        length, dtype=np.uint32)), None))
    np.testing.assert_array_equal(out, plain)
    unit = o.two_adic_generator(2)
    quads = layer.reshape(-1, 16)
    got = out.reshape(-1, 4)
    idx = list(range(min(32, quads.shape[0]))) + list(rng.integers(0, quads.shape[0], size=32))
    for i in idx:
        inv = o.field_inv(o.point_at(lde, depth, 4 * int(i)))
        np.testing.assert_array_equal(got[i], o.fold_values(quads[i], inv, ch, unit), err_msg="quad %d" % i)
    tree = o.merkle_flat64(out.astype("<u4").tobytes())
    assert lv.tobytes() == tree[:len(lv)]


@pytest.mark.parametrize("bits", [0, 1, 8, 14, 18])
def test_grind_smallest_nonce(ctx, bits):
    rng = np.random.default_rng(800 + bits)
    head = rng.integers(0, 256, size=47, dtype=np.uint8).tobytes()
    nonce = ctypes.c_uint64()
    print("grind bits", bits, flush=True)
    _check(_ffi.lib().z0p_test_grind(ctx.handle, _ffi.buf(head), 47, bits, ctypes.byref(nonce)))
    assert nonce.value == o.grind_head(head, bits)
