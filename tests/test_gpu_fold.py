"""IVC accumulation (src/fold.rs) on the GPU: every proof must be
byte-identical to the oracle's, advance the transcript as the oracle does,
and verify exactly when the oracle's does. Cases cover the reference bench
(4 segments of 8 rows), production shapes up to 2^16-row segments, caps and
grinding, a poisoned segment, broken chains and the error codes."""
import pytest

import oracle_ffi as o
from helpers import P

import z0p_amd as z

pytestmark = pytest.mark.gpu


def _both(op, zp, traces, label, ctx):
    ot, zt = o.Transcript(label), z.Transcript(label)
    oacc, zacc = o.FoldAccumulator(op), z.fold.Accumulator(zp, ctx)
    for ca, cb in traces:
        oacc.add_segment(ca, cb, ot)
        zacc.add_segment(ca, cb, zt)
    assert len(zacc) == len(traces)
    want = oacc.finalize(ot)
    got = zacc.finalize(zt)
    assert len(zacc) == 0
    assert zt.challenge_bytes(b"after") == ot.challenge_bytes(b"after")
    return got, want


def _verdict(op, zp, label, proof):
    ok_o = o.fold_verify(op, o.Transcript(label), proof) == 0
    try:
        z.fold.verify(zp, proof, z.Transcript(label))
        ok_z = True
    except z.Z0pError:
        ok_z = False
    assert ok_o == ok_z
    return ok_o


def test_bench_instance(ctx):
    # benches/speed_size.rs:432-480: the reference's 16845-byte proof.
    traces = o.chained_traces(8, 4)
    out_b = traces[-1][1][-1]
    op, zp = o.fold_params_new(8, 8, 8, 4, 1, 1, out_b), z.fold.params_new(8, 8, 8, 4, 1, 1, out_b)
    got, want = _both(op, zp, traces, b"bench/fold", ctx)
    assert len(got) == 16845 and got == want
    assert _verdict(op, zp, b"bench/fold", got)


@pytest.mark.parametrize("seg,count,a,b", [(16, 2, 1, 1), (1 << 10, 5, 3, 7), (1 << 14, 3, 1, 2), (1 << 16, 2, 5, 8)])
def test_for_security_byte_identical(ctx, seg, count, a, b):
    traces = o.chained_traces(seg, count, a, b)
    out_b = traces[-1][1][-1]
    op, zp = o.fold_params_for_security(seg, a, b, out_b), z.fold.for_security(seg, a, b, out_b)
    print("fold for_security", seg, count, zp.shape(), flush=True)
    got, want = _both(op, zp, traces, b"fs", ctx)
    assert got == want
    assert _verdict(op, zp, b"fs", got)


@pytest.mark.parametrize("blowup,fs,cap,grind", [(4, 2, 1, 3), (8, 8, 3, 0), (16, 4, 8, 5)])
def test_shapes_byte_identical(ctx, blowup, fs, cap, grind):
    seg = 64
    if (((seg * blowup) // fs).bit_length() - 1) % 2:
        fs *= 2
    traces = o.chained_traces(seg, 3)
    out_b = traces[-1][1][-1]
    op = o.fold_params_new(seg, blowup, 9, fs, 1, 1, out_b).replace(cap_height=cap, grinding_bits=grind)
    zp = z.fold.params_new(seg, blowup, 9, fs, 1, 1, out_b).replace(cap_height=cap, grinding_bits=grind)
    print("fold shape", zp.shape(), flush=True)
    got, want = _both(op, zp, traces, b"shape", ctx)
    assert got == want
    assert _verdict(op, zp, b"shape", got)


def test_bad_segment_and_broken_chain(ctx):
    # fold.rs:1151-1170: a poisoned middle segment; claims that do not chain.
    traces = o.chained_traces(64, 3)
    out_b = traces[-1][1][-1]
    op, zp = o.fold_params_new(64, 8, 12, 8, 1, 1, out_b), z.fold.params_new(64, 8, 12, 8, 1, 1, out_b)
    poisoned = [traces[0], ([*traces[1][0][:4], (traces[1][0][4] + 1) % P, *traces[1][0][5:]], traces[1][1]),
                traces[2]]
    got, want = _both(op, zp, poisoned, b"badseg", ctx)
    assert got == want and not _verdict(op, zp, b"badseg", got)
    unchained = [traces[0], o.chained_traces(64, 1, 5, 6)[0], traces[2]]
    got, want = _both(op, zp, unchained, b"chain", ctx)
    assert got == want and not _verdict(op, zp, b"chain", got)


def test_errors_match_oracle(ctx):
    op, zp = o.fold_params_new(8, 8, 8, 4, 1, 1, 1), z.fold.params_new(8, 8, 8, 4, 1, 1, 1)
    zacc, oacc = z.fold.Accumulator(zp, ctx), o.FoldAccumulator(op)
    with pytest.raises(z.Z0pError) as ze:
        zacc.finalize(z.Transcript(b"e"))
    with pytest.raises(o.OracleError) as oe:
        oacc.finalize(o.Transcript(b"e"))
    assert ze.value.code == oe.value.code == 1
    with pytest.raises(z.Z0pError) as ze:
        zacc.add_segment([1] * 7, [1] * 7, z.Transcript(b"e"))
    assert ze.value.code == 8
    with pytest.raises(z.Z0pError) as ze:
        zacc.add_segment([P] * 8, [1] * 8, z.Transcript(b"e"))
    assert ze.value.code == 8
    assert len(zacc) == 0
    # final_size 8 over 64 points: no power-of-four fold ratio
    traces = o.chained_traces(8, 2)
    op8, zp8 = op.replace(final_size=8), zp.replace(final_size=8)
    ot, zt = o.Transcript(b"e"), z.Transcript(b"e")
    oacc, zacc = o.FoldAccumulator(op8), z.fold.Accumulator(zp8, ctx)
    for ca, cb in traces:
        oacc.add_segment(ca, cb, ot)
        zacc.add_segment(ca, cb, zt)
    with pytest.raises(o.OracleError) as oe:
        oacc.finalize(ot)
    with pytest.raises(z.Z0pError) as ze:
        zacc.finalize(zt)
    assert ze.value.code == oe.value.code


def test_smallest_segments(ctx):
    # Segments of two rows (the smallest), a 16-point LDE per segment.
    traces = o.chained_traces(2, 3)
    out_b = traces[-1][1][-1]
    op, zp = o.fold_params_new(2, 8, 4, 4, 1, 1, out_b), z.fold.params_new(2, 8, 4, 4, 1, 1, out_b)
    got, want = _both(op, zp, traces, b"tiny", ctx)
    assert got == want
    assert _verdict(op, zp, b"tiny", got)


def test_accumulator_reused_after_finalize(ctx):
    # finalize consumes the segments; the same accumulator then proves anew.
    traces = o.chained_traces(16, 2)
    out_b = traces[-1][1][-1]
    op, zp = o.fold_params_new(16, 8, 6, 8, 1, 1, out_b), z.fold.params_new(16, 8, 6, 8, 1, 1, out_b)
    zacc = z.fold.Accumulator(zp, ctx)
    for label in [b"r1", b"r2"]:
        zt, ot = z.Transcript(label), o.Transcript(label)
        oacc = o.FoldAccumulator(op)
        for ca, cb in traces:
            zacc.add_segment(ca, cb, zt)
            oacc.add_segment(ca, cb, ot)
        assert zacc.finalize(zt) == oacc.finalize(ot)
        assert len(zacc) == 0


def test_cap_covers_every_row(ctx):
    traces = o.chained_traces(8, 3)
    out_b = traces[-1][1][-1]
    op = o.fold_params_new(8, 8, 8, 4, 1, 1, out_b).replace(cap_height=8)
    zp = z.fold.params_new(8, 8, 8, 4, 1, 1, out_b).replace(cap_height=8)
    got, want = _both(op, zp, traces, b"c8", ctx)
    assert got == want
    assert _verdict(op, zp, b"c8", got)
