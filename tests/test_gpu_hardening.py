"""The prove-side halves of the reference's hardening and red-team cases
(tests/hardening.rs:21-171, tests/redteam.rs:374-416) on the GPU prover:
hand-built invalid parameter structs are refused with an error before any
device work, the FRI rate gate refuses a bound one past the ceiling, and the
context keeps proving correctly afterwards."""
import pytest

import oracle_ffi as o

import z0p_amd as z

pytestmark = pytest.mark.gpu

HARD = b"hardening"
P = 2013265921


def _air_valid():
    bl = o.fib_b_last(1, 1, 32)
    return lambda: z.air.params_new(32, 4, 8, 8, 1, 1, bl)


@pytest.mark.parametrize("change", [dict(n_steps=0), dict(n_steps=3), dict(blowup=1), dict(blowup=3),
                                    dict(blowup=(1 << 64) - 1), dict(num_queries=0),
                                    dict(num_queries=(1 << 64) - 1), dict(final_size=0), dict(final_size=3),
                                    dict(grinding_bits=64)])
def test_air_prove_refuses_invalid_params(ctx, change):
    print("air prove with", change, flush=True)
    bad = _air_valid()()
    for k, v in change.items():
        setattr(bad, k, v)
    with pytest.raises(z.Z0pError):
        z.air.prove(bad, z.Transcript(HARD), ctx)


def test_range_table_fold_prove_refuse_invalid_params(ctx):
    for change in [dict(n_rows=0), dict(n_rows=3), dict(n_bits=0), dict(n_bits=33), dict(blowup=1),
                   dict(num_queries=0), dict(final_size=0)]:
        print("range prove with", change, flush=True)
        bad = z.range.params_new(32, 8, 4, 8, 8, 5)
        for k, v in change.items():
            setattr(bad, k, v)
        with pytest.raises(z.Z0pError):
            z.range.prove(bad, z.Transcript(HARD), ctx)
    table = [10, 20, 30, 40]
    for change in [dict(n_rows=0), dict(n_rows=3), dict(blowup=1), dict(num_queries=0),
                   dict(num_queries=(1 << 64) - 1), dict(final_size=3)]:
        print("table prove with", change, flush=True)
        bad = z.table.params_new(8, 8, 8, 4, 30)
        for k, v in change.items():
            setattr(bad, k, v)
        with pytest.raises(z.Z0pError):
            z.table.prove(bad, table, z.Transcript(HARD), ctx)
    ca, cb = o.chained_traces(32, 1)[0]
    for change in [dict(seg_steps=0), dict(seg_steps=3), dict(blowup=1), dict(blowup=3), dict(num_queries=0),
                   dict(final_size=0)]:
        print("fold segment with", change, flush=True)
        bad = z.fold.params_new(32, 4, 8, 8, ca[0], cb[0], cb[-1])
        for k, v in change.items():
            setattr(bad, k, v)
        with pytest.raises(z.Z0pError):
            acc = z.fold.Accumulator(bad, ctx)
            acc.add_segment(ca, cb, z.Transcript(HARD))
    print("recovery proof", flush=True)
    zp, op = _air_valid()(), o.air_params_new(32, 4, 8, 8, 1, 1, o.fib_b_last(1, 1, 32))
    assert z.air.prove(zp, z.Transcript(HARD), ctx) == o.air_prove(op, o.Transcript(HARD))


def test_rate_boundary_exact_prove_side(ctx):
    zp, op = z.fri.params_new(4, 8, 4), o.fri_params_new(4, 8, 4)
    coeffs = [((i * 0x9E3779B9 + 5) & 0xFFFFFFFF) % P for i in range(16)]
    lde = o.evaluate_on_coset(coeffs, 4)
    print("commit_evals at bound 31", flush=True)
    got = z.fri.commit_open_evals(lde, 31, zp, z.Transcript(b"redteam/edge"), ctx)
    assert got == o.fri_commit_open_kind(o.FRI_KIND_EVALS, lde, op, o.Transcript(b"redteam/edge"), 31)
    z.fri.verify(got[0], got[1], zp, z.Transcript(b"redteam/edge"))
    print("commit_evals at bound 32", flush=True)
    with pytest.raises(z.Z0pError):
        z.fri.commit_open_evals(lde, 32, zp, z.Transcript(b"redteam/edge"), ctx)
    with pytest.raises(z.Z0pError):
        z.fri.commit_evals(lde, 32, zp, z.Transcript(b"redteam/edge"), ctx)
