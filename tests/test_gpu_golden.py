"""GPU proofs against the committed golden fixtures (tests/golden/): full
bytes for the reference bench rows, SHA-256 for the rest, including the
2^20-step AIR proof and the 2^22-coefficient FRI commitment."""
import hashlib
import json
import os

import pytest

from helpers import fri_bench_coeffs

import z0p_amd as z

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(GOLDEN, "manifest.json")))


@pytest.mark.parametrize("case", MANIFEST["air"], ids=lambda c: "air%d-seed%s" % (c["n"], c.get("seed")))
def test_gpu_air_matches_golden(ctx, case):
    n, a0, b0 = case["n"], case["a0"], case["b0"]
    print("golden air", n, case.get("seed"), flush=True)
    proof = z.air.prove(z.air.for_security(n, a0, b0, z.air.b_last(a0, b0, n)),
                        z.Transcript(case["label"].encode()), ctx)
    assert len(proof) == case["bytes"]
    assert hashlib.sha256(proof).hexdigest() == case["sha256"]
    if "file" in case:
        assert proof == open(os.path.join(GOLDEN, case["file"]), "rb").read()


@pytest.mark.parametrize("case", MANIFEST["fri"], ids=lambda c: "fri%d" % c["n"])
def test_gpu_fri_matches_golden(ctx, case):
    n = case["n"]
    print("golden fri", n, flush=True)
    c, p = z.fri.commit_open(fri_bench_coeffs(n), z.fri.for_security(n), z.Transcript(case["label"].encode()), ctx)
    assert (len(c), len(p)) == (case["commitment_bytes"], case["proof_bytes"])
    assert hashlib.sha256(c + p).hexdigest() == case["sha256"]
