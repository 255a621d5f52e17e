"""The committed golden fixtures (tests/golden/, written by the oracle with
tests/golden/make_fixtures.py): the oracle must keep reproducing them and
accept them, and their sizes are the reference's published sizes
(tools/baseline.json:78-112)."""
import hashlib
import json
import os

import pytest

import oracle_ffi as o
from helpers import fri_bench_coeffs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(GOLDEN, "manifest.json")))
PUBLISHED = {64: 33061, 1024: 68573, 65536: 141641, 1 << 20: 203553}


def _air(c):
    bl = o.fib_b_last(c["a0"], c["b0"], c["n"])
    return o.air_params_for_security(c["n"], c["a0"], c["b0"], bl)


@pytest.mark.parametrize("case", [c for c in MANIFEST["air"] if "file" in c], ids=lambda c: c["file"])
def test_fixture_bytes_reproduced_and_accepted(case):
    data = open(os.path.join(GOLDEN, case["file"]), "rb").read()
    assert len(data) == case["bytes"] == PUBLISHED[case["n"]]
    assert hashlib.sha256(data).hexdigest() == case["sha256"]
    p = _air(case)
    assert o.air_prove(p, o.Transcript(case["label"].encode())) == data
    assert o.air_verify(p, o.Transcript(case["label"].encode()), data) == 0


@pytest.mark.parametrize("case", [c for c in MANIFEST["air"] if "file" not in c and c["n"] <= 65536],
                         ids=lambda c: "air%d-seed%s" % (c["n"], c.get("seed")))
def test_air_digests_reproduced(case):
    proof = o.air_prove(_air(case), o.Transcript(case["label"].encode()))
    assert len(proof) == case["bytes"]
    assert hashlib.sha256(proof).hexdigest() == case["sha256"]


@pytest.mark.parametrize("case", [c for c in MANIFEST["fri"] if c["n"] <= 1024], ids=lambda c: "fri%d" % c["n"])
def test_fri_digests_reproduced(case):
    n = case["n"]
    c, p = o.fri_commit_open(fri_bench_coeffs(n), o.fri_params_for_security(n), o.Transcript(case["label"].encode()))
    assert (len(c), len(p)) == (case["commitment_bytes"], case["proof_bytes"])
    assert hashlib.sha256(c + p).hexdigest() == case["sha256"]
