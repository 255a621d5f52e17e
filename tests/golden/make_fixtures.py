"""Writes the golden proof fixtures of tests/golden/ with the oracle (the C++
restatement of the reference prover, oracle/).

Fixtures: full proof bytes for the reference bench rows N = 64 (label
bench/air-prod) and N = 1024 (bench/air-scale) with a0 = b0 = 1
(benches/speed_size.rs:654, 724), and SHA-256 digests with sizes for larger
or seeded cases: seeded N = 1024 claims, N = 2^16 and 2^20 bench rows, the FRI
bench rows (256 and 1024 coefficients) and the 2^22-coefficient FRI
commitment. Run from the repository root:

    python tests/golden/make_fixtures.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import oracle_ffi as o  # noqa: E402
from helpers import fri_bench_coeffs, seeded_claims  # noqa: E402


def air_case(n, a0, b0, label):
    bl = o.fib_b_last(a0, b0, n)
    return o.air_prove(o.air_params_for_security(n, a0, b0, bl), o.Transcript(label))


def fri_case(n, label):
    c, p = o.fri_commit_open(fri_bench_coeffs(n), o.fri_params_for_security(n), o.Transcript(label))
    return c + p, len(c)


def main():
    o.set_threads(os.cpu_count() or 1)
    manifest = {"air": [], "fri": []}
    for n, label in [(64, b"bench/air-prod"), (1024, b"bench/air-scale")]:
        proof = air_case(n, 1, 1, label)
        name = "air_%d_%s.bin" % (n, label.decode().replace("/", "-"))
        with open(os.path.join(HERE, name), "wb") as f:
            f.write(proof)
        manifest["air"].append({"n": n, "a0": 1, "b0": 1, "label": label.decode(), "bytes": len(proof),
                                "sha256": hashlib.sha256(proof).hexdigest(), "file": name})
    cases = [(1024, seed) for seed in (0, 1, 2)] + [(1 << 16, None), (1 << 20, None)]
    for n, seed in cases:
        a0, b0 = (1, 1) if seed is None else seeded_claims(seed)
        label = b"bench/air-scale"
        proof = air_case(n, a0, b0, label)
        manifest["air"].append({"n": n, "a0": a0, "b0": b0, "label": label.decode(), "seed": seed,
                                "bytes": len(proof), "sha256": hashlib.sha256(proof).hexdigest()})
        print("air", n, seed, len(proof), flush=True)
    for n, label in [(256, b"bench/fri-prod"), (1024, b"bench/fri-prod-big"), (1 << 22, b"bench/fri-prod-big")]:
        both, clen = fri_case(n, label)
        manifest["fri"].append({"n": n, "label": label.decode(), "commitment_bytes": clen,
                                "proof_bytes": len(both) - clen, "sha256": hashlib.sha256(both).hexdigest()})
        print("fri", n, len(both), flush=True)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
