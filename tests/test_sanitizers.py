"""The host half of the library (verifiers, parameter constructors,
transcripts, Merkle helpers and the verify_batch thread pool) built with
g++ under the sanitizers (`make -C z0p_amd/csrc sanitize`, which
__graft_entry__.build() runs), driven in a subprocess with the sanitizer
runtime preloaded and Z0P_HOST_LIBRARY selecting that build: the
adversarial suites and the verifier's host tests under ASan + UBSan, and a
threaded verify_batch under TSan. Any report aborts the subprocess."""
import os
import pathlib
import subprocess
import sys
import textwrap

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
SAN = ROOT / "z0p_amd" / "_san"


def _runtime(name):
    out = subprocess.run(["g++", "-print-file-name=%s" % name], capture_output=True, text=True).stdout.strip()
    return out if os.path.isabs(out) and os.path.exists(out) else None


def _env(lib, runtime, options):
    libpath = SAN / lib
    rt, std = _runtime(runtime), _runtime("libstdc++.so")
    if not libpath.exists() or not rt or not std:
        pytest.skip("sanitizer build or runtime missing: run `make -C z0p_amd/csrc sanitize`")
    env = dict(os.environ)
    # libstdc++ is preloaded too: the runtime's exception interceptors need
    # it resolved before the first throw (the Python binary does not link it).
    env.update({"LD_PRELOAD": "%s %s" % (rt, std), "Z0P_HOST_LIBRARY": str(libpath)}, **options)
    return env


def test_adversarial_suites_under_asan_ubsan():
    env = _env("libz0p_host_asan.so", "libasan.so", {
        "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
    })
    cmd = [sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
           str(ROOT / "tests" / "test_adversarial.py"),
           str(ROOT / "tests" / "test_host.py") + "::test_verifier_rejects_tampering",
           str(ROOT / "tests" / "test_host.py") + "::test_fri_verifier_accepts_and_rejects"]
    r = subprocess.run(cmd, env=env, cwd=ROOT, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])


def test_verify_batch_threads_under_tsan():
    env = _env("libz0p_host_tsan.so", "libtsan.so", {"TSAN_OPTIONS": "halt_on_error=1"})
    script = textwrap.dedent("""
        import sys
        sys.path.insert(0, %r)
        sys.path.insert(0, %r)
        import oracle_ffi as o
        import z0p_amd as z
        bl = o.fib_b_last(1, 1, 64)
        zp = z.air.params_new(64, 16, 8, 16, 1, 1, bl)
        op = o.air_params_new(64, 16, 8, 16, 1, 1, bl)
        proofs, labels = [], []
        for i in range(32):
            lab = b"tsan/%%d" %% i
            p = o.air_prove(op, o.Transcript(lab))
            if i %% 5 == 0:
                p = p[:100] + bytes([p[100] ^ 1]) + p[101:]
            proofs.append(p)
            labels.append(lab)
        for threads in (8, 3):
            res = z.air.verify_batch([zp] * 32, proofs, [z.Transcript(l) for l in labels], threads=threads)
            assert [r is None for r in res] == [i %% 5 != 0 for i in range(32)], res
        print("tsan verify_batch ok")
    """) % (str(ROOT), str(ROOT / "tests"))
    r = subprocess.run([sys.executable, "-c", script], env=env, cwd=ROOT, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0 and "tsan verify_batch ok" in r.stdout, (r.stdout[-3000:], r.stderr[-3000:])
