"""Host sanitizers: the oracle (oracle/rvpf_oracle.c) and the synthetic
workload generator (rivet-av1_amd/csrc/synth.c) built with AddressSanitizer
and UndefinedBehaviorSanitizer and run over small frames of every layout
(8/10/12 bits; 4:2:0, 4:2:2, 4:4:4, mono; super-resolution; odd sizes down
to 1x1):
the chain, each stage, and the encoder searches. Any out-of-bounds access,
leak, signed overflow or invalid shift fails the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None, reason="needs gcc and make")
def test_oracle_and_generator_under_asan_ubsan(tmp_path):
    out = str(tmp_path)
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize", "SAN_OUT=" + out], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([os.path.join(out, "sanitize_oracle")], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("rc 0") == 14, r.stdout
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, r.stderr
