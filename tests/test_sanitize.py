"""The FP64 oracle and the host build of the SASA core under AddressSanitizer
and UndefinedBehaviorSanitizer (oracle/Makefile `sanitize`), CPU only."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_and_sasa_core_clean_under_asan_ubsan():
    r = subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "sanitize"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "ran clean" in r.stdout
