"""The CPU producers and the CPU oracle under AddressSanitizer and
UndefinedBehaviorSanitizer (tests/sanitize/): round trips of every codec
and width, plus damaged and truncated streams through the oracle decoders.
Any sanitizer report aborts the driver, failing this test."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_producers_and_oracle_under_asan_ubsan(tmp_path):
    mk = os.path.join(ROOT, "tests", "sanitize", "Makefile")
    p = subprocess.run(["make", "-s", "-f", mk, f"OUT={tmp_path}", "run"], capture_output=True,
                       text=True, timeout=900,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
                                UBSAN_OPTIONS="print_stacktrace=1"))
    assert p.returncode == 0, (p.stdout + p.stderr)[-4000:]
    assert "0 failures" in p.stdout
