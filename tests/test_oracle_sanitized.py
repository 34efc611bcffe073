"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer:
oracle/san_check.c runs every oracle function once on small lattices (tiny
periodic grids where the neighbourhood wraps onto itself, channels with
solids on the inlet and outlet columns, the learned step with a saturating
network, training samples, the restriction), built with
-fsanitize=address,undefined -fno-sanitize-recover=all, so an
out-of-bounds access, a leak or undefined behaviour ends it with a report
and a non-zero status."""
import os
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_oracle_clean_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "lib/san_check"], check=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(ROOT / "oracle" / "lib" / "san_check")], capture_output=True, text=True, timeout=300,
                       env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "san_check ok" in r.stdout
