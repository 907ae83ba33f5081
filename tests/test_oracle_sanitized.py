"""The oracle under AddressSanitizer, LeakSanitizer and UndefinedBehaviorSanitizer (the
reference's debug configurations segf and nans, makefiles/make.gcc.flags:14-20): `make -C
oracle sanitize` builds the oracle with a driver that calls each of its entry points on small
periodic, mixed-radix and non-periodic grids, with bodies and every planted mistake; any
sanitizer report ends the run with a non-zero status. CPU only."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_runs_clean_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    r = subprocess.run([os.path.join(ROOT, "oracle", "_build", "oracle_selftest_san")], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.stdout.strip().endswith("ok")
