"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (CPU): `make -C oracle
sanitize` builds it with a driver that calls every entry point the tests use (all-pairs and
cell-list boxes, loss blocks with truncation and snapshots, MSD, MLP, energy, reweighting,
barostat, the stored-trajectory adjoint); the run must finish without a sanitizer report."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    env = dict(os.environ, OMP_NUM_THREADS="4", ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([os.path.join(ROOT, "oracle", "_san", "oracle_san")], env=env, capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "sanitized oracle run ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
