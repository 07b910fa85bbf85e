"""The oracle's C sources built with AddressSanitizer and
UndefinedBehaviorSanitizer and driven through every configuration they
restate (oracle/sanitize_check.c): any out-of-bounds access, leak or undefined
operation in the checker aborts the run."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'oracle')


@pytest.mark.skipif(shutil.which('gcc') is None, reason='needs gcc')
def test_oracle_under_address_and_ub_sanitizers():
    subprocess.run(['make', '-s', 'sanitize'], cwd=ORACLE, check=True)
    env = dict(os.environ, ASAN_OPTIONS='detect_leaks=1:abort_on_error=0', UBSAN_OPTIONS='print_stacktrace=1',
               OMP_NUM_THREADS='2')
    r = subprocess.run([os.path.join(ORACLE, 'build', 'sanitize_check')], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert 'runtime error' not in r.stderr, r.stderr
    assert 'all ran' in r.stdout
