"""The voss C ABI driven from plain C (tests/c/test_capi.c, built by the
package build): known answers through every entry point, as a C caller of the
reference library sees them (library/tests/test_prime_count.cpp)."""

import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parent / "c" / "test_capi"


def test_c_program_known_answers():
    if not EXE.exists():
        pytest.fail("tests/c/test_capi is not built: run `make -C tests/c` (the package build does)")
    r = subprocess.run([str(EXE), "full"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "ALL TESTS PASSED" in r.stdout
