"""AddressSanitizer / UndefinedBehaviorSanitizer runs of host code (CPU):

* the oracle (gcc, -fsanitize=address,undefined) over every entry point on
  small cases of orders 1 to 3 (the case files are the shim driver's);
* the assembly library built with the sanitizers on its host pass: the
  threaded CSR symbolic phase against a brute-force pattern, and the
  argument checks of init / init_batch / residual_static on a handle
  without a device.

Both programs are built into a temporary directory by
tests/sanitize/Makefile; the sanitizers stop the program at the first
invalid access or undefined operation.
"""

from __future__ import annotations

import os
import subprocess
from pathlib import Path

import pytest

from cases import case
from test_shim import write_case

ROOT = Path(__file__).resolve().parents[1]
SAN = ROOT / "tests" / "sanitize"
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


@pytest.fixture(scope="module")
def san_build(tmp_path_factory):
    out = tmp_path_factory.mktemp("san")
    r = subprocess.run(["make", "-s", "-f", str(SAN / "Makefile"), f"OUT={out}"],
                       capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return out


@pytest.mark.parametrize("name,p", [("grating", 1), ("hole", 2), ("grating", 3)])
def test_oracle_under_sanitizers(san_build, tmp_path, name, p):
    c, ad, I, J = case(name, p, 6.0, "tiny")
    write_case(tmp_path, ad, c.problem, I, J)
    r = subprocess.run([str(san_build / "san_oracle"), str(tmp_path)], capture_output=True,
                       text=True, timeout=600, env=ENV)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "san_oracle ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr


def test_library_host_code_under_sanitizers(san_build):
    r = subprocess.run([str(san_build / "san_lib")], capture_output=True, text=True, timeout=600,
                       env=ENV)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "san_lib ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
