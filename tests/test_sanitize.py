"""Sanitizer runs on the CPU: the oracle's C restatement driven through every entry
point under AddressSanitizer and UndefinedBehaviorSanitizer (oracle/sanitize_main.c,
`make -C oracle sanitize`), and the host code of libolx.so (file readers and writers,
host-side rules, argument validation) under the same sanitizers
(tests/cpp/Makefile.san; about five minutes of compilation, so it runs when
OLX_SANITIZE=1 is set)."""
import os
import subprocess

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def test_oracle_under_sanitizers():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"],
                       capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert "oracle sanitize run: ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr


@pytest.mark.skipif(os.environ.get("OLX_SANITIZE") != "1",
                    reason="the sanitized build of libolx takes minutes: set OLX_SANITIZE=1")
def test_libolx_host_code_under_sanitizers():
    r = subprocess.run(["make", "-s", "-f", "tests/cpp/Makefile.san", "-j8", "run"], cwd=ROOT,
                       capture_output=True, text=True, timeout=1800)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert "host sanitize run: ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
