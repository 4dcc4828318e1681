"""The host-side C++ (Init.txt reader, reference and tube set-up, text output, the serving
scheduler on its host-only fake device, the C ABI's argument checks) and the oracle, built
with g++ under AddressSanitizer + UndefinedBehaviorSanitizer and run by a driver
(tests/sanitize/host_driver.cpp).  Any out-of-bounds access, use after free, leak in this
code or undefined behaviour aborts the driver.  No GPU is touched."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SAN = ROOT / "tests" / "sanitize"


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_code_under_asan_and_ubsan(tmp_path):
    # the driver links the package's kernel objects (built, never launched here)
    subprocess.run(["make", "-s", "-j8", "-C", str(ROOT / "dustyshock3d_amd" / "csrc")], check=True,
                   timeout=900)
    subprocess.run(["make", "-s", "-j8", "-C", str(SAN)], check=True, timeout=600)
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:exitcode=23"
    env["LSAN_OPTIONS"] = f"suppressions={SAN / 'lsan.supp'}"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    r = subprocess.run([str(SAN / "_build" / "host_driver"), str(ROOT / "tests" / "golden" / "Init.txt"),
                        str(tmp_path)], env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-6000:]
    assert "host driver: ok" in r.stdout
    # the text output was written
    assert any(p.name.startswith("G") for p in tmp_path.iterdir())
