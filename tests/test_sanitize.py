"""AddressSanitizer + UndefinedBehaviorSanitizer build of the CPU code:
the oracle, the host format library of libvbt_hip.so and the synthetic
encoder (make -C oracle san, driver tests/native/sanitize_check.cpp), run
over every golden fixture, its truncations and byte-flipped copies, and over
small encodes of every field kind.  A sanitizer finding aborts the driver
with its report; valid inputs decode without an error."""
import os
import re
import shutil
import subprocess

import pytest

import golden_io

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "oracle", "_san", "sanitize_check")
FIXTURES = ["smoke64", "liquid_mini", "sci_mini", "sci_dense_mini", "sci_tile_mini", "sci_patch_mini"]


@pytest.fixture(scope="module")
def driver():
    if shutil.which("g++") is None or shutil.which("gcc") is None:
        pytest.skip("gcc/g++ not available")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True, capture_output=True,
                   timeout=600)
    return DRIVER


def test_sanitized_oracle_format_library_and_encoder(driver, tmp_path):
    paths = []
    for name in FIXTURES:
        p = tmp_path / f"{name}.vbtp"
        p.write_bytes(golden_io.vbtp(name))
        paths.append(str(p))
    env = dict(os.environ, ASAN_OPTIONS="abort_on_error=1:detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([driver] + paths + ["--encode"], capture_output=True, text=True, timeout=600, env=env)
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    lines = {m.group(1): (int(m.group(2)), int(m.group(3)))
             for m in re.finditer(r"^ok (.+?): (\d+) calls (\d+) errors on the file", r.stdout, re.M)}
    assert len(lines) == len(FIXTURES) + 5
    for name, (calls, errors) in lines.items():
        assert calls > 0
        if "sci_patch_mini" in name:
            assert errors > 0  # its poisoned leaves are errors for the reference too
        else:
            assert errors == 0, name
