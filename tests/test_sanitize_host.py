"""Host code of libpyqed_amd under AddressSanitizer and UndefinedBehaviorSanitizer: the hierarchy table
builder on shapes up to 58905 ADOs (checked against the defining relations of the tables), the plan
builder's host path on a 4096-instance batch with 64-bit element offsets, and its rejection of
non-finite and out-of-range inputs (tests/sanitize/host_driver.cpp, built by tests/sanitize/build.sh).
Runs on the CPU; needs hipcc."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(not (shutil.which("hipcc") or os.path.exists("/opt/rocm/bin/hipcc")), reason="needs hipcc")
@pytest.mark.skipif(not os.path.exists(os.path.join(HERE, "sanitize", "build.sh")),
                    reason="sanitizer build files not present (they are not shipped to GPU runs)")
def test_host_code_under_address_and_ub_sanitizers(tmp_path):
    build = subprocess.run(["bash", os.path.join(HERE, "sanitize", "build.sh"), str(tmp_path)],
                           capture_output=True, text=True, timeout=900)
    assert build.returncode == 0, build.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([str(tmp_path / "host_driver")], capture_output=True, text=True, timeout=900, env=env)
    out = run.stdout + run.stderr
    assert run.returncode == 0, out[-3000:]
    assert "0 failed checks" in out
    assert "AddressSanitizer" not in out and "runtime error" not in out
