"""Host sanitizer run (AddressSanitizer + UndefinedBehaviorSanitizer, CPU
only): the C ABI's argument checks called directly from C
(tests/sanitize/abi_args.c, linked against a sanitized build of the library's
sources) and every oracle function on the reference's known-answer tests
(tests/sanitize/oracle_kats.c).  Any memory error, undefined behaviour or
leak aborts the run."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None or not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="needs gcc and hipcc")
def test_host_sanitizers(tmp_path):
    r = subprocess.run(["make", "-s", "-C", os.path.join(HERE, "sanitize"), f"OUT={tmp_path}",
                        "-j8", "run"], capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "abi_args: all argument checks passed" in out
    assert "oracle_kats: all checks passed" in out
