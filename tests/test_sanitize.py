"""AddressSanitizer + UndefinedBehaviorSanitizer build of the host-only C code
(bksp_amd/csrc/host/io.c: the Matrix Market readers; host/gen.c: the seeded
generators of every workload) and of the whole CPU oracle (oracle/*.c), run on the
reference's fixture matrices by tests/sanitize/san_driver.c. Any out-of-bounds
access, leak or undefined operation aborts the run (-fno-sanitize-recover=all).
The HIP layer is not part of this build (no GPU sanitizers on this platform)."""
import glob
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT

SRC = ([os.path.join(ROOT, "tests", "sanitize", "san_driver.c"),
        os.path.join(ROOT, "bksp_amd", "csrc", "host", "io.c"),
        os.path.join(ROOT, "bksp_amd", "csrc", "host", "gen.c")]
       + sorted(glob.glob(os.path.join(ROOT, "oracle", "orc_*.c"))))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_host_code_and_oracle_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "san_driver")
    cmd = ["gcc", "-std=gnu11", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"), "-o", exe,
           *SRC, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, os.path.join(GOLDEN, "testmat", "bfwb398.mtx"),
                        os.path.join(GOLDEN, "testmat", "pde2961.mtx")],
                       capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "san_driver done" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
