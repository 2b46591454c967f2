"""The library's host C code (csrc/host_init.c: the reference initializer,
Fold3dPMM.c; csrc/io.c: the reference's %f dump and the VTK writer) built with
gcc under AddressSanitizer and UndefinedBehaviorSanitizer, every finding
fatal, and run on cubic, box and tiny grids plus the file-error paths.  CPU
only; the GPU code is not part of this build."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "fgcs-openacc-advanced_amd", "csrc")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    exe = tmp_path / "host_driver"
    cmd = ["gcc", "-std=gnu99", "-O1", "-g", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer",
           os.path.join(REPO, "tests", "sanitize", "host_driver.c"),
           os.path.join(CSRC, "host_init.c"), os.path.join(CSRC, "io.c"),
           "-o", str(exe), "-lm"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, env=env,
                       timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
    assert r.stdout.strip().endswith("ok")
