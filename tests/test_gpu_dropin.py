"""The drop-in boundary from C: a driver of the reference's shape
(tests/dropin/driver.c) compiled against include/ and linked with
libfold3d.so, calling the reference-signature plane steps.  Its "%f" dump at
16^3 has the digest of the reference main()'s own dump, and at 32^3 it is the
golden field's dump."""
import hashlib
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LIBDIR = os.path.join(REPO, "fgcs-openacc-advanced_amd")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("dropin") / "driver")
    subprocess.run(["gcc", "-O2", "-std=gnu99", "-I", os.path.join(REPO, "include"),
                    os.path.join(HERE, "dropin", "driver.c"), "-L", LIBDIR, "-lfold3d",
                    f"-Wl,-rpath,{LIBDIR}", "-lm", "-o", exe], check=True)
    return exe


def _run(driver, n, iters, out):
    r = subprocess.run([driver, str(n), str(iters), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    return out.read_bytes()


def test_c_driver_dump_matches_reference_main(driver, tmp_path):
    ref = json.load(open(os.path.join(HERE, "golden", "ref_dump_16.json")))
    data = _run(driver, 16, 8, tmp_path / "d16.txt")
    assert hashlib.md5(data).hexdigest() == ref["md5"]


def test_c_driver_matches_golden_32(driver, golden, tmp_path):
    T = golden["solve8_32"]
    H = 34
    T3 = T.reshape(H, H, H)
    want = "".join("%f\n" % T3[k, j, i] for i in range(H) for j in range(H) for k in range(H))
    data = _run(driver, 32, 8, tmp_path / "d32.txt")
    assert data.decode() == want
