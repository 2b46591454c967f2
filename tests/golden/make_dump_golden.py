"""Write tests/golden/ref_dump_16.json: digests of the reference's own solution
dump (its main() built with PRI at 16^3, AN_ISO_1_OpenACC.c:1244-1260), so the
text writer can be checked without the reference tree.

Run from the repository root after `make -C oracle ref`:
    python tests/golden/make_dump_golden.py
"""
import ctypes as C
import hashlib
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(REPO, "oracle", "_ref", "libfoldref_pri_16.so")


def main() -> None:
    if not os.path.exists(LIB):
        sys.exit(f"{LIB} is not built: make -C oracle ref")
    with tempfile.TemporaryDirectory() as d:
        # main() writes into the working directory; run it in a child so the
        # directory change and its stdout stay there
        code = ("import ctypes, os; os.chdir(%r); L = ctypes.CDLL(%r); "
                "L.ref_main.restype = ctypes.c_int; raise SystemExit(L.ref_main())" % (d, LIB))
        subprocess.run([sys.executable, "-c", code], check=True, stdout=subprocess.DEVNULL)
        data = open(os.path.join(d, "an_iso_1_openacc.txt"), "rb").read()
    out = {"n": 16, "iterations": 8, "lines": data.count(b"\n"),
           "md5": hashlib.md5(data).hexdigest(), "sha256": hashlib.sha256(data).hexdigest(),
           "first_lines": [x.decode() for x in data.split(b"\n")[:4]],
           "source": "Stencil-OpenACC/AN_ISO_1_OpenACC.c main() with -DPRI, 16^3"}
    with open(os.path.join(HERE, "ref_dump_16.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out)


if __name__ == "__main__":
    main()
