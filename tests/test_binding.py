"""The C++ binding a reference host uses (include/idx2g_idx2.hpp), compiled
against the reference's own headers and objects: error codes equal the
reference's enum values, and parameter validation returns the same error as
the reference's Finalize (Source/Core/idx2Common.cpp:358-754).  Needs the
reference sources and their objects (oracle/ref.mk); skipped elsewhere."""
import glob
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/Source"
OBJ = os.path.join(ROOT, "oracle", "_ref", "obj")
LIBDIR = os.path.join(ROOT, "idx2_amd", "_build")


def _objects():
    return sorted(glob.glob(os.path.join(OBJ, "*.o")))


@pytest.mark.skipif(not os.path.isdir(REF) or not os.path.isdir(OBJ)
                    or not os.path.exists(os.path.join(LIBDIR, "libidx2g.so")),
                    reason="reference sources/objects or libidx2g.so not present")
def test_binding_matches_reference_finalize():
    src = os.path.join(ROOT, "tests", "binding", "check_validate.cpp")
    with tempfile.TemporaryDirectory() as wd:
        exe = os.path.join(wd, "check_validate")
        cmd = ["g++", "-std=c++17", "-O1", "-w", "-fopenmp", "-mavx2", "-Didx2_Avx2", "-I" + REF,
               "-I" + os.path.join(REF, "Core"), "-I" + os.path.join(ROOT, "include"), src, *_objects(),
               "-L" + LIBDIR, "-lidx2g", "-Wl,-rpath," + LIBDIR, "-lpthread", "-o", exe]
        subprocess.run(cmd, check=True, timeout=300)
        r = subprocess.run([exe, "1500"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 mismatches" in r.stdout
