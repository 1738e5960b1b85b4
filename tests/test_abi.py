"""C-ABI checks that need no GPU: the library loads, exports every function
include/tomhip.h declares, ctypes signatures match, MEX gateways compile."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "tomhip.h")


def declared():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tomhip_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import tom_amd
    lib = tom_amd.load()  # raises if missing: no CPU fallback
    names = declared()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", tom_amd.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(lib, n)


def test_ctypes_table_covers_header():
    from tom_amd._lib import SIGNATURES
    assert sorted(SIGNATURES) == declared()


def test_errors_cross_the_abi_as_codes():
    """No GPU here: init must fail with a HIP error code, not crash."""
    import ctypes

    import tom_amd
    lib = tom_amd.load()
    h = ctypes.c_void_p()
    rc = lib.tomhip_init(0, ctypes.byref(h))
    if rc == 0:
        lib.tomhip_destroy(h)
        pytest.skip("a GPU is present")
    assert rc < 0
    assert lib.tomhip_last_error(None)


def test_mex_gateways_compile():
    for f in sorted(os.listdir(os.path.join(ROOT, "tom_amd", "mex"))):
        if f.endswith(".c"):
            subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-I",
                            os.path.join(ROOT, "tests", "mex_stub"), "-I", os.path.join(ROOT, "include"),
                            os.path.join(ROOT, "tom_amd", "mex", f)], check=True)


def test_no_oracle_in_product():
    """The product package never imports or links the oracle."""
    for dp, _, fs in os.walk(os.path.join(ROOT, "tom_amd")):
        for f in fs:
            if f.endswith((".py", ".hip", ".h", ".c", ".cpp")):
                txt = open(os.path.join(dp, f)).read()
                assert "oracle" not in txt.replace("oracles", ""), os.path.join(dp, f)
