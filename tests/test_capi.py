"""CPU checks of the C ABI boundary: the library loads, exports every symbol
include/codag.h declares, and its status codes agree with the oracle's."""
import ctypes
import os
import re
import subprocess

import oracle
import codag_amd as ca

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "codag.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(codag_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 35
    out = subprocess.run(["nm", "-D", "--defined-only", ca.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (codag_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    L = ca.lib()
    for n in names:
        assert hasattr(L, n)


def test_status_codes_agree_with_oracle():
    assert (ca.CHUNK_OK, ca.CHUNK_TRUNCATED, ca.CHUNK_OVERFLOW, ca.CHUNK_CORRUPT) == \
        (oracle.OK, oracle.TRUNCATED, oracle.OVERFLOW, oracle.CORRUPT)
    hdr = open(os.path.join(ROOT, "include", "codag.h")).read()
    for name, val in [("CODAG_CHUNK_OK", 0), ("CODAG_CHUNK_TRUNCATED", 1),
                      ("CODAG_CHUNK_OVERFLOW", 2), ("CODAG_CHUNK_CORRUPT", 3),
                      ("CODAG_ERR_DECODE", 6)]:
        assert re.search(rf"{name}\s*=\s*{val}\b", hdr), name


def test_no_gpu_paths_fail_loudly():
    # without a GPU the context cannot be created; no silent CPU fallback exists
    if ca.device_count() > 0:
        return
    try:
        ca.Context(0)
    except ca.CodagError as e:
        assert e.status == ca.ERR_UNSUPPORTED
    else:
        raise AssertionError("context created without a GPU")


def test_status_strings():
    L = ca.lib()
    for s in range(10):
        assert L.codag_status_string(s)
    assert L.codag_version().startswith(b"codag")


def test_cli_usage_errors():
    for tool, args in [("rlev1", []), ("rlev2", []), ("zlib", ["a"])]:
        p = subprocess.run([os.path.join(ROOT, "codag_amd", "bin", tool)] + args,
                           capture_output=True, text=True)
        assert p.returncode != 0
