"""C ABI: the library builds, loads and exports every symbol its header declares (CPU)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "banshee_hip.h")).read()
    return sorted(set(re.findall(r"^(?:bsh_status|const char\*|size_t)\s+(bsh_\w+)\(", src, re.M)))


def test_header_matches_binding_list():
    from banshee_amd import hip
    assert declared_symbols() == sorted(hip.EXPORTS)


def test_library_exports_every_declared_symbol():
    from banshee_amd import hip
    if not os.path.exists(hip.LIB_PATH):
        pytest.skip("library not built")
    out = subprocess.run(["nm", "-D", "--defined-only", hip.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (bsh_\w+)", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing
    hip.lib()


def test_fails_loudly_without_device():
    from banshee_amd import hip
    if not os.path.exists(hip.LIB_PATH):
        pytest.skip("library not built")
    try:
        c = hip.Context(0)
    except hip.BshError as e:
        assert "device" in str(e)
    else:
        c.close()   # a GPU is present: creation succeeded
