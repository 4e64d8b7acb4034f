"""The C++ host layer (vierkant/intersection.hpp, RayBuilder.hpp, PBRPathTracer.hpp) end to end: the
compiled program tests/cpp/test_host_api.cpp checks GPU traversal against the scalar vierkant::intersect
loop bit for bit and the path tracer's batch semantics."""
import pathlib
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = pathlib.Path(__file__).resolve().parent / "cpp" / "test_host_api"


def test_cpp_host_layer():
    if not EXE.exists():
        pytest.fail("tests/cpp/test_host_api missing: run __graft_entry__.build()")
    res = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "host api ok" in res.stdout
