"""The C-ABI library loads, exports every entry point include/vkrt.h declares, and its host-only
geometry functions reproduce the reference generators. No GPU computation happens here."""
import ctypes
import pathlib
import re

import numpy as np
import pytest

import oracle_binding as orc
import vierkant_amd as vk

HEADER = pathlib.Path(__file__).resolve().parent.parent / "include" / "vkrt.h"


def declared_in_header():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(vkrt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(vk.LIB_PATH))
    names = declared_in_header()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_covers_header():
    assert set(declared_in_header()) == set(vk.declared_symbols())


def test_version_and_no_device_behaviour():
    lib = vk.lib()
    assert lib.vkrt_version() == 1
    if lib.vkrt_device_count() == 0:
        h = ctypes.c_void_p()
        assert lib.vkrt_create(0, None, ctypes.byref(h)) == vk.E_NO_DEVICE
        with pytest.raises(vk.VkrtError):
            vk.Context(0)


def test_null_context_is_an_error():
    lib = vk.lib()
    assert lib.vkrt_sync(None) == vk.E_INVALID_ARG
    assert lib.vkrt_last_error(None) == b"null context"


@pytest.mark.parametrize("subdiv", [0, 1, 3, 5])
@pytest.mark.parametrize("radius", [1.0, 0.037])
def test_icosphere_bitwise_equals_oracle(subdiv, radius):
    pos, idx = vk.icosphere(radius, subdiv)
    opos, oidx = orc.icosphere(radius, subdiv)
    assert np.array_equal(idx, oidx)
    assert pos.tobytes() == opos.tobytes()


def test_plane_bitwise_equals_oracle():
    pos, idx = vk.plane(10.0, 10.0, 512, 512)
    opos, oidx = orc.plane(10.0, 10.0, 512, 512)
    assert np.array_equal(idx, oidx) and pos.tobytes() == opos.tobytes()
    assert len(idx) // 3 == 524288


def test_heightfield_spheres_scene_counts():
    from vierkant_amd import workloads as W
    pos, idx = W.heightfield_spheres_scene(grid=64, n_spheres=20)
    assert len(idx) // 3 == 64 * 64 * 2 + 20 * 1280
    assert idx.max() < len(pos)
