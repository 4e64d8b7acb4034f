"""A scene loaded from a file renders like the same scene built in memory: the room scene written as glTF
(one mesh per object, float positions and normals, u32 indices, binary buffer) loads back into identical
arrays, and the path tracer's image of it equals the in-memory scene's bit for bit."""
import json
import struct

import numpy as np
import pytest

import vierkant_amd as vk
from vierkant_amd import workloads as W

pytestmark = pytest.mark.gpu


def write_glb(path, sc):
    """glTF 2.0 binary: one node, mesh and primitive per object, vertices of that object only"""
    pos, nrm, idx, obj = sc["positions"], sc["normals"], sc["indices"].reshape(-1, 3), sc["tri_object"]
    blob, views, accessors, meshes, nodes = b"", [], [], [], []

    def add(arr, ctype, typ, count):
        nonlocal blob
        views.append({"buffer": 0, "byteOffset": len(blob), "byteLength": arr.nbytes})
        accessors.append({"bufferView": len(views) - 1, "componentType": ctype, "count": count, "type": typ})
        blob += arr.tobytes()
        blob += b"\0" * (-len(blob) % 4)
        return len(accessors) - 1

    for o in range(len(sc["object_material"])):
        tris = idx[obj == o]
        lo, hi = int(tris.min()), int(tris.max())
        p = add(np.ascontiguousarray(pos[lo:hi + 1]), 5126, "VEC3", hi + 1 - lo)
        n = add(np.ascontiguousarray(nrm[lo:hi + 1]), 5126, "VEC3", hi + 1 - lo)
        i = add((tris - lo).astype(np.uint32).reshape(-1), 5125, "SCALAR", tris.size)
        meshes.append({"primitives": [{"attributes": {"POSITION": p, "NORMAL": n}, "indices": i,
                                       "material": int(sc["object_material"][o])}]})
        nodes.append({"mesh": o})
    mats = []
    for m in sc["materials"]:
        mats.append({"pbrMetallicRoughness": {"baseColorFactor": [float(x) for x in m["color"]],
                                              "metallicFactor": float(m["metalness"]),
                                              "roughnessFactor": float(m["roughness"])},
                     "emissiveFactor": [float(x) for x in m["emission"][:3]], "doubleSided": bool(m["two_sided"]),
                     "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": float(m["emission"][3])},
                                    "KHR_materials_transmission": {"transmissionFactor": float(m["transmission"])},
                                    "KHR_materials_ior": {"ior": float(m["ior"])}}})
    doc = {"asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": list(range(len(nodes)))}],
           "nodes": nodes, "meshes": meshes, "materials": mats, "buffers": [{"byteLength": len(blob)}],
           "bufferViews": views, "accessors": accessors}
    js = json.dumps(doc).encode()
    js += b" " * (-len(js) % 4)
    glb = struct.pack("<4sII", b"glTF", 2, 12 + 8 + len(js) + 8 + len(blob))
    glb += struct.pack("<I4s", len(js), b"JSON") + js + struct.pack("<I4s", len(blob), b"BIN\0") + blob
    path.write_bytes(glb)


def test_loaded_gltf_renders_like_the_scene_in_memory(ctx, tmp_path):
    sc = W.room_scene()
    write_glb(tmp_path / "room.glb", sc)
    m = vk.load_model(str(tmp_path / "room.glb"), environment=sc.get("environment", 1.0))
    for k in ("positions", "normals", "indices", "tri_object", "object_material"):
        assert np.array_equal(m[k], sc[k].reshape(m[k].shape)), k
    for f in ("color", "emission", "metalness", "roughness", "transmission", "ior", "two_sided"):
        assert np.array_equal(m["materials"][f], sc["materials"][f]), f
    images = []
    for scene in (sc, m):
        s = vk.default_settings()
        s.resolution[0], s.resolution[1] = 160, 90
        s.num_samples, s.max_trace_depth = 4, 8
        pt = vk.PathTracer(ctx, s, seed=0)
        pt.set_scene(scene)
        pt.render(W.room_camera())
        images.append(pt.pixels()["radiance"])
        pt.close()
    assert np.array_equal(images[0], images[1])
