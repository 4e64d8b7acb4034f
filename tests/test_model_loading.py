"""Scene files (vkrt_model_load): Wavefront OBJ + MTL and glTF 2.0 (.gltf with base64 and external buffers,
.glb), against values worked out here from the file contents and the reference's mapping to geometry and
material_t (src/model/wavefront_obj.cpp:13-137, src/model/gltf.cpp:136-660, src/Geometry.cpp:106-137)."""
import base64
import json
import struct

import numpy as np
import pytest

import vierkant_amd as vk


def write(path, text):
    path.write_text(text)
    return str(path)


def test_obj_quad_fan_computed_normals_and_default_material(tmp_path):
    """a quad fans into (0,1,2), (0,2,3) with one vertex per corner; without vn lines the normals are the
    normalized face normals; without materials the scene gets material_t's defaults"""
    p = write(tmp_path / "quad.obj", "v 0 0 0\nv 2 0 0\nv 2 1 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\n"
                                    "f 1/1 2/2 3/3 4/4\n")
    m = vk.load_model(p)
    want = np.array([[0, 0, 0], [2, 0, 0], [2, 1, 0], [0, 0, 0], [2, 1, 0], [0, 1, 0]], np.float32)
    assert np.array_equal(m["positions"], want)
    assert np.array_equal(m["indices"], np.arange(6, dtype=np.uint32))
    assert np.array_equal(m["normals"], np.tile(np.float32([0, 0, 1]), (6, 1)))
    assert np.array_equal(m["texcoords"], np.float32([[0, 0], [1, 0], [1, 1], [0, 0], [1, 1], [0, 1]]))
    assert np.array_equal(m["tri_object"], [0, 0]) and np.array_equal(m["object_material"], [0])
    mat = m["materials"][0]
    assert np.array_equal(mat["color"], [1, 1, 1, 1]) and mat["roughness"] == 1 and mat["metalness"] == 0
    assert mat["ior"] == np.float32(1.5) and mat["emission"][3] == 1


def test_obj_shapes_negative_indices_normals_and_mtl(tmp_path):
    """'o' starts a shape; negative indices count back from the last vertex; vn lines are used as given;
    each shape takes its first face's material; MTL maps as the reference does: alpha = clamp(1 - d),
    roughness = clamp(max(Pr, (1 - Ns)^2), 0, 1), metalness Pm, ior Ni, emission Ke, clearcoat roughness Pcr"""
    write(tmp_path / "m.mtl", "newmtl red\nKd 0.8 0.1 0.1\nd 0.25\nNs 0.5\nNi 1.33\nPm 0.75\nKe 1 2 3\nPcr 0.4\n"
                              "newmtl rough\nKd 0.2 0.3 0.4\nPr 0.9\nNs 0.0\n")
    p = write(tmp_path / "s.obj", "mtllib m.mtl\n"
                                  "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nvn 0 0.6 0.8\n"
                                  "o first\nusemtl rough\nf 1//1 2//1 3//2\n"
                                  "o second\nusemtl red\nv 0 0 1\nv 1 0 1\nv 0 1 1\nf -3//1 -2//2 -1//2\n"
                                  "usemtl rough\nf -1//1 -2//1 -3//1\n")
    m = vk.load_model(p)
    assert len(m["object_material"]) == 2 and np.array_equal(m["tri_object"], [0, 1, 1])
    assert np.array_equal(m["object_material"], [1, 0])      # 'rough' is material 1, 'red' material 0
    assert np.array_equal(m["positions"][3:6], np.float32([[0, 0, 1], [1, 0, 1], [0, 1, 1]]))
    assert np.array_equal(m["positions"][6:9], np.float32([[0, 1, 1], [1, 0, 1], [0, 0, 1]]))
    assert np.array_equal(m["normals"][2], np.float32([0, 0.6, 0.8]))
    red, rough = m["materials"][0], m["materials"][1]
    assert np.array_equal(red["color"], np.float32([0.8, 0.1, 0.1, 0.75]))
    assert red["roughness"] == np.float32(0.25)                  # (1 - 0.5)^2
    assert red["metalness"] == np.float32(0.75) and red["ior"] == np.float32(1.33)
    assert np.array_equal(red["emission"], np.float32([1, 2, 3, 1])) and red["clearcoat_roughness"] == np.float32(0.4)
    assert rough["roughness"] == 1.0 and rough["color"][3] == 0.0  # max(0.9, 1) clamped; d = 1 -> alpha 0
    assert rough["ior"] == 1.0                                     # tinyobj's default Ni


def test_obj_computed_normals_follow_compute_vertex_normals(tmp_path):
    """per-corner vertices: each normal is its face's normalize(cross(b - a, c - a)), computed in float32"""
    rng = np.random.default_rng(1)
    v = rng.uniform(-1, 1, (30, 3)).astype(np.float32)
    lines = [f"v {x!r} {y!r} {z!r}" for x, y, z in v.tolist()]
    faces = np.stack([rng.choice(30, 3, replace=False) + 1 for _ in range(20)])
    lines += [f"f {a} {b} {c}" for a, b, c in faces]
    m = vk.load_model(write(tmp_path / "r.obj", "\n".join(lines) + "\n"))
    p = m["positions"].reshape(-1, 3, 3)
    assert np.array_equal(p, v[faces - 1])
    e1, e2 = p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]
    c = np.stack([e1[:, 1] * e2[:, 2] - e2[:, 1] * e1[:, 2], e1[:, 2] * e2[:, 0] - e2[:, 2] * e1[:, 0],
                  e1[:, 0] * e2[:, 1] - e2[:, 0] * e1[:, 1]], -1).astype(np.float32)
    n = c * (np.float32(1) / np.sqrt(((c[:, 0] * c[:, 0] + c[:, 1] * c[:, 1]) + c[:, 2] * c[:, 2])))[:, None]
    n = n.astype(np.float32)
    # a normalized vector normalized again moves by at most an ulp
    assert np.allclose(m["normals"].reshape(-1, 3, 3), n[:, None, :], rtol=0, atol=2e-7)


def gltf_doc(buffer_uri=None, byte_length=0):
    """one triangle mesh (u16 indices, float positions and normals) used by two nodes, one with TRS and a
    child, and materials with KHR extensions; the second primitive has no material"""
    return {
        "asset": {"version": "2.0"},
        "scene": 0,
        "scenes": [{"nodes": [0, 2]}],
        "nodes": [{"mesh": 0, "translation": [1, 2, 3], "rotation": [0, 0, 0.7071067811865476, 0.7071067811865476],
                   "scale": [2, 2, 2], "children": [1]},
                  {"mesh": 0, "translation": [0, 0, 1]},
                  {"mesh": 1, "matrix": [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 5, 0, 0, 1]}],
        "meshes": [{"name": "tri", "primitives": [{"attributes": {"POSITION": 0, "NORMAL": 1}, "indices": 2,
                                                    "material": 1}]},
                   {"name": "bare", "primitives": [{"attributes": {"POSITION": 0}}]}],
        "materials": [{"name": "plain"},
                      {"name": "glass", "pbrMetallicRoughness": {"baseColorFactor": [0.5, 0.6, 0.7, 0.8],
                                                                 "metallicFactor": 0.25, "roughnessFactor": 0.125},
                       "emissiveFactor": [0.1, 0.2, 0.3], "doubleSided": True,
                       "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 4.0},
                                      "KHR_materials_transmission": {"transmissionFactor": 0.9},
                                      "KHR_materials_ior": {"ior": 1.45},
                                      "KHR_materials_clearcoat": {"clearcoatFactor": 0.3,
                                                                  "clearcoatRoughnessFactor": 0.2},
                                      "KHR_materials_sheen": {"sheenColorFactor": [0.1, 0.2, 0.3],
                                                              "sheenRoughnessFactor": 0.4},
                                      "KHR_materials_volume": {"thicknessFactor": 0.5,
                                                               "attenuationColor": [0.25, 0.0, 1.0],
                                                               "attenuationDistance": 2.5}}}],
        "buffers": [dict({"byteLength": byte_length}, **({"uri": buffer_uri} if buffer_uri else {}))],
        "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": 36},
                        {"buffer": 0, "byteOffset": 36, "byteLength": 36},
                        {"buffer": 0, "byteOffset": 72, "byteLength": 6}],
        "accessors": [{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3"},
                      {"bufferView": 1, "componentType": 5126, "count": 3, "type": "VEC3"},
                      {"bufferView": 2, "componentType": 5123, "count": 3, "type": "SCALAR"}],
    }


TRI = np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0]])
NRM = np.float32([[0, 0, 1], [0, 0, 1], [0, 0, 1]])


def gltf_bin():
    b = TRI.tobytes() + NRM.tobytes() + np.uint16([0, 1, 2]).tobytes()
    return b + b"\0" * (-len(b) % 4)


def check_gltf_scene(m):
    # node 0: T(1,2,3) R(90 deg about z) S(2): (x, y, z) -> (1 - 2y, 2 + 2x, 3 + 2z); node 1 adds z + 1 before
    r = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    node0 = (TRI.astype(np.float64) * 2) @ r.T + [1, 2, 3]
    node1 = ((TRI.astype(np.float64) + [0, 0, 1]) * 2) @ r.T + [1, 2, 3]
    node2 = TRI.astype(np.float64) + [5, 0, 0]
    want = np.concatenate([node0, node1, node2]).astype(np.float32)
    assert np.allclose(m["positions"], want, rtol=0, atol=1e-6)
    assert np.allclose(m["normals"][:6], np.tile(np.float32([0, 0, 1]), (6, 1)), atol=1e-7)
    assert np.array_equal(m["normals"][6:], np.tile(np.float32([0, 0, 1]), (3, 1)))  # computed
    assert np.array_equal(m["indices"], np.arange(9, dtype=np.uint32))
    assert np.array_equal(m["tri_object"], [0, 1, 2])
    assert np.array_equal(m["object_material"], [1, 1, 2])      # the bare primitive gets the appended default
    plain, glass, default = m["materials"]
    assert np.array_equal(plain["color"], [1, 1, 1, 1]) and plain["metalness"] == 1.0  # glTF's default factor
    assert np.array_equal(glass["color"], np.float32([0.5, 0.6, 0.7, 0.8]))
    assert glass["metalness"] == np.float32(0.25) and glass["roughness"] == np.float32(0.125)
    assert np.array_equal(glass["emission"], np.float32([0.1, 0.2, 0.3, 4.0]))
    assert glass["transmission"] == np.float32(0.9) and glass["ior"] == np.float32(1.45)
    assert glass["clearcoat"] == np.float32(0.3) and glass["clearcoat_roughness"] == np.float32(0.2)
    assert np.array_equal(glass["sheen_color"][:3], np.float32([0.1, 0.2, 0.3]))
    assert glass["sheen_roughness"] == np.float32(0.4) and glass["two_sided"] == 1
    assert default["metalness"] == 0.0 and default["ior"] == np.float32(1.5)
    # KHR_materials_volume: the attenuation, a zero channel raised to 1e-6; no extension: no medium
    assert glass["attenuation_distance"] == np.float32(2.5)
    assert np.array_equal(glass["attenuation_color"], np.float32([0.25, 1e-6, 1.0]))
    assert plain["attenuation_distance"] == 0.0 and default["attenuation_distance"] == 0.0


def test_gltf_base64_buffer(tmp_path):
    data = gltf_bin()
    uri = "data:application/octet-stream;base64," + base64.b64encode(data).decode()
    p = write(tmp_path / "s.gltf", json.dumps(gltf_doc(uri, len(data))))
    check_gltf_scene(vk.load_model(p))


def test_gltf_external_buffer_and_glb_agree(tmp_path):
    data = gltf_bin()
    (tmp_path / "s.bin").write_bytes(data)
    a = vk.load_model(write(tmp_path / "s.gltf", json.dumps(gltf_doc("s.bin", len(data)))))
    check_gltf_scene(a)
    js = json.dumps(gltf_doc(None, len(data))).encode()
    js += b" " * (-len(js) % 4)
    glb = struct.pack("<4sII", b"glTF", 2, 12 + 8 + len(js) + 8 + len(data))
    glb += struct.pack("<I4s", len(js), b"JSON") + js + struct.pack("<I4s", len(data), b"BIN\0") + data
    (tmp_path / "s.glb").write_bytes(glb)
    b = vk.load_model(str(tmp_path / "s.glb"))
    for k in ("positions", "normals", "indices", "tri_object", "object_material"):
        assert np.array_equal(a[k], b[k]), k
    assert a["materials"].tobytes() == b["materials"].tobytes()


def test_load_errors_carry_a_message(tmp_path):
    with pytest.raises(vk.VkrtError, match="cannot open"):
        vk.load_model(str(tmp_path / "missing.obj"))
    with pytest.raises(vk.VkrtError, match="out of range"):
        vk.load_model(write(tmp_path / "bad.obj", "v 0 0 0\nf 1 2 3\n"))
    with pytest.raises(vk.VkrtError, match="JSON"):
        vk.load_model(write(tmp_path / "bad.gltf", "{\"asset\": "))
    with pytest.raises(vk.VkrtError, match="unknown model format"):
        vk.load_model(write(tmp_path / "x.fbx", ""))


@pytest.mark.parametrize("name, edit, message", [
    ("negative count", lambda d: d["accessors"][0].update(count=-1), "invalid accessor count"),
    ("fractional count", lambda d: d["accessors"][0].update(count=2.5), "invalid accessor count"),
    ("count past the view", lambda d: d["accessors"][0].update(count=4), "past its bufferView"),
    ("huge count", lambda d: d["accessors"][0].update(count=2 ** 40), "count above"),
    ("offset past the view", lambda d: d["accessors"][0].update(byteOffset=28), "past its bufferView"),
    ("negative offset", lambda d: d["accessors"][0].update(byteOffset=-12), "invalid accessor byteOffset"),
    ("stride below the element", lambda d: d["bufferViews"][0].update(byteStride=8), "byteStride below"),
    ("stride overflowing", lambda d: d["bufferViews"][0].update(byteStride=2 ** 39), "past its bufferView"),
    ("view past the buffer", lambda d: d["bufferViews"][2].update(byteLength=4096), "past its buffer"),
    ("view offset past the buffer", lambda d: d["bufferViews"][1].update(byteOffset=2 ** 36), "past its buffer"),
    ("negative view length", lambda d: d["bufferViews"][1].update(byteLength=-36),
     "invalid bufferView byteLength"),
    ("missing buffer", lambda d: d["bufferViews"][0].update(buffer=3), "buffer out of range"),
    ("fractional buffer index", lambda d: d["bufferViews"][0].update(buffer=0.5), "buffer out of range"),
    ("huge view index", lambda d: d["accessors"][0].update(bufferView=1e300), "bufferView out of range"),
])
def test_gltf_malformed_accessors_are_rejected(tmp_path, name, edit, message):
    """accessor and bufferView bounds: negative, fractional and oversized sizes, strides below the element,
    ranges past the bufferView or the buffer (checked without overflow) fail with a message, never read out
    of bounds"""
    data = gltf_bin()
    (tmp_path / "s.bin").write_bytes(data)
    doc = gltf_doc("s.bin", len(data))
    edit(doc)
    with pytest.raises(vk.VkrtError, match=message):
        vk.load_model(write(tmp_path / "s.gltf", json.dumps(doc)))


def test_gltf_strided_view_within_bounds(tmp_path):
    """a bufferView holding positions and normals interleaved (byteStride 24) reads both correctly, and an
    accessor ending exactly at the view's end is accepted"""
    inter = np.concatenate([TRI, NRM], axis=1).astype(np.float32)
    data = inter.tobytes() + np.uint16([0, 1, 2]).tobytes()
    data += b"\0" * (-len(data) % 4)
    (tmp_path / "s.bin").write_bytes(data)
    doc = gltf_doc("s.bin", len(data))
    doc["bufferViews"] = [{"buffer": 0, "byteOffset": 0, "byteLength": 72, "byteStride": 24},
                          {"buffer": 0, "byteOffset": 72, "byteLength": 6}]
    doc["accessors"][0].update(bufferView=0, byteOffset=0)
    doc["accessors"][1].update(bufferView=0, byteOffset=12)
    doc["accessors"][2].update(bufferView=1)
    check_gltf_scene(vk.load_model(write(tmp_path / "s.gltf", json.dumps(doc))))


def test_obj_round_trip_of_the_room_scene(tmp_path):
    """the room scene written as OBJ (shortest float repr, per-vertex normals, one shape and material per
    object) loads back with the same triangles, normals and object grouping"""
    from vierkant_amd import workloads as W
    sc = W.room_scene(n_spheres=12)
    pos, nrm, idx = sc["positions"], sc["normals"], sc["indices"].reshape(-1, 3)
    lines = ["mtllib room.mtl"]
    lines += [f"v {x!r} {y!r} {z!r}" for x, y, z in pos.tolist()]
    lines += [f"vn {x!r} {y!r} {z!r}" for x, y, z in nrm.tolist()]
    obj = sc["tri_object"]
    for o in range(len(sc["object_material"])):
        lines.append(f"o obj{o}")
        lines.append(f"usemtl m{sc['object_material'][o]}")
        for a, b, c in (idx[obj == o] + 1).tolist():
            lines.append(f"f {a}//{a} {b}//{b} {c}//{c}")
    mtl = []
    for k, mat in enumerate(sc["materials"]):
        mtl += [f"newmtl m{k}", "Kd {!r} {!r} {!r}".format(*mat["color"][:3].tolist()), "d 1", "Ns 0"]
    (tmp_path / "room.mtl").write_text("\n".join(mtl) + "\n")
    m = vk.load_model(write(tmp_path / "room.obj", "\n".join(lines) + "\n"))
    order = np.concatenate([np.nonzero(obj == o)[0] for o in range(len(sc["object_material"]))])
    assert np.array_equal(m["positions"].reshape(-1, 3, 3), pos[idx[order]])
    assert np.array_equal(m["normals"].reshape(-1, 3, 3), nrm[idx[order]])
    assert np.array_equal(m["tri_object"], obj[order])
    assert np.array_equal(m["object_material"], sc["object_material"])
    assert np.array_equal(m["materials"]["color"][:, :3], sc["materials"]["color"][:, :3])
