"""JSON scene manifests read as LoadSceneFile does (scene.go:56-133), host side.

Pins: float32 fields round the decimal once (strconv.ParseFloat(s, 32)), shown
by a decimal just above a float32 midpoint that float64 rounds onto the
midpoint; the degree conversion (Vec3.ToRadians, vector.go:82-85) against
exact rationals; Go's struct decoding rules (case-insensitive keys, the last
duplicate wins, unknown keys and trailing content ignored, short arrays
zero-padded, null elements as zero structs); every error message of
LoadSceneFile. Meshes come from OBJ files checked by test_assets.
"""
from decimal import Decimal, getcontext
from fractions import Fraction
import json

import numpy as np
import pytest

from gorender_amd import _lib
from gorender_amd import scenes as S
from gorender_amd.assets import ObjFile, SceneFile, SceneFileError, go_float32

import pngwrite

QUAD_OBJ = ("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\n"
            "vn 0 0 1\nf 1/1/1 2/2/1 3/3/1\nf 1/1/1 3/3/1 4/4/1\n")
TRI_OBJ = "v 0 0 0\nv 2 0 0\nv 0 2 1\nf 1 2 3\n"


def bits(x):
    return int(np.array(np.float32(x)).view(np.uint32))


def rn32(q: Fraction) -> np.float32:
    """Nearest-even float32 of an exact rational, by brute comparison."""
    f = np.float32(float(q))
    cands = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
    return min(cands, key=lambda c: (abs(Fraction(float(c)) - q), bits(c) & 1))


def test_go_float32_rounds_once():
    assert bits(go_float32("0.1")) == 0x3DCCCCCD
    assert bits(go_float32("3")) == 0x40400000
    assert bits(go_float32("-0")) == 0x80000000
    assert go_float32("1e-50") == 0
    # 1 + 2^-24 + 2^-60: float64 drops the 2^-60 and lands on the midpoint
    # between 1 and 1 + 2^-23, which then rounds to even (1.0); rounding the
    # decimal once gives 1 + 2^-23.
    getcontext().prec = 120
    x = Decimal(2 ** 60 + 2 ** 36 + 1) / Decimal(2 ** 60)
    assert bits(np.float32(float(str(x)))) == 0x3F800000
    assert bits(go_float32(str(x))) == 0x3F800001
    with pytest.raises(SceneFileError):
        go_float32("1e39")


def write_assets(tmp):
    d = tmp / "assets"
    d.mkdir()
    (d / "quad.obj").write_text(QUAD_OBJ)
    (d / "tri.obj").write_text(TRI_OBJ)
    rng = np.random.default_rng(3)
    (d / "t4.png").write_bytes(pngwrite.encode(rng.integers(0, 256, (4, 4, 4)), 6, 8))
    (d / "t35.png").write_bytes(pngwrite.encode(rng.integers(0, 256, (5, 3, 4)), 6, 8))
    return d


def manifest(tmp, doc, name="scene.json"):
    p = tmp / name
    p.write_text(doc if isinstance(doc, str) else json.dumps(doc))
    return str(p)


def test_scene_manifest(tmp_path):
    write_assets(tmp_path)
    doc = ('{"name": "demo", "unknown": [1, {"x": 2}],'
           ' "meshes": ['
           '  {"id": "q", "objFile": "assets/quad.obj", "texture": "assets/t4.png",'
           '   "textureScale": 2.5},'
           '  {"ID": "t", "objfile": "assets/../assets/tri.obj"},'
           '  {"id": "q3", "objFile": "assets/quad.obj", "texture": "assets/t35.png"}],'
           ' "objects": ['
           '  {"meshID": "q", "position": [1, 2, 3], "rotation": [90, 45.5, -30],'
           '   "scale": [1, 1, 1]},'
           '  {"MESHID": "t", "position": [1], "scale": [1, 1, 1], "scale": [2, 0.5, 3]},'
           '  {"meshID": "q3", "rotation": [180, 0, 0, 99], "scale": [1, 2, null]}]}'
           ' trailing text is not read')
    sf = SceneFile(manifest(tmp_path, doc))
    assert sf.name == "demo"
    assert len(sf.meshes) == 3
    quad = ObjFile(str(tmp_path / "assets" / "quad.obj"), single_mesh=True).meshes[0]
    np.testing.assert_array_equal(sf.meshes[0].vertices, quad.vertices)
    np.testing.assert_array_equal(sf.meshes[0].vertex_indices, quad.vertex_indices)
    # Textures in load order: the 4x4 file, the shared solid default, the 3x5 file.
    t = sf.textures
    assert [x.type for x in t] == [_lib.GR_TEXTURE_IMAGE_FAST, _lib.GR_TEXTURE_SOLID,
                                   _lib.GR_TEXTURE_IMAGE]
    assert t[0].scale == 2.5 and t[2].scale == 1.0
    assert t[1].color == (200, 200, 200, 255)
    assert (t[0].width, t[0].height, t[2].width, t[2].height) == (4, 4, 3, 5)
    assert [int(m.face_textures[0]) for m in sf.meshes] == [0, 1, 2]
    assert all((m.face_textures == m.face_textures[0]).all() for m in sf.meshes)
    o = sf.objects
    assert [x.mesh for x in o] == [0, 1, 2]
    np.testing.assert_array_equal(o[0].translation, [1, 2, 3])
    np.testing.assert_array_equal(o[1].translation, [1, 0, 0])
    np.testing.assert_array_equal(o[1].scale, np.float32([2, 0.5, 3]))
    np.testing.assert_array_equal(o[2].scale, [1, 2, 0])
    # ToRadians: v * RN32(pi32 / 180), each step rounded to float32.
    pi32 = Fraction(float(np.float32(np.pi)))
    f = Fraction(float(rn32(pi32 / 180)))
    want = [rn32(Fraction(v) * f) for v in (90, Fraction(91, 2), -30)]
    assert [bits(v) for v in o[0].rotation] == [bits(v) for v in want]
    assert bits(o[2].rotation[0]) == bits(rn32(180 * f))
    assert sf.warnings == []
    assert (sf.num_objects(), sf.num_triangles()) == (3, 2 + 1 + 2)
    assert sf.num_vertices() == 4 + 3 + 4


def test_scene_camera_matrices(tmp_path):
    write_assets(tmp_path)
    sf = SceneFile(manifest(tmp_path, {
        "meshes": [{"id": "t", "objFile": "assets/tri.obj"}],
        "objects": [{"meshID": "t", "rotation": [0, 30, 0], "scale": [1, 1, 1],
                     "position": [0, 0, -1]}]}))
    fr = S.FrameData(64, 48)
    sc = sf.scene(fr, (0, 0, 3), (0, 0, -1))
    o = sf.objects[0]
    ref = S.camera_object(fr, 0, (0, 0, 3), (0, 0, -1),
                          world=S.world_matrix(o.scale, o.rotation, o.translation))
    np.testing.assert_array_equal(sc.objects[0].mvp, ref.mvp)
    np.testing.assert_array_equal(sc.objects[0].world, ref.world)


def test_scene_zero_scale_and_duplicates(tmp_path):
    write_assets(tmp_path)
    sf = SceneFile(manifest(tmp_path, {
        "meshes": [{"id": "m", "objFile": "assets/quad.obj"},
                   {"id": "m", "objFile": "assets/tri.obj"}],
        "objects": [{"meshID": "m"}]}))
    # The later mesh entry owns the id; a missing scale is zero and only warns.
    assert sf.meshes[sf.objects[0].mesh].num_faces == 1
    assert sf.warnings == ["[WARN] object scale is zero: m"]
    assert SceneFile(manifest(tmp_path, "null", "null.json")).objects == []


@pytest.mark.parametrize("doc,msg", [
    ("{", "failed to read scene manifest"),
    ('{"meshes": [{"id": "a", "textureScale": NaN}]}', "failed to read scene manifest"),
    ('{"objects": [{"meshID": "a", "position": "up"}]}', "failed to read scene manifest"),
    ('{"objects": [{"meshID": 3}]}', "failed to read scene manifest"),
    ('{"meshes": {"id": "a"}}', "failed to read scene manifest"),
    ('{"objects": [{"meshID": "a", "scale": [1e39, 1, 1]}]}', "failed to read scene manifest"),
    ('{"objects": [{"meshID": "nope"}]}', "mesh id not found: nope"),
    ('{"meshes": [{"id": "p", "objFile": "assets/x.ply"}]}',
     "failed to load mesh 'p': unsupported mesh format: .ply"),
    ('{"meshes": [null]}', "failed to load mesh '': unsupported mesh format: "),
    ('{"meshes": [{"id": "g", "objFile": "assets/gone.obj"}]}', "failed to load mesh 'g'"),
    ('{"meshes": [{"id": "q", "objFile": "assets/quad.obj", "texture": "assets/gone.png"}]}',
     "failed to load texture q"),
])
def test_scene_errors(tmp_path, doc, msg):
    write_assets(tmp_path)
    with pytest.raises(SceneFileError) as e:
        SceneFile(manifest(tmp_path, doc))
    assert str(e.value).startswith(msg)


def test_scene_missing_file(tmp_path):
    with pytest.raises(SceneFileError) as e:
        SceneFile(str(tmp_path / "none.json"))
    assert str(e.value).startswith("failed to open level file")
