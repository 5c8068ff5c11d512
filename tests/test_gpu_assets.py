"""OBJ + MTL + PNG ingestion into the resident store (gr_obj_upload), then a
draw compared bit for bit with the oracle on the same parsed meshes/textures."""
import numpy as np
import pytest

from gorender_amd import scenes as S
from gorender_amd.assets import ObjFile
from oracle import refcpu
from parity import assert_same

import pngwrite

pytestmark = pytest.mark.gpu


def write_obj(path, mesh, mtl=None, material=None):
    """OBJ text with one 'vt' per face corner and f v/vt/vn faces."""
    lines = []
    if mtl:
        lines.append(f"mtllib {mtl}")
    lines.append("o Ico")
    for v in mesh.vertices:
        lines.append("v " + " ".join(f"{float(c):.9g}" for c in v[:3]))
    for n in mesh.vertex_normals:
        lines.append("vn " + " ".join(f"{float(c):.9g}" for c in n[:3]))
    for f in range(mesh.num_faces):
        for k in range(3):
            lines.append("vt " + " ".join(f"{float(c):.9g}" for c in mesh.uvs[f, k]))
    if material:
        lines.append(f"usemtl {material}")
    for f, (a, b, c) in enumerate(mesh.vertex_indices):
        t = 3 * f + 1
        lines.append(f"f {a + 1}/{t}/{a + 1} {b + 1}/{t + 1}/{b + 1} {c + 1}/{t + 2}/{c + 1}")
    path.write_text("\n".join(lines) + "\n")


def test_obj_upload_and_draw(ctx, tmp_path):
    mesh = S.icosphere(3)
    p = mesh.vertices[:, :3].astype(np.float64)
    u = (np.arctan2(p[:, 2], p[:, 0]) / (2 * np.pi) + 0.5).astype(np.float32)
    v = (np.arccos(np.clip(p[:, 1], -1, 1)) / np.pi).astype(np.float32)
    mesh.uvs = np.stack([u, v], -1)[mesh.vertex_indices]
    img = np.random.default_rng(7).integers(0, 256, (12, 20, 3))
    (tmp_path / "skin.png").write_bytes(pngwrite.encode(img, 2, 8, interlace=True))
    (tmp_path / "ico.mtl").write_text("newmtl skin\nmap_Kd skin.png\n")
    write_obj(tmp_path / "ico.obj", mesh, "ico.mtl", "skin")

    obj = ObjFile(str(tmp_path / "ico.obj"))
    handles = obj.upload(ctx)
    assert len(handles) == 1 and len(obj.textures) == 1
    fr = S.FrameData(160, 120)
    scene = S.SceneData(fr, obj.meshes, obj.textures,
                        [S.camera_object(fr, 0, (0.3, 0.2, 2.2), (-0.13, -0.09, -1.0))],
                        name="obj-ico3")
    got = ctx.draw(scene, handles)
    assert_same(scene, got, refcpu.draw(scene))
    assert got[2]["pixels_shaded"] > 1000


def test_scene_file_draw(ctx, tmp_path):
    """A JSON scene manifest (scene.go:56-133): two OBJ meshes, one with a
    PNG texture and scale, one on the default solid texture, three objects
    with rotations in degrees; drawn from a camera, bit for bit the oracle's."""
    import json
    from gorender_amd.assets import SceneFile
    ico = S.icosphere(2)
    p = ico.vertices[:, :3].astype(np.float64)
    u = (np.arctan2(p[:, 2], p[:, 0]) / (2 * np.pi) + 0.5).astype(np.float32)
    v = (np.arccos(np.clip(p[:, 1], -1, 1)) / np.pi).astype(np.float32)
    ico.uvs = np.stack([u, v], -1)[ico.vertex_indices]
    write_obj(tmp_path / "ico.obj", ico)
    cube = S.icosphere(1)
    cube.uvs = np.zeros((cube.num_faces, 3, 2), np.float32)
    write_obj(tmp_path / "ico1.obj", cube)
    img = np.random.default_rng(9).integers(0, 256, (8, 16, 4))
    (tmp_path / "skin.png").write_bytes(pngwrite.encode(img, 6, 8))
    (tmp_path / "scene.json").write_text(json.dumps({
        "name": "two-meshes",
        "meshes": [{"id": "ico", "objFile": "ico.obj", "texture": "skin.png",
                    "textureScale": 2},
                   {"id": "plain", "objFile": "ico1.obj"}],
        "objects": [{"meshID": "ico", "position": [-0.6, 0, 0], "rotation": [10, 35, 0],
                     "scale": [0.5, 0.5, 0.5]},
                    {"meshID": "plain", "position": [0.6, 0.1, -0.2], "rotation": [0, 0, 45],
                     "scale": [0.4, 0.6, 0.4]},
                    {"meshID": "ico", "position": [0, -0.5, -1], "rotation": [-20, 190, 5],
                     "scale": [0.7, 0.7, 0.7]}]}))
    sf = SceneFile(str(tmp_path / "scene.json"))
    fr = S.FrameData(200, 130)
    scene = sf.scene(fr, (0, 0.4, 3), (0, -0.13, -1))
    got = ctx.draw(scene)
    assert_same(scene, got, refcpu.draw(scene))
    assert got[2]["pixels_shaded"] > 2000
