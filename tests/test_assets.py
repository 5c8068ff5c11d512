"""Asset ingestion (host side, no GPU): OBJ/MTL parsing as LoadObjFile does it
(obj.go:42-309) and PNG decoding as LoadTextureFile does it (texture.go:35-63,
91-103: image/png, then color.RGBAModel).

Pins: a PNG encoder written for these tests (tests/pngwrite.py) round-trips
every colour type, bit depth, filter type and Adam7; colour conversion checks
use closed-form values of Go's NRGBA premultiplication; OBJ cases are small
hand-written files whose meshes are written out in full; the reference's own
model files (models/cube.obj + cube.mtl + textures-16.png, models/suzanne.obj)
are parsed when the reference tree is present, with counts taken from the text.
"""
import os

import numpy as np
import pytest

from gorender_amd import _lib
from gorender_amd.assets import ObjFile, png_decode

import pngwrite

REF_MODELS = "/root/reference/models"


# ---- PNG ------------------------------------------------------------------

def premul8(v, a):
    """color.NRGBA.RGBA then rgbaModel's >> 8 (image/color)."""
    v = np.asarray(v, np.int64)
    a = np.asarray(a, np.int64)
    return (((v | (v << 8)) * a) // 0xFF) >> 8


def premul16(v, a):
    return ((np.asarray(v, np.int64) * np.asarray(a, np.int64)) // 0xFFFF) >> 8


def test_premultiply_closed_forms():
    assert premul8(255, 128) == 128 and premul8(255, 255) == 255 and premul8(200, 0) == 0
    assert premul8(1, 255) == 1 and premul8(128, 128) == 64
    assert premul16(0xFFFF, 0x8000) == 0x80 and premul16(0x1234, 0xFFFF) == 0x12


def expected_rgba(samples, ct, depth, palette=None, trns=None):
    s = np.asarray(samples, np.int64)
    h, w, _ = s.shape
    out = np.zeros((h, w, 4), np.int64)
    if ct == 0:
        y = s[..., 0]
        y8 = y >> 8 if depth == 16 else (y * (255 // ((1 << depth) - 1)) if depth < 8 else y)
        out[..., 0] = out[..., 1] = out[..., 2] = y8
        out[..., 3] = 255
        if trns is not None:
            t = trns[0] << 8 | trns[1]
            out[y == t] = 0
    elif ct == 2:
        rgb = s >> 8 if depth == 16 else s
        out[..., :3] = rgb
        out[..., 3] = 255
        if trns is not None:
            t = [trns[0] << 8 | trns[1], trns[2] << 8 | trns[3], trns[4] << 8 | trns[5]]
            out[np.all(s == t, axis=-1)] = 0
    elif ct == 3:
        pal = np.asarray(palette, np.int64)
        idx = s[..., 0]
        for i in range(h):
            for j in range(w):
                k = idx[i, j]
                if k >= len(pal):
                    out[i, j] = (0, 0, 0, 255)
                elif trns is not None and k < len(trns):
                    a = trns[k]
                    out[i, j, :3] = premul8(pal[k], a)
                    out[i, j, 3] = a
                else:
                    out[i, j, :3] = pal[k]
                    out[i, j, 3] = 255
    elif ct == 4:
        y, a = s[..., 0], s[..., 1]
        for c in range(3):
            out[..., c] = premul16(y, a) if depth == 16 else premul8(y, a)
        out[..., 3] = a >> 8 if depth == 16 else a
    else:
        a = s[..., 3]
        for c in range(3):
            out[..., c] = premul16(s[..., c], a) if depth == 16 else premul8(s[..., c], a)
        out[..., 3] = a >> 8 if depth == 16 else a
    return out.astype(np.uint8)


CASES = [(0, 1), (0, 2), (0, 4), (0, 8), (0, 16), (2, 8), (2, 16), (3, 1), (3, 2), (3, 4),
         (3, 8), (4, 8), (4, 16), (6, 8), (6, 16)]


@pytest.mark.parametrize("interlace", [False, True])
@pytest.mark.parametrize("ct,depth", CASES)
def test_png_round_trip(tmp_path, ct, depth, interlace):
    rng = np.random.default_rng(ct * 100 + depth + interlace)
    h, w = 13, 11  # odd sizes exercise every Adam7 pass and partial bytes
    ch = pngwrite.CHANNELS[ct]
    palette = trns = None
    if ct == 3:
        npal = min(1 << depth, 6)
        palette = rng.integers(0, 256, (npal, 3))
        hi = 1 << depth  # indexes past the palette read opaque black
        samples = rng.integers(0, hi, (h, w, 1))
        trns = list(rng.integers(0, 256, min(npal, 3)))
    else:
        samples = rng.integers(0, 1 << depth, (h, w, ch))
        if depth == 8 and ct in (0, 2):
            # some pixels equal to the transparent colour
            trns_val = samples[0, 0]
            trns = [b for v in trns_val for b in (0, int(v))]
    data = pngwrite.encode(samples, ct, depth, palette, trns, interlace, seed=depth)
    p = tmp_path / "t.png"
    p.write_bytes(data)
    got = png_decode(str(p))
    assert got.shape == (h, w, 4)
    np.testing.assert_array_equal(got, expected_rgba(samples, ct, depth, palette, trns))


def test_png_errors(tmp_path):
    p = tmp_path / "x.png"
    good = pngwrite.encode(np.zeros((2, 2, 4), int), 6, 8)
    p.write_bytes(good[:-20])  # truncated
    with pytest.raises(_lib.GorenderError):
        png_decode(str(p))
    bad = bytearray(good)
    bad[30] ^= 0xFF  # IHDR checksum no longer matches
    p.write_bytes(bytes(bad))
    with pytest.raises(_lib.GorenderError):
        png_decode(str(p))
    p.write_bytes(b"\xff\xd8\xff\xe0" + b"\0" * 16)  # a JPEG header (image/jpeg)
    with pytest.raises(_lib.GorenderError) as e:
        png_decode(str(p))
    assert e.value.status == _lib.GR_ERR_UNSUPPORTED
    with pytest.raises(_lib.GorenderError):
        png_decode(str(tmp_path / "missing.png"))


# ---- OBJ / MTL --------------------------------------------------------------

def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_obj_face_formats_and_offsets(tmp_path):
    """Each face format of parseFace, two 'o' groups with index offsets
    (ObjContext.Clear, obj.go:31-40), UVs resolved while parsing."""
    text = """# comment
o A
v 0 0 0
v 1 0 0
v 0 1 0
v 0 0 1
vt 0.25 0.5
vt 0.75 0.125
vn 0 0 1
vn 0 1 0
f 1 2 3
f 1/1 2/2 3/1
f 2/2/2 3/1/1 4/2/1
o B
v 5 5 5
v 6 5 5
v 5 6 5
vt 1 1
vn 1 0 0
f 5/3/3 6/3/3 7/3/3
f 5 7 6
"""
    obj = ObjFile(write(tmp_path, "m.obj", text))
    assert len(obj.meshes) == 2 and obj.textures == []
    a, b = obj.meshes
    np.testing.assert_array_equal(a.vertices, [[0, 0, 0, 1], [1, 0, 0, 1], [0, 1, 0, 1],
                                               [0, 0, 1, 1]])
    np.testing.assert_array_equal(a.vertex_indices, [[0, 1, 2], [0, 1, 2], [1, 2, 3]])
    np.testing.assert_array_equal(a.normal_indices, [[0, 0, 0], [0, 0, 0], [1, 0, 0]])
    np.testing.assert_array_equal(a.uvs, [[[0, 0], [0, 0], [0, 0]],
                                          [[0.25, 0.5], [0.75, 0.125], [0.25, 0.5]],
                                          [[0.75, 0.125], [0.25, 0.5], [0.75, 0.125]]])
    np.testing.assert_array_equal(a.vertex_normals, [[0, 0, 1, 1], [0, 1, 0, 1]])
    np.testing.assert_array_equal(a.face_textures, [-1, -1, -1])
    # second group: indices relative to its own first vertex / uv / normal
    np.testing.assert_array_equal(b.vertex_indices, [[0, 1, 2], [0, 2, 1]])
    np.testing.assert_array_equal(b.normal_indices, [[0, 0, 0], [0, 0, 0]])
    np.testing.assert_array_equal(b.uvs[0], [[1, 1], [1, 1], [1, 1]])
    # single mesh: one mesh with all faces, offsets never advance
    one = ObjFile(write(tmp_path, "m.obj", text), single_mesh=True)
    assert len(one.meshes) == 1
    np.testing.assert_array_equal(one.meshes[0].vertex_indices[3:], [[4, 5, 6], [4, 6, 5]])


def test_obj_double_slash_normal_bug(tmp_path):
    """f v//vn scans the third normal into vn1 (obj.go:77-82)."""
    text = "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 1 0 0\nvn 0 1 0\nvn 0 0 1\nf 1//1 2//2 3//3\n"
    m = ObjFile(write(tmp_path, "m.obj", text)).meshes[0]
    np.testing.assert_array_equal(m.normal_indices, [[0, 2, -1]])


def test_obj_numbers(tmp_path):
    """%f tokens: exponents, hex mantissas, NaN/Inf, signs; extra fields ignored."""
    text = "v 1e3 -2.5 0x1p-2\nv nan +inf -Inf\nv 0.1 16777217 3.4028236e38 9\n"
    v = ObjFile(write(tmp_path, "m.obj", text)).meshes[0].vertices
    assert v[0].tolist() == [1000.0, -2.5, 0.25, 1.0]
    assert np.isnan(v[1, 0]) and v[1, 1] == np.inf and v[1, 2] == -np.inf
    assert v[2, 0] == np.float32(0.1) and v[2, 1] == np.float32(16777216.0)
    assert v[2, 2] == np.inf  # rounds past float32's maximum, as ParseFloat(.., 32) does


def test_obj_materials(tmp_path):
    """usemtl picks the material's texture; a material without map_Kd gets the
    solid magenta default; a shared image is decoded once; unknown names are nil."""
    img = np.random.default_rng(3).integers(0, 256, (4, 6, 4))
    (tmp_path / "tex.png").write_bytes(pngwrite.encode(img, 6, 8))
    write(tmp_path, "m.mtl", "newmtl plain\nKd 1 1 1\n\nnewmtl wood\nmap_Kd tex.png\n"
                             "newmtl wood2\nmap_Kd tex.png\n")
    text = ("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\nusemtl wood\nf 1 2 3\n"
            "usemtl plain\nf 1 2 3\nusemtl wood2\nf 1 2 3\nusemtl nosuch\nf 1 2 3\n")
    obj = ObjFile(write(tmp_path, "m.obj", text))
    assert len(obj.textures) == 2
    solid = [t for t in obj.textures if t.rgba is None][0]
    image = [t for t in obj.textures if t.rgba is not None][0]
    assert solid.type == _lib.GR_TEXTURE_SOLID and solid.color == (255, 0, 255, 255)
    assert image.type == _lib.GR_TEXTURE_IMAGE and (image.width, image.height) == (6, 4)
    np.testing.assert_array_equal(image.rgba, expected_rgba(img, 6, 8))
    ti, ts = obj.textures.index(image), obj.textures.index(solid)
    np.testing.assert_array_equal(obj.meshes[0].face_textures, [-1, ti, ts, ti, -1])


@pytest.mark.parametrize("text,status", [
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nf 1 2 3 4\n", "GR_ERR_INVALID_ARGUMENT"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nf 1/1 2/2 3/1\n", "GR_ERR_INDEX_OUT_OF_RANGE"),
    ("v 0 0\n", "GR_ERR_INVALID_ARGUMENT"),
    ("v 0 0 0\nf 1 x 2\n", "GR_ERR_INVALID_ARGUMENT"),
    ("# nothing\n", "GR_ERR_INVALID_ARGUMENT"),
    ("mtllib missing.mtl\nv 0 0 0\n", "GR_ERR_INVALID_ARGUMENT"),
])
def test_obj_errors(tmp_path, text, status):
    with pytest.raises(_lib.GorenderError) as e:
        ObjFile(write(tmp_path, "bad.obj", text))
    assert e.value.status == getattr(_lib, status)


@pytest.mark.skipif(not os.path.isdir(REF_MODELS), reason="reference model files absent")
def test_reference_models():
    cube = ObjFile(os.path.join(REF_MODELS, "cube.obj"))
    lines = open(os.path.join(REF_MODELS, "cube.obj")).read().split("\n")
    assert len(cube.meshes) == 1
    m = cube.meshes[0]
    assert len(m.vertices) == sum(l.startswith("v ") for l in lines)
    assert len(m.vertex_indices) == sum(l.startswith("f ") for l in lines)
    assert len(m.vertex_normals) == sum(l.startswith("vn ") for l in lines)
    assert len(cube.textures) == 1  # three materials share textures-16.png
    t = cube.textures[0]
    assert t.type == _lib.GR_TEXTURE_IMAGE_FAST and t.rgba.shape == (t.height, t.width, 4)
    np.testing.assert_array_equal(png_decode(os.path.join(REF_MODELS, "textures-16.png")), t.rgba)
    assert set(m.face_textures.tolist()) == {0}
    # first face "f 4/8/6 6/9/6 8/10/6"
    assert m.vertex_indices[0].tolist() == [3, 5, 7] and m.normal_indices[0].tolist() == [5, 5, 5]
    vts = [l for l in lines if l.startswith("vt ")]
    assert m.uvs[0, 0].tolist() == [np.float32(float(x)) for x in vts[7].split()[1:]]
    suz = ObjFile(os.path.join(REF_MODELS, "suzanne.obj"))
    slines = open(os.path.join(REF_MODELS, "suzanne.obj")).read().split("\n")
    assert len(suz.meshes) == 1
    assert len(suz.meshes[0].vertex_indices) == sum(l.startswith("f ") for l in slines)
    assert len(suz.meshes[0].vertices) == sum(l.startswith("v ") for l in slines)
