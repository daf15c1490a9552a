"""OBJX container (objcvt/objxio.cpp) and the demo's scene preparation
(test2/loadscene.cpp:137-260), on the reference's own assets (src/test/sphere.objx,
src/test2/plane.objx, gzipped in tests/golden) and on closed forms.  CPU only."""
import gzip
import os
import struct

import numpy as np
import pytest

from puresoft3d_amd import native as N
from puresoft3d_amd import objx as X
import scenegen

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def asset(name):
    with gzip.open(os.path.join(GOLDEN, "%s.objx.gz" % name), "rb") as f:
        return f.read()


def test_layout_sizes():
    # header 4 u32 + 8 f32 + 4 lights x 8 f32; mesh header as packed in objxio.cpp:25-44
    assert X.HEADER_BYTES == 176
    assert X.MESH_HEADER_BYTES == 260 + 8 + 3 + 13 * 4 + 5 * 260 + 4 == 1627


def test_reference_assets_parse():
    s = X.read_objx(asset("sphere"))
    assert len(s.meshes) == 1
    m = s.meshes[0]
    assert m.name == "sphere" and m.vertices.shape == (1584, 4) and m.indices is None
    assert m.normals.shape == (1584, 4) and m.tangents.shape == (1584, 4) and m.texcoords.shape == (1584, 2)
    np.testing.assert_allclose(m.diffuse, [0.8, 0.8, 0.8, 1.0], rtol=1e-6)
    # unit sphere: positions on the sphere, normals unit length
    r = np.linalg.norm(m.vertices[:, :3], axis=1)
    assert abs(r.mean() - r[0]) < 1e-3 * r[0]
    np.testing.assert_allclose(np.linalg.norm(m.normals[:, :3], axis=1), 1.0, atol=1e-3)
    p = X.read_objx(asset("plane"))
    names = [q.name for q in p.meshes]
    assert len(names) == 13 and names[0] == "book/cover" and names[-1] == "plane/top@noshadow"
    assert all(q.vertices.shape[0] % 3 == 0 for q in p.meshes)
    jar = p.meshes[names.index("jar/body")]
    assert jar.vertices.shape[0] == 21402 and jar.diffuse_file == "marmite.jpg"
    assert jar.programme == "VP_DiffuseOnly:IP_DiffuseOnly:FP_DiffuseOnly"
    np.testing.assert_allclose(p.camera_pos[:3], [-0.83, 0.47, 0.44], atol=0.01)
    assert p.light_types() == ["omni"] * 4


@pytest.mark.parametrize("name", ["sphere", "plane"])
def test_round_trip_is_byte_identical(name):
    data = asset(name)
    assert X.write_objx(X.read_objx(data)) == data


def test_colour_order_and_clamp():
    m = X.ObjxMesh("o/m", np.zeros((3, 4)), diffuse=(0.1, 0.2, 0.3, 0.4), ambient=(2.0, -1.0, 0.5, 1.0))
    data = X.write_objx(X.ObjxScene([m]))
    # stored B, G, R, A right after the three flags
    off = X.HEADER_BYTES + 260 + 8 + 3
    amb = struct.unpack_from("<4f", data, off)
    dif = struct.unpack_from("<4f", data, off + 16)
    np.testing.assert_allclose(dif, [0.3, 0.2, 0.1, 0.4], rtol=1e-6)
    np.testing.assert_allclose(amb, [0.5, -1.0, 2.0, 1.0], rtol=1e-6)
    back = X.read_objx(data).meshes[0]
    np.testing.assert_allclose(back.diffuse, [0.1, 0.2, 0.3, 0.4], rtol=1e-6)
    np.testing.assert_allclose(back.ambient, [1.0, 0.0, 0.5, 1.0], rtol=1e-6)   # clamped to [0, 1]


def test_rejects_other_versions_and_truncation():
    data = bytearray(asset("sphere"))
    bad = bytearray(data)
    struct.pack_into("<I", bad, 0, 0x010002)
    with pytest.raises(ValueError):
        X.read_objx(bytes(bad))
    with pytest.raises(ValueError):
        X.read_objx(bytes(data[:-100]))
    with pytest.raises(ValueError):
        X.read_objx(bytes(data[:100]))


def test_model_space_split():
    rng = np.random.default_rng(0)
    v = np.concatenate([rng.uniform(-3, 5, (30, 3)), np.zeros((30, 1))], 1).astype(np.float32)
    mv, c = X.model_space(v)
    assert np.all(mv[:, 3] == 1.0)
    lo, hi = v[:, :3].min(0), v[:, :3].max(0)
    np.testing.assert_allclose(c[:3], (lo + hi) / 2, rtol=1e-6)
    # the box is centred on the origin, and translation restores the positions
    np.testing.assert_allclose(mv[:, :3].min(0), -mv[:, :3].max(0), atol=1e-5)
    np.testing.assert_allclose(mv[:, :3] + c[:3], v[:, :3], atol=1e-5)


def test_triangle_tangents_closed_form():
    # texture u along +x, v along +y: the tangent is +x for any such triangle
    v = np.array([[0, 0, 0, 1], [2, 0, 0, 1], [0, 3, 0, 1]], np.float32)
    tc = np.array([[0, 0], [1, 0], [0, 1]], np.float32)
    t = X.triangle_tangents(v, tc)
    np.testing.assert_allclose(t[:, :3], [[1, 0, 0]] * 3, atol=1e-7)
    # mirrored u: the tangent flips
    tc2 = np.array([[1, 0], [0, 0], [1, 1]], np.float32)
    np.testing.assert_allclose(X.triangle_tangents(v, tc2)[:, :3], [[-1, 0, 0]] * 3, atol=1e-7)


def test_loader_skips_unprefixed_and_expands_indices():
    tri = np.array([[0, 0, 0, 1], [1, 0, 0, 1], [0, 1, 0, 1], [1, 1, 0, 1]], np.float32)
    n = np.tile(np.array([0, 0, 1, 0], np.float32), (4, 1))
    a = X.ObjxMesh("plain", tri[:3], n[:3])
    b = X.ObjxMesh("obj/quad", tri, n, indices=[0, 1, 2, 2, 1, 3], diffuse=(1, 0.5, 0.25, 1))
    sc = X.load_scene(X.ObjxScene([a, b], camera_pos=(0.5, 0.5, 3, 0)), 64, 48)
    assert list(sc.meshes) == ["obj/quad"]
    m = sc.meshes["obj/quad"]
    assert m[0].shape == (6, 4) and np.all(m[0][:, 3] == 1)
    np.testing.assert_array_equal(m[2][0], np.float32([255, 127.5, 63.75, 255]))
    assert sc.draws[0].prog == N.PROC_DEF02


def test_binormals_of_the_jar_are_unit_and_orthogonal():
    p = X.read_objx(asset("plane"))
    sc = X.load_scene(p, 64, 48)
    m = sc.meshes["jar/body"]
    b, nrm, t = m[2][:, :3], m[3][:, :3], m[1][:, :3]
    # about a third of the jar's stored tangents are NaN (as exported); their
    # binormals are NaN, the others unit length
    live = ~np.isnan(t).any(axis=1)
    assert 0.5 < live.mean() < 1.0 and np.isnan(b[~live]).all() and not np.isnan(b[live]).any()
    np.testing.assert_allclose(np.linalg.norm(b[live], axis=1), 1.0, atol=1e-5)
    # b = norm(n x t) is perpendicular to the normal and the tangent (which the
    # exporter did not make perpendicular to each other)
    b, nrm, t = b[live], nrm[live], t[live]
    tn = t / np.linalg.norm(t, axis=1, keepdims=True)
    nn = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    assert np.abs(np.einsum("ij,ij->i", b, nn)).max() < 1e-5
    assert np.abs(np.einsum("ij,ij->i", b, tn)).max() < 1e-5


def test_load_picture_rows_and_channels(tmp_path):
    """picldr.cpp:33-133: 32-bit ARGB texels (bytes B, G, R, A) with the
    picture's rows flipped (RotateNoneFlipY), alpha 255 where the picture has
    none, palettes expanded"""
    from PIL import Image
    rgba = np.array([[[10, 20, 30, 40], [50, 60, 70, 80], [90, 100, 110, 120]],
                     [[130, 140, 150, 160], [170, 180, 190, 200], [210, 220, 230, 240]]], np.uint8)
    Image.fromarray(rgba, "RGBA").save(str(tmp_path / "a.png"))
    got = X.load_picture(str(tmp_path / "a.png"))
    assert got.shape == (2, 3, 4) and got.dtype == np.uint8 and got.flags.c_contiguous
    np.testing.assert_array_equal(got[0, 0], [150, 140, 130, 160])   # bottom-left pixel, B G R A
    np.testing.assert_array_equal(got[1, 0], [30, 20, 10, 40])        # top-left
    np.testing.assert_array_equal(got[1, 2], [110, 100, 90, 120])     # top-right
    Image.fromarray(rgba[..., :3], "RGB").save(str(tmp_path / "b.png"))
    assert (X.load_picture(str(tmp_path / "b.png"))[..., 3] == 255).all()
    pal = Image.fromarray(rgba[..., :3], "RGB").convert("P")
    pal.save(str(tmp_path / "c.png"))
    want = np.asarray(pal.convert("RGB"))[::-1, :, ::-1]
    np.testing.assert_array_equal(X.load_picture(str(tmp_path / "c.png"))[..., :3], want)


def test_reference_pictures_decode():
    """the reference's demo pictures kept as fixtures decode to the sizes the
    demos rely on (decoded with PIL; agreement with GDI+ is not pinned)"""
    import os
    pdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pictures")
    sizes = {"marmite.jpg": (1000, 1500), "penhold.jpg": (460, 818), "top.jpg": (1356, 2048),
             "earth.jpg": (1024, 2048), "earth.dot3.png": (1024, 2048), "purplenebula_up.png": (512, 512)}
    for name, (h, w) in sizes.items():
        img = X.load_picture(os.path.join(pdir, name))
        assert img.shape == (h, w, 4), name
        assert img[..., :3].std() > 5, name


def test_objx_scenes_take_their_pictures(oracle):
    """plane.objx's textured meshes take marmite.jpg, top.jpg and penhold.jpg
    (diffuse only: tex1light1), each loaded once; the sphere takes the demo
    Earth's diffuse and bump maps (tex1bump1light1); both render"""
    sc = scenegen.objx_scene("plane", 320, 240, pictures=True)
    assert len(sc.textures) == 2 + 3
    progs = {d.prog for d in sc.draws}
    assert progs == {N.PROC_DEF01, N.PROC_DEF02}
    f, st = oracle.render_scene(sc)
    assert st.frags_shaded > 10000
    sp = scenegen.objx_scene("sphere", 320, 240, pictures=True)
    assert [d.prog for d in sp.draws] == [N.PROC_DEF03]
    f2, st2 = oracle.render_scene(sp)
    assert st2.frags_shaded > 3000
