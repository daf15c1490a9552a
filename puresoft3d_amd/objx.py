"""OBJX meshes and scenes: the reference's mesh container (objcvt/objxio.cpp)
and the demo's scene loading (test2/loadscene.cpp:137-260), as Scenes for the
raster path.

File layout (objxio.cpp:9-45, little-endian, packed):
  header  u32 version (0x010003), u32 mesh count, u32 mesh-header bytes,
          u32 reserved, f32 camera position[4], f32 camera yaw/pitch/roll[4],
          4 lights x (f32 position[4], f32 direction[4])          176 bytes
  per mesh: header (mesh-header bytes; 1627 in version 0x010003)
          char name[260], u32 vertices, u32 indices, bool texcoords, bool
          normals, bool tangents, f32 ambient/diffuse/specular[4] (B,G,R,A),
          f32 specular exponent, char diffuse/bump/spc/spe file[260],
          char programme[260], u32 data bytes
        data: positions f32x4, normals f32x4 (if), tangents f32x4 (if),
          texcoords f32x2 (if), indices i32 (if any)
Colours come out as (R, G, B, A) clamped to [0, 1] (objxio.cpp:264-275).
"""
import struct

import os

import numpy as np

from . import native as N
from . import scenes as S

VERSION = 0x010003
NAME_LEN = 260
_HDR = struct.Struct("<4I4f4f32f")
_MESH = struct.Struct("<%ds2I3?13f%ds%ds%ds%ds%dsI" % ((NAME_LEN,) * 6))
HEADER_BYTES = _HDR.size          # 176
MESH_HEADER_BYTES = _MESH.size    # 1627


class ObjxMesh:
    def __init__(self, name, vertices, normals=None, tangents=None, texcoords=None, indices=None,
                 ambient=(0, 0, 0, 0), diffuse=(0, 0, 0, 0), specular=(0, 0, 0, 0), specular_exponent=0.0,
                 diffuse_file="", bump_file="", spc_file="", spe_file="", programme=""):
        self.name = name
        self.vertices = np.ascontiguousarray(vertices, np.float32).reshape(-1, 4)
        n = self.vertices.shape[0]
        def opt(a, k):
            return None if a is None else np.ascontiguousarray(a, np.float32).reshape(n, k)
        self.normals, self.tangents, self.texcoords = opt(normals, 4), opt(tangents, 4), opt(texcoords, 2)
        self.indices = None if indices is None or len(indices) == 0 else np.ascontiguousarray(indices, np.int32)
        self.ambient = np.asarray(ambient, np.float32)
        self.diffuse = np.asarray(diffuse, np.float32)
        self.specular = np.asarray(specular, np.float32)
        self.specular_exponent = float(specular_exponent)
        self.diffuse_file, self.bump_file, self.spc_file, self.spe_file = diffuse_file, bump_file, spc_file, spe_file
        self.programme = programme


class ObjxScene:
    def __init__(self, meshes, camera_pos=(0, 0, 0, 0), camera_ypr=(0, 0, 0, 0), lights=None):
        self.meshes = list(meshes)
        self.camera_pos = np.asarray(camera_pos, np.float32)
        self.camera_ypr = np.asarray(camera_ypr, np.float32)
        # (position[4], direction[4]); a light without direction is omni (objxio.cpp:219-225)
        self.lights = lights or [(np.zeros(4, np.float32), np.zeros(4, np.float32))] * 4

    def light_types(self):
        return ["omni" if float(np.sqrt(np.dot(d[:3], d[:3]))) < 1e-7 else "dir" for _, d in self.lights]


def _cstr(b):
    return b.split(b"\0", 1)[0].decode("latin-1")


def _bgra_to_rgba(c):
    return np.clip(np.array([c[2], c[1], c[0], c[3]], np.float32), 0.0, 1.0)


def read_objx(data):
    """Parse OBJX bytes (or a path).  Raises ValueError for another version or
    a mesh header shorter than this version's."""
    if isinstance(data, (str, bytes)) and not isinstance(data, bytes):
        with open(data, "rb") as f:
            data = f.read()
    if len(data) < HEADER_BYTES:
        raise ValueError("not an OBJX file: %d bytes" % len(data))
    h = _HDR.unpack_from(data, 0)
    version, n_meshes, mh_bytes = h[0], h[1], h[2]
    if version != VERSION or mh_bytes < MESH_HEADER_BYTES:
        raise ValueError("unsupported OBJX version 0x%x / mesh header %d bytes" % (version, mh_bytes))
    cam_pos, cam_ypr = h[4:8], h[8:12]
    lf = h[12:44]
    lights = [(np.array(lf[8 * i:8 * i + 4], np.float32), np.array(lf[8 * i + 4:8 * i + 8], np.float32))
              for i in range(4)]
    off = HEADER_BYTES
    meshes = []
    for _ in range(n_meshes):
        if off + mh_bytes > len(data):
            raise ValueError("truncated OBJX mesh header at byte %d" % off)
        m = _MESH.unpack_from(data, off)
        name, nv, ni, has_tc, has_n, has_t = _cstr(m[0]), m[1], m[2], m[3], m[4], m[5]
        amb, dif, spc, sexp = m[6:10], m[10:14], m[14:18], m[18]
        files = [_cstr(x) for x in m[19:24]]
        nbytes = m[24]
        p = off + mh_bytes
        need = nv * 16 * (1 + int(has_n) + int(has_t)) + nv * 8 * int(has_tc) + ni * 4
        if nbytes != need or p + need > len(data):
            raise ValueError("mesh %r: %d data bytes recorded, %d implied" % (name, nbytes, need))
        def take(count, dtype, k):
            nonlocal p
            a = np.frombuffer(data, dtype, count * k, p).reshape(count, k).copy()
            p += count * k * 4
            return a
        verts = take(nv, "<f4", 4)
        norms = take(nv, "<f4", 4) if has_n else None
        tans = take(nv, "<f4", 4) if has_t else None
        tcs = take(nv, "<f4", 2) if has_tc else None
        idx = take(ni, "<i4", 1).ravel() if ni > 0 else None
        meshes.append(ObjxMesh(name, verts, norms, tans, tcs, idx, _bgra_to_rgba(amb), _bgra_to_rgba(dif),
                               _bgra_to_rgba(spc), sexp, *files))
        off = off + mh_bytes + nbytes
    return ObjxScene(meshes, cam_pos, cam_ypr, lights)


def write_objx(scene):
    """OBJX bytes of a scene (objxio.cpp:90-170 layout); colours are stored B,G,R,A."""
    lf = []
    for pos, d in scene.lights:
        lf += list(np.asarray(pos, np.float32)) + list(np.asarray(d, np.float32))
    out = [_HDR.pack(VERSION, len(scene.meshes), MESH_HEADER_BYTES, 0, *scene.camera_pos, *scene.camera_ypr, *lf)]
    for m in scene.meshes:
        nv = m.vertices.shape[0]
        ni = 0 if m.indices is None else m.indices.size
        body = [m.vertices.astype("<f4").tobytes()]
        for a in (m.normals, m.tangents, m.texcoords):
            if a is not None:
                body.append(a.astype("<f4").tobytes())
        if ni:
            body.append(m.indices.astype("<i4").tobytes())
        body = b"".join(body)
        col = lambda c: (c[2], c[1], c[0], c[3])  # noqa: E731
        enc = lambda s: s.encode("latin-1")[:NAME_LEN - 1]  # noqa: E731
        out.append(_MESH.pack(enc(m.name), nv, ni, m.texcoords is not None, m.normals is not None,
                              m.tangents is not None, *col(m.ambient), *col(m.diffuse), *col(m.specular),
                              m.specular_exponent, enc(m.diffuse_file), enc(m.bump_file), enc(m.spc_file),
                              enc(m.spe_file), enc(m.programme), len(body)))
        out.append(body)
    return b"".join(out)


# ---------------------------------------------------------------- vector ops in FP32, op for op
def _f32(x):
    return np.asarray(x, np.float32)


def _norm3(v):
    """mcemaths_norm_3_4 (vector.cpp): v * (1 / sqrt(dot3(v, v))), FP32"""
    v = _f32(v)
    d = (v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]) + v[..., 2] * v[..., 2]
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.float32(1.0) / np.sqrt(d.astype(np.float32))
    out = v.copy()
    out[..., :3] = v[..., :3] * r[..., None]
    return out


def _cross3(a, b):
    """mcemaths_cross_3 (vector.cpp:114-144): r = a x b in xyz, w = 0"""
    a, b = _f32(a), _f32(b)
    r = np.zeros_like(a)
    r[..., 0] = a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1]
    r[..., 1] = a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2]
    r[..., 2] = a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]
    return r


def triangle_tangents(vertices, texcoords):
    """Per-triangle tangents of a non-indexed mesh, normalised and copied to
    its three vertices (objxio.cpp generate_tangents, the non-indexed branch).
    The formula is the one generate_tangent intends, t = ((c-a)*dv2 - (b-a)*dv1)
    / (du1*dv2 - dv1*du2) with du1,dv1 = tc-ta and du2,dv2 = tb-ta; the
    reference itself reads an unset vector there (objxio.cpp generate_tangent
    writes b-a over c-a), so meshes stored without tangents do not match it.
    None of the reference's own assets lacks tangents."""
    v = _f32(vertices).reshape(-1, 3, 4)
    t = _f32(texcoords).reshape(-1, 3, 2)
    c1 = t[:, 2] - t[:, 0]
    c2 = t[:, 1] - t[:, 0]
    e1 = v[:, 2] - v[:, 0]
    e2 = v[:, 1] - v[:, 0]
    tan = np.zeros_like(v[:, 0])
    tan[:, :3] = e1[:, :3] * c2[:, 1:2] - e2[:, :3] * c1[:, 1:2]
    with np.errstate(divide="ignore", invalid="ignore"):
        f = np.float32(1.0) / (c1[:, 0] * c2[:, 1] - c1[:, 1] * c2[:, 0])
    tan[:, :3] *= f[:, None]
    tan = _norm3(tan)
    return np.repeat(tan, 3, axis=0)


def model_space(vertices):
    """The demo's world-to-model split of a mesh (loadscene.cpp:212-237): the
    translation is the centre of the bounding box (minps/maxps from
    (FLT_MAX,..,0) and (-FLT_MAX,..,0), centre = max + (min - max) * 0.5,
    vector.cpp:320-350,410-420); positions lose it in xyz and get w = 1."""
    v = _f32(vertices)
    lo = np.array([np.finfo(np.float32).max] * 3 + [0.0], np.float32)
    hi = np.array([-np.finfo(np.float32).max] * 3 + [0.0], np.float32)
    lo = np.minimum(lo, v.min(axis=0))
    hi = np.maximum(hi, v.max(axis=0))
    centre = hi + (lo - hi) * np.float32(0.5)
    out = v.copy()
    out[:, :3] = v[:, :3] - centre[:3]
    out[:, 3] = 1.0
    return out, centre


def split_name(name):
    """'object/mesh' -> (object, mesh); names without '/' are skipped by the loader"""
    i = name.find("/")
    return (None, None) if i < 0 else (name[:i], name[i + 1:])


def load_picture(path):
    """PuresoftDefaultPictureLoader (picldr.cpp:33-133) with PIL in place of
    GDI+: the picture as 32-bit ARGB texels (bytes B, G, R, A; alpha 255 for
    pictures without one), rows flipped top to bottom (RotateNoneFlipY,
    picldr.cpp:47), so memory row 0 holds the picture's bottom row; the
    scanline is width * 4 (WIDTHBYTES of 32 bits).  Returns (h, w, 4) uint8."""
    try:
        from PIL import Image
    except ImportError as e:   # the product decodes no pictures itself
        raise ImportError("loading pictures needs PIL (Pillow)") from e
    with Image.open(path) as im:
        rgba = np.asarray(im.convert("RGBA"))
    return np.ascontiguousarray(rgba[::-1, :, [2, 1, 0, 3]])


def load_scene(objx, width=640, height=480, fov_deg=60.0, tex_size=256, seed=5, prog_textured=None,
               name=None, picture_dir=None, picture_wrap=N.CLAMP):
    """A Scene of an OBJX file, meshes as the demo loader prepares them:
      * meshes named 'object/mesh' only (loadscene.cpp:151-158), indices expanded;
      * model-space positions (w = 1) and the bounding-box translation, drawn
        with M = translate(translation) so that world positions are restored;
      * binormal = norm(cross(normal, tangent)) (loadscene.cpp:176-184);
      * slots 0 position, 1 tangent, 2 binormal, 3 normal, 4 texcoord (or the
        material colour x 255 in slot 2 for colour-only meshes).
    Textured meshes (a diffuse file) take their pictures from picture_dir
    when it is given, as the demo loader does (loadscene.cpp:316-330): the
    diffuse picture, and the bump picture when the mesh names one, each
    loaded once (load_picture) into a texture of picture_wrap (CLAMP there,
    loadscene.cpp:362; WRAP in the other demo, test/scenobj.cpp:177); a mesh with a bump
    picture is drawn by tex1bump1light1 (DEF03), one without by tex1light1
    (DEF01).  Without picture_dir, textured meshes use DEF03 (or
    prog_textured) with seeded albedo and normal-map textures.  The others
    use colr1light1 (DEF02) with their diffuse colour.  The camera looks from
    the file's camera position at the scene's centre."""
    sc_in = objx if isinstance(objx, ObjxScene) else read_objx(objx)
    prog_textured = N.PROC_DEF03 if prog_textured is None else prog_textured
    sc = S.Scene(width, height)
    sc.name = name or "objx_%dmesh_%dx%d" % (len(sc_in.meshes), width, height)
    centres = []
    items = []
    for m in sc_in.meshes:
        obj, part = split_name(m.name)
        if obj is None:
            continue
        verts, norms, tans, tcs = m.vertices, m.normals, m.tangents, m.texcoords
        if m.indices is not None:
            verts = verts[m.indices]
            norms = None if norms is None else norms[m.indices]
            tans = None if tans is None else tans[m.indices]
            tcs = None if tcs is None else tcs[m.indices]
        n3 = verts.shape[0] // 3 * 3
        verts = verts[:n3]
        if tans is None and tcs is not None:
            tans = triangle_tangents(verts, tcs[:n3])
        mv, centre = model_space(verts)
        centres.append(centre[:3])
        items.append((obj, part, m, mv, centre, None if norms is None else norms[:n3],
                      None if tans is None else tans[:n3], None if tcs is None else tcs[:n3]))
    if not items:
        raise ValueError("no 'object/mesh' meshes in the OBJX scene")
    ctr = np.mean(np.array(centres, np.float64), axis=0)
    eye = np.asarray(sc_in.camera_pos[:3], np.float64)
    if not np.all(np.isfinite(eye)) or np.linalg.norm(eye - ctr) < 1e-6:
        eye = ctr + np.array([0.0, 0.0, 3.0])
    PV = S.perspective(np.deg2rad(fov_deg), width / height, 0.1 * np.linalg.norm(eye - ctr) / 3.0,
                       10.0 * np.linalg.norm(eye - ctr) + 10.0) @ S.look_at(eye, ctr)
    sc.uniforms[3] = S.colmajor(PV)
    sc.uniforms[5] = S.colmajor(np.eye(4))
    lp = sc_in.lights[0][0]
    sc.uniforms[7] = S.vec4(*(lp[:3] if np.any(lp[:3] != 0) else eye + np.array([1.0, 2.0, 0.5])), 0.0)
    sc.uniforms[8] = S.vec4(*eye, 0.0)
    sc.textures.append(S.TextureSpec([S.noise_texture(tex_size, seed, cells=16)], wrap=N.WRAP))
    sc.textures.append(S.TextureSpec([S.normal_map(tex_size, seed + 1, cells=16)], wrap=N.WRAP))
    sc.uniforms[9] = np.array([0], np.int32)
    sc.uniforms[10] = np.array([1], np.int32)
    sc.uniforms[11] = np.array([16], np.int32)
    pictures = {}   # file name -> texture index (findOrCreateTexture, loadscene.cpp:349-366)

    def picture(fname):
        if fname not in pictures:
            pictures[fname] = len(sc.textures)
            sc.textures.append(S.TextureSpec([load_picture(os.path.join(picture_dir, fname))], wrap=picture_wrap))
        return pictures[fname]

    for obj, part, m, mv, centre, norms, tans, tcs in items:
        key = "%s/%s" % (obj, part)
        n = mv.shape[0]
        normals = norms if norms is not None else np.tile(np.array([0, 0, 1, 0], np.float32), (n, 1))
        textured = bool(m.diffuse_file) and tcs is not None and tans is not None
        uni = {4: S.colmajor(S.translate(*[float(x) for x in centre[:3]]))}
        if textured:
            binorm = _norm3(_cross3(normals, tans))
            sc.meshes[key] = {0: mv, 1: _f32(tans), 2: binorm, 3: _f32(normals), 4: _f32(tcs)}
            prog = prog_textured
            if picture_dir is not None:
                uni[9] = np.array([picture(m.diffuse_file)], np.int32)
                if m.bump_file:
                    uni[10] = np.array([picture(m.bump_file)], np.int32)
                    prog = N.PROC_DEF03
                else:
                    prog = N.PROC_DEF01
        else:
            col = np.tile(m.diffuse * np.float32(255.0), (n, 1)).astype(np.float32)
            sc.meshes[key] = {0: mv, 1: _f32(normals), 2: col}
            prog = N.PROC_DEF02
        sc.draws.append(S.DrawSpec(key, prog, uniforms=uni))
    return sc
