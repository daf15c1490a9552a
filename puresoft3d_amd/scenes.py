"""Seeded procedural scenes for the benchmark workloads and the parity tests.

A Scene holds non-indexed triangle streams in the reference's VBO slot
convention (test/scenobj.cpp:135-157: 0 position float4, 1 tangent, 2 binormal,
3 normal float4, 4 texcoord float2), level-0 texture images (BGRA8 bytes),
uniform slots and an ordered list of draws.  Matrices are column-major float32,
as mcemaths stores them (matrix.cpp:515-557).
"""
import numpy as np

from . import native as N


class TextureSpec:
    def __init__(self, images, wrap=N.WRAP, levels=1):
        """images: list (one per layer) of uint8 arrays (h, w, 4) in BGRA order."""
        self.images = [np.ascontiguousarray(im, np.uint8) for im in images]
        self.height, self.width = self.images[0].shape[:2]
        self.wrap = wrap
        self.levels = levels
        self.layers = len(self.images)


class DrawSpec:
    def __init__(self, mesh, prog, uniforms=None, behavior=None):
        self.mesh = mesh
        self.prog = prog
        self.uniforms = uniforms or {}
        self.behavior = behavior


class ShadowPass:
    """A depth-only pass (DEF05, shadow.cpp) rendered before the scene's own
    draws into a size x size FP32 depth texture, which becomes the scene's
    texture number len(scene.textures) (clamped).  light_pv: the light's PV
    (column-major); draws: DrawSpecs of DEF05 with their model matrix in
    uniform 4."""

    def __init__(self, size, light_pv, draws):
        self.size, self.light_pv, self.draws = size, light_pv, draws


class Scene:
    def __init__(self, width, height):
        self.width, self.height = width, height
        self.meshes = {}
        self.textures = []
        self.uniforms = {}
        self.draws = []
        self.shadow = None
        self.clear_depth = 1.0
        self.clear_colour = 0
        self.name = ""

    def triangles(self):
        return sum(min(a.shape[0] for a in self.meshes[d.mesh].values()) // 3 for d in self.draws)

    def vertex_bytes(self):
        tot = 0
        for d in self.draws:
            m = self.meshes[d.mesh]
            n = min(a.shape[0] for a in m.values()) // 3 * 3
            tot += sum(n * a.shape[1] * 4 for a in m.values())
        return tot


# ---------------------------------------------------------------- matrices
def perspective(fov_rad, aspect, near, far):
    """mcemaths_make_proj_perspective (matrxgl.cpp:9-22), as a 4x4 row-major math matrix."""
    m = np.eye(4)
    h = 1.0 / np.tan(fov_rad / 2.0)
    m[0, 0] = h / aspect
    m[1, 1] = h
    m[2, 2] = (far + near) / (near - far)
    m[3, 2] = -1.0
    m[2, 3] = 2.0 * (near * far) / (near - far)
    m[3, 3] = 0.0
    return m


def look_at(eye, target, up=(0.0, 1.0, 0.0)):
    eye, target, up = (np.asarray(v, np.float64) for v in (eye, target, up))
    f = target - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[0, 3], m[1, 3], m[2, 3] = -s @ eye, -u @ eye, f @ eye
    return m


def translate(x, y, z):
    m = np.eye(4)
    m[:3, 3] = (x, y, z)
    return m


def colmajor(m):
    """row-major 4x4 math matrix -> 16 float32 in mcemaths column-major order"""
    return np.ascontiguousarray(np.asarray(m, np.float64).T.ravel(), np.float32)


def vec4(x, y, z, w=0.0):
    return np.array([x, y, z, w], np.float32)


# ---------------------------------------------------------------- noise
def value_noise(h, w, cells, rng, channels=1, octaves=1, persistence=0.5):
    """Seeded value noise in [0,1]: random lattice values, smoothstep interpolation,
    optional fBm octaves; tiles periodically."""
    out = np.zeros((h, w, channels), np.float64)
    amp, tot = 1.0, 0.0
    for o in range(octaves):
        c = cells * (2 ** o)
        lat = rng.random((c + 1, c + 1, channels))
        lat[c, :] = lat[0, :]
        lat[:, c] = lat[:, 0]
        fy = (np.arange(h) + 0.5) / h * c
        fx = (np.arange(w) + 0.5) / w * c
        iy, ix = np.floor(fy).astype(int), np.floor(fx).astype(int)
        ty, tx = fy - iy, fx - ix
        ty, tx = ty * ty * (3 - 2 * ty), tx * tx * (3 - 2 * tx)
        a = lat[iy][:, ix]
        b = lat[iy][:, ix + 1]
        cc = lat[iy + 1][:, ix]
        d = lat[iy + 1][:, ix + 1]
        tx3 = tx[None, :, None]
        ty3 = ty[:, None, None]
        top = a + (b - a) * tx3
        bot = cc + (d - cc) * tx3
        out += amp * (top + (bot - top) * ty3)
        tot += amp
        amp *= persistence
    return out / tot


def noise_texture(size, seed, cells=16, octaves=1):
    rng = np.random.default_rng(seed)
    v = value_noise(size, size, cells, rng, channels=3, octaves=octaves)
    img = np.empty((size, size, 4), np.uint8)
    img[..., :3] = np.clip(v * 255.0 + 0.5, 0, 255).astype(np.uint8)
    img[..., 3] = 255
    return img


def normal_map(size, seed, cells=32, strength=4.0):
    """tangent-space normal map of a seeded value-noise height field; r,g,b = x,y,z"""
    rng = np.random.default_rng(seed)
    hgt = value_noise(size, size, cells, rng, channels=1, octaves=3)[..., 0] * strength
    dx = (np.roll(hgt, -1, 1) - np.roll(hgt, 1, 1)) * 0.5 * (cells / 8.0)
    dy = (np.roll(hgt, -1, 0) - np.roll(hgt, 1, 0)) * 0.5 * (cells / 8.0)
    n = np.stack([-dx, -dy, np.ones_like(dx)], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    img = np.empty((size, size, 4), np.uint8)
    enc = np.clip((n * 0.5 + 0.5) * 255.0 + 0.5, 0, 255).astype(np.uint8)
    img[..., 2], img[..., 1], img[..., 0] = enc[..., 0], enc[..., 1], enc[..., 2]
    img[..., 3] = 255
    return img


def mip_levels(w, h):
    n = 1
    while (w >> n) > 0 or (h >> n) > 0:
        n += 1
    return n


# ---------------------------------------------------------------- meshes
def _orient_ccw(tris_pos, outward):
    """flip triangles whose winding is not counter-clockwise seen from `outward`"""
    a, b, c = tris_pos[:, 0, :3], tris_pos[:, 1, :3], tris_pos[:, 2, :3]
    n = np.cross(b - a, c - a)
    return (n * outward).sum(-1) < 0


def uv_sphere(stacks, slices, radius=1.0, center=(0, 0, 0), seed=None, tangents=False):
    """Lat-long sphere, two triangles per quad including the poles."""
    i = np.arange(stacks + 1)[:, None]
    j = np.arange(slices + 1)[None, :]
    th = np.pi * i / stacks
    ph = 2 * np.pi * j / slices
    nx = np.sin(th) * np.cos(ph)
    ny = np.cos(th) * np.ones_like(ph)
    nz = np.sin(th) * np.sin(ph)
    nrm = np.stack(np.broadcast_arrays(nx, ny, nz), -1)
    uv = np.stack(np.broadcast_arrays(j / slices + 0 * i, i / stacks + 0 * j), -1)
    tan = np.stack(np.broadcast_arrays(-np.sin(ph) + 0 * th, 0 * th + 0 * ph, np.cos(ph) + 0 * th), -1)
    q = []
    for a in range(stacks):
        for b in range(slices):
            q.append(((a, b), (a + 1, b), (a + 1, b + 1)))
            q.append(((a, b), (a + 1, b + 1), (a, b + 1)))
    idx = np.array(q)  # (T, 3, 2)
    n = nrm[idx[..., 0], idx[..., 1]]
    t_uv = uv[idx[..., 0], idx[..., 1]]
    t_tan = tan[idx[..., 0], idx[..., 1]]
    pos = n * radius + np.asarray(center, np.float64)
    flip = _orient_ccw(pos, (pos - np.asarray(center, np.float64)).mean(1))
    for arr in (n, t_uv, t_tan, pos):
        arr[flip] = arr[flip][:, [0, 2, 1]]
    if seed is not None:
        perm = np.random.default_rng(seed).permutation(len(idx))
        n, t_uv, t_tan, pos = n[perm], t_uv[perm], t_tan[perm], pos[perm]
    nv = n.shape[0] * 3
    P = np.ones((nv, 4), np.float32)
    P[:, :3] = pos.reshape(-1, 3)
    Nn = np.zeros((nv, 4), np.float32)
    Nn[:, :3] = n.reshape(-1, 3)
    UV = t_uv.reshape(-1, 2).astype(np.float32)
    mesh = {0: P, 3: Nn, 4: UV}
    if tangents:
        T = np.zeros((nv, 4), np.float32)
        T[:, :3] = t_tan.reshape(-1, 3)
        B = np.zeros((nv, 4), np.float32)
        B[:, :3] = np.cross(Nn[:, :3], T[:, :3])
        mesh[1], mesh[2] = T, B
    return mesh


def displaced_sphere(stacks=125, slices=200, amp=0.08, seed=41):
    """Lat-long sphere (2*stacks*slices triangles) displaced along the radius
    by seeded periodic value noise; normals, tangents and binormals from the
    displaced grid.  Slots 0-4 as the Blinn-Phong programme reads them."""
    rng = np.random.default_rng(seed)
    hgt = value_noise(stacks + 1, slices + 1, 8, rng, channels=1, octaves=3)[..., 0]
    i = np.arange(stacks + 1)[:, None]
    j = np.arange(slices + 1)[None, :]
    th = np.pi * i / stacks
    ph = 2 * np.pi * j / slices
    rad = np.stack(np.broadcast_arrays(np.sin(th) * np.cos(ph), np.cos(th) + 0 * ph, np.sin(th) * np.sin(ph)), -1)
    pos = rad * (1.0 + amp * (hgt[..., None] - 0.5))
    dth = np.gradient(pos, axis=0)
    dph = np.gradient(pos, axis=1)
    nrm = np.cross(dph, dth)
    ln = np.linalg.norm(nrm, axis=-1, keepdims=True)
    pole = ln[..., 0] < 1e-9
    nrm = np.where(pole[..., None], rad, nrm / np.maximum(ln, 1e-30))
    nrm *= np.sign((nrm * rad).sum(-1, keepdims=True) + 1e-30)
    tan = np.stack(np.broadcast_arrays(-np.sin(ph) + 0 * th, 0 * th + 0 * ph, np.cos(ph) + 0 * th), -1)
    uv = np.stack(np.broadcast_arrays(j / slices + 0 * i, i / stacks + 0 * j), -1)
    q = []
    for a in range(stacks):
        for b in range(slices):
            q.append(((a, b), (a + 1, b), (a + 1, b + 1)))
            q.append(((a, b), (a + 1, b + 1), (a, b + 1)))
    idx = np.array(q)
    P = pos[idx[..., 0], idx[..., 1]]
    flip = _orient_ccw(P, P.mean(1))
    idx[flip] = idx[flip][:, [0, 2, 1]]
    nv = idx.shape[0] * 3
    out = {}
    for slot, arr, w in ((0, pos, 1.0), (1, tan, 0.0), (3, nrm, 0.0)):
        a4 = np.full((nv, 4), w, np.float32)
        a4[:, :3] = arr[idx[..., 0], idx[..., 1]].reshape(-1, 3)
        out[slot] = a4
    B = np.zeros((nv, 4), np.float32)
    B[:, :3] = np.cross(out[3][:, :3], out[1][:, :3])
    out[2] = B
    out[4] = uv[idx[..., 0], idx[..., 1]].reshape(-1, 2).astype(np.float32)
    return out


def terrain(cells, size, seed, height=8.0, uv_repeat=8.0):
    rng = np.random.default_rng(seed)
    n = cells + 1
    hgt = value_noise(n, n, 6, rng, channels=1, octaves=6)[..., 0]
    hgt = (hgt - hgt.mean()) * height * 2.0
    xs = np.linspace(-size / 2, size / 2, n)
    zs = np.linspace(-size / 2, size / 2, n)
    X, Z = np.meshgrid(xs, zs)
    step = size / cells
    dhdx = np.gradient(hgt, step, axis=1)
    dhdz = np.gradient(hgt, step, axis=0)
    nrm = np.stack([-dhdx, np.ones_like(hgt), -dhdz], -1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    tan = np.stack([np.ones_like(hgt), dhdx, np.zeros_like(hgt)], -1)
    tan /= np.linalg.norm(tan, axis=-1, keepdims=True)
    bin_ = np.cross(nrm, tan)
    pos = np.stack([X, hgt, Z], -1)
    uv = np.stack([np.arange(n)[None, :] / cells * uv_repeat + 0 * Z, np.arange(n)[:, None] / cells * uv_repeat + 0 * X], -1)
    a = np.arange(cells)
    ii, jj = np.meshgrid(a, a, indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    # two triangles per cell, counter-clockwise seen from +y
    t1 = np.stack([np.stack([ii, jj], -1), np.stack([ii + 1, jj], -1), np.stack([ii + 1, jj + 1], -1)], 1)
    t2 = np.stack([np.stack([ii, jj], -1), np.stack([ii + 1, jj + 1], -1), np.stack([ii, jj + 1], -1)], 1)
    idx = np.stack([t1, t2], 1).reshape(-1, 3, 2)
    P3 = pos[idx[..., 0], idx[..., 1]]
    flip = _orient_ccw(P3, np.array([0.0, 1.0, 0.0]))
    idx[flip] = idx[flip][:, [0, 2, 1]]
    nv = idx.shape[0] * 3
    out = {}
    for slot, arr, w in ((0, pos, 1.0), (1, tan, 0.0), (2, bin_, 0.0), (3, nrm, 0.0)):
        a4 = np.full((nv, 4), w, np.float32)
        a4[:, :3] = arr[idx[..., 0], idx[..., 1]].reshape(-1, 3)
        out[slot] = a4
    out[4] = uv[idx[..., 0], idx[..., 1]].reshape(-1, 2).astype(np.float32)
    return out, hgt, xs, zs


# ---------------------------------------------------------------- workloads
def sphere_scene(width=640, height=480, seed=1, tex_size=512):
    """640x480 UV sphere (32x32, 2,048 triangles), Lambert x bilinear 512^2
    value-noise texture, directional light normalize(1,1,1), camera at z=3, 60 deg."""
    sc = Scene(width, height)
    sc.name = "uv_sphere_2048_%dx%d" % (width, height)
    sc.meshes["sphere"] = uv_sphere(32, 32, 1.0, seed=seed)
    sc.textures.append(TextureSpec([noise_texture(tex_size, seed)], wrap=N.WRAP, levels=1))
    P = perspective(np.radians(60.0), width / height, 0.1, 10.0)
    V = look_at((0, 0, 3), (0, 0, 0))
    sc.uniforms[3] = colmajor(P @ V)
    sc.uniforms[4] = colmajor(np.eye(4))
    sc.uniforms[5] = colmajor(np.eye(4))
    L = np.array([1.0, 1.0, 1.0]) / np.sqrt(3.0)
    sc.uniforms[7] = vec4(*L)
    sc.uniforms[9] = np.array([0], np.int32)
    sc.draws.append(DrawSpec("sphere", N.PROC_LAMBERT))
    return sc


def terrain_scene(width=1920, height=1080, seed=7, cells=300, n_spheres=64, tex_size=2048, spec_exp=32):
    """1080p terrain (300x300 cells, 180k triangles) + 64 instanced spheres of
    320 triangles, 2048^2 albedo + normal map with full mip chains, trilinear,
    per-pixel Blinn-Phong with normal mapping, one point light."""
    sc = Scene(width, height)
    sc.name = "terrain_%dk_%dx%d" % ((cells * cells * 2 + n_spheres * 320) // 1000, width, height)
    size = 100.0
    ter, hgt, xs, zs = terrain(cells, size, seed)
    sc.meshes["terrain"] = ter
    sc.meshes["sphere"] = uv_sphere(10, 16, 1.5, tangents=True)
    lv = mip_levels(tex_size, tex_size)
    sc.textures.append(TextureSpec([noise_texture(tex_size, seed + 1, cells=32, octaves=4)], wrap=N.WRAP, levels=lv))
    sc.textures.append(TextureSpec([normal_map(tex_size, seed + 2)], wrap=N.WRAP, levels=lv))
    eye = np.array([0.0, 22.0, 62.0])
    P = perspective(np.radians(60.0), width / height, 0.5, 200.0)
    V = look_at(eye, (0, 0, -5))
    sc.uniforms[3] = colmajor(P @ V)
    sc.uniforms[4] = colmajor(np.eye(4))
    sc.uniforms[5] = colmajor(np.eye(4))
    sc.uniforms[7] = vec4(30.0, 60.0, 40.0, 0.0)
    sc.uniforms[8] = vec4(*eye, 0.0)
    sc.uniforms[9] = np.array([0], np.int32)
    sc.uniforms[10] = np.array([1], np.int32)
    sc.uniforms[11] = np.array([spec_exp], np.int32)
    sc.draws.append(DrawSpec("terrain", N.PROC_BLINN))
    rng = np.random.default_rng(seed + 3)
    for k in range(n_spheres):
        x, z = rng.uniform(-size * 0.45, size * 0.45, 2)
        iz = int(np.clip(np.searchsorted(zs, z), 0, len(zs) - 1))
        ix = int(np.clip(np.searchsorted(xs, x), 0, len(xs) - 1))
        y = hgt[iz, ix] + 2.0
        sc.draws.append(DrawSpec("sphere", N.PROC_BLINN, {4: colmajor(translate(x, y, z))}))
    return sc


def shadow_terrain_scene(width=480, height=270, seed=7, cells=60, n_spheres=8, tex_size=256, shadow_size=256,
                         light=(30.0, 60.0, 40.0)):
    """The terrain scene lit through a shadow map: a DEF05 pass renders the
    terrain and spheres from the point light into a shadow_size^2 depth
    texture, then PSGPU_PROC_SHADOWED shades with the Lambert term times the
    projective sampler's factor.  Uniform 12 maps world positions to the
    shadow map: texture x, y in [0, 1] (a 0.5 scale and offset on the light's
    clip x, y) and the light's clip z, which is what the map stores."""
    sc = terrain_scene(width, height, seed, cells, n_spheres, tex_size)
    sc.name = "shadowed_" + sc.name
    lp = np.asarray(light, np.float64)
    sc.uniforms[7] = vec4(*lp, 0.0)
    LPV = perspective(np.radians(75.0), 1.0, 20.0, 200.0) @ look_at(lp, (0.0, 0.0, 0.0))
    bias = np.array([[0.5, 0, 0, 0.5], [0, 0.5, 0, 0.5], [0, 0, 1, 0], [0, 0, 0, 1]], np.float64)
    draws = [DrawSpec(d.mesh, N.PROC_DEF05, {4: d.uniforms.get(4, sc.uniforms[4])}) for d in sc.draws]
    sc.shadow = ShadowPass(shadow_size, colmajor(LPV), draws)
    for d in sc.draws:
        d.prog = N.PROC_SHADOWED
    sc.uniforms[12] = colmajor(bias @ LPV)
    sc.uniforms[13] = np.array([len(sc.textures)], np.int32)
    return sc


def cloud_texture(size, seed):
    """cloud density in the red channel (what PSGPU_PROC_CLOUD takes as alpha
    and PSGPU_PROC_CLOUD_SHADOW tests against 150); the other channels 255"""
    rng = np.random.default_rng(seed)
    n = value_noise(size, size, 8, rng, channels=1, octaves=4)[..., 0]
    img = np.full((size, size, 4), 255, np.uint8)
    img[..., 2] = (np.clip((n - 0.3) * 2.5, 0.0, 1.0) * 255.0).astype(np.uint8)
    return img


def cloud_terrain_scene(width=480, height=270, seed=7, cells=60, n_spheres=8, tex_size=256, shadow_size=256,
                        cloud_cells=24, cloud_y=14.0, cloud_tex=128):
    """The shadowed terrain under a cloud layer, drawn the way the reference's
    demo draws its clouds (test/puresoft.cpp:179-205, testobjs.cpp:93-112): a
    rolling sheet of cloud_cells^2 cells at height cloud_y casts shadows
    through PSGPU_PROC_CLOUD_SHADOW in the shadow pass (texels with red below
    150 discarded), then is drawn last with PSGPU_PROC_CLOUD and blending on:
    white lit by 2 (L . n), times the shadow factor, alpha = the texel's red."""
    sc = shadow_terrain_scene(width, height, seed, cells, n_spheres, tex_size, shadow_size)
    sc.name = "clouded_" + sc.name
    sheet = terrain(cloud_cells, 70.0, seed + 5, height=1.5, uv_repeat=1.0)[0]
    sc.meshes["cloud"] = {k: sheet[k] for k in (0, 3, 4)}
    ci = np.array([len(sc.textures)], np.int32)
    sc.textures.append(TextureSpec([cloud_texture(cloud_tex, seed + 6)], wrap=N.CLAMP))
    shadow_idx = np.array([len(sc.textures)], np.int32)   # the map follows the scene's textures
    sc.uniforms[13] = shadow_idx
    sc.uniforms[15] = shadow_idx
    sc.uniforms[16] = sc.uniforms[12]
    M = colmajor(translate(0.0, cloud_y, 0.0))
    sc.shadow.draws.append(DrawSpec("cloud", N.PROC_CLOUD_SHADOW, {4: M, 9: ci}))
    sc.draws.append(DrawSpec("cloud", N.PROC_CLOUD, {4: M, 9: ci}, behavior=N.DEFAULT_BEHAVIOR | N.ALPHABLEND))
    return sc


# ---------------------------------------------------------------- GPU upload
class GpuScene:
    """A Scene resident on one psgpu context (VBOs, VAOs, textures, programmes)."""

    def __init__(self, ctx, scene, gen_mips=True):
        self.ctx, self.scene = ctx, scene
        self.tex = []
        for t in scene.textures:
            idx = ctx.texture_create(t.width, t.height, 4, extra_layers=t.layers - 1, wrap=t.wrap,
                                     mip_levels=t.levels, pixels=t.images[0])
            for L in range(1, t.layers):
                ctx.texture_upload(idx, t.images[L], layer=L)
            if t.levels > 1 and gen_mips:
                ctx.texture_generate_mips(idx)
            self.tex.append(idx)
        self.vao = {}
        for name, slots in scene.meshes.items():
            v = ctx.vao_create()
            for s, arr in slots.items():
                ctx.vao_attach(v, s, ctx.vbo_create(arr))
            self.vao[name] = v
        self.progs = {}
        for d in scene.draws + (scene.shadow.draws if scene.shadow else []):
            if d.prog not in self.progs:
                self.progs[d.prog] = ctx.program_create(d.prog)
        self.shadow_tex = None
        if scene.shadow is not None:
            n = scene.shadow.size
            self.shadow_tex = ctx.texture_create(n, n, 4, wrap=N.CLAMP)

    def shadow_pass(self):
        """the shadow map: depth of the DEF05 draws, seen from the light"""
        ctx, sp = self.ctx, self.scene.shadow
        ctx.depth_target(self.shadow_tex)
        ctx.viewport(sp.size, sp.size)
        ctx.clear_depth(1.0)
        ctx.enable(N.DEFAULT_BEHAVIOR)
        ctx.uniform(3, sp.light_pv)
        cur = None
        for d in sp.draws:
            if d.prog != cur:
                ctx.program_use(self.progs[d.prog])
                cur = d.prog
            for k, v in d.uniforms.items():
                ctx.uniform(k, v)
            ctx.draw(self.vao[d.mesh])
        ctx.depth_target(N.DEFAULT_DEPTH)
        ctx.viewport(self.scene.width, self.scene.height)

    def frame(self, clear=True, flush=True):
        """clearDepth + clearColour + the scene's drawVAO calls (queued, then
        flushed unless flush=False)"""
        ctx, sc = self.ctx, self.scene
        if sc.shadow is not None:
            self.shadow_pass()
        if clear:
            ctx.clear_depth(sc.clear_depth)
            ctx.clear_colour(sc.clear_colour)
        for k, v in sc.uniforms.items():
            ctx.uniform(k, v)
        cur, beh = None, N.DEFAULT_BEHAVIOR
        for d in sc.draws:
            if d.prog != cur:
                ctx.program_use(self.progs[d.prog])
                cur = d.prog
            want = N.DEFAULT_BEHAVIOR if d.behavior is None else d.behavior
            if want != beh:
                ctx.enable(want)
                ctx.disable(~want & (N.UPDATE_DEPTH | N.TEST_DEPTH | N.FACE_CULLING | N.ALPHABLEND))
                beh = want
            for k, v in d.uniforms.items():
                ctx.uniform(k, v)
            ctx.draw(self.vao[d.mesh])
            for k in d.uniforms:
                if k in sc.uniforms:
                    ctx.uniform(k, sc.uniforms[k])
        if beh != N.DEFAULT_BEHAVIOR:
            ctx.enable(N.DEFAULT_BEHAVIOR)
            ctx.disable(~N.DEFAULT_BEHAVIOR & (N.UPDATE_DEPTH | N.TEST_DEPTH | N.FACE_CULLING | N.ALPHABLEND))
        if flush:
            ctx.flush()


# ---------------------------------------------------------------- batches of views
class SceneAsset:
    """A resident scene for batched views: one mesh, albedo + normal textures."""

    def __init__(self, mesh, textures, center, radius, name):
        self.mesh, self.textures, self.center, self.radius, self.name = mesh, textures, center, radius, name

    def triangles(self):
        return min(a.shape[0] for a in self.mesh.values()) // 3


def batch_assets(n_scenes=16, seed=11, tri_lo=10_000, tri_hi=500_000, tex_size=1024):
    """n_scenes seeded terrain patches with log-uniform triangle counts and
    log2(tex_size)+1-level mipmapped albedo + normal maps."""
    rng = np.random.default_rng(seed)
    assets = []
    lv = mip_levels(tex_size, tex_size)
    for k in range(n_scenes):
        T = float(np.exp(rng.uniform(np.log(tri_lo), np.log(tri_hi))))
        cells = max(2, int(round(np.sqrt(T / 2.0))))
        mesh, hgt, xs, zs = terrain(cells, 40.0, seed * 100 + k, height=4.0, uv_repeat=4.0)
        tex = [TextureSpec([noise_texture(tex_size, seed * 100 + k + 1, cells=16, octaves=3)], wrap=N.WRAP, levels=lv),
               TextureSpec([normal_map(tex_size, seed * 100 + k + 2, cells=16)], wrap=N.WRAP, levels=lv)]
        assets.append(SceneAsset(mesh, tex, np.array([0.0, 0.0, 0.0]), 28.0, "patch%d_%dk" % (k, cells * cells * 2 // 1000)))
    return assets


def tiled_texture(make, size, seed, tile=512):
    """a size x size texture that repeats a seeded periodic tile of side
    min(size, tile): full-size memory footprint, bounded generation time"""
    t = min(size, tile)
    img = make(t, seed)
    reps = size // t
    return np.ascontiguousarray(np.tile(img, (reps, reps, 1)))


def serving_assets(n_scenes=32, seed=21, tri_lo=1_000, tri_hi=1_000_000, tex_lo=512, tex_hi=4096):
    """resident scenes of the serving workload: terrain patches with
    log-uniform triangle counts and power-of-two textures with log-uniform
    side in [tex_lo, tex_hi], mipmapped"""
    rng = np.random.default_rng(seed)
    assets = []
    lo, hi = int(np.log2(tex_lo)), int(np.log2(tex_hi))
    for k in range(n_scenes):
        T = float(np.exp(rng.uniform(np.log(tri_lo), np.log(tri_hi))))
        cells = max(2, int(round(np.sqrt(T / 2.0))))
        ts = 1 << int(rng.integers(lo, hi + 1))
        lv = mip_levels(ts, ts)
        mesh, _, _, _ = terrain(cells, 40.0, seed * 100 + k, height=4.0, uv_repeat=4.0)
        tex = [TextureSpec([tiled_texture(lambda t, sd: noise_texture(t, sd, cells=8, octaves=3), ts, seed * 100 + k + 1)],
                           wrap=N.WRAP, levels=lv),
               TextureSpec([tiled_texture(lambda t, sd: normal_map(t, sd, cells=16), ts, seed * 100 + k + 2)],
                           wrap=N.WRAP, levels=lv)]
        assets.append(SceneAsset(mesh, tex, np.array([0.0, 0.0, 0.0]), 28.0,
                                 "srv%d_%dk_tex%d" % (k, cells * cells * 2 // 1000, ts)))
    return assets


class ViewSpec:
    def __init__(self, scene, width, height, eye, target, light, spec_exp=24):
        self.scene, self.width, self.height = scene, width, height
        self.eye, self.target, self.light, self.spec_exp = eye, target, light, spec_exp

    def uniforms(self):
        P = perspective(np.radians(60.0), self.width / self.height, 0.5, 200.0)
        V = look_at(self.eye, self.target)
        return {3: colmajor(P @ V), 4: colmajor(np.eye(4)), 5: colmajor(np.eye(4)),
                7: vec4(*self.light, 0.0), 8: vec4(*self.eye, 0.0), 9: np.array([0], np.int32),
                10: np.array([1], np.int32), 11: np.array([self.spec_exp], np.int32)}


def batch_views(assets, n_views=256, seed=12, side_lo=256, side_hi=1024, aspect_jitter=0.25):
    """views with area log-uniform in [side_lo^2, side_hi^2]; cameras seeded on
    an upper sphere around their scene"""
    rng = np.random.default_rng(seed)
    views = []
    for i in range(n_views):
        a = float(np.exp(rng.uniform(np.log(side_lo * side_lo), np.log(side_hi * side_hi))))
        asp = float(np.exp(rng.uniform(-aspect_jitter, aspect_jitter)))
        w = max(8, int(round(np.sqrt(a * asp))))
        h = max(8, int(round(a / w)))
        s = int(rng.integers(0, len(assets)))
        th = rng.uniform(0, 2 * np.pi)
        el = rng.uniform(0.35, 1.1)
        r = assets[s].radius * rng.uniform(1.4, 2.0)
        eye = assets[s].center + r * np.array([np.cos(el) * np.cos(th), np.sin(el), np.cos(el) * np.sin(th)])
        light = assets[s].center + np.array([20.0, 50.0, 10.0])
        views.append(ViewSpec(s, w, h, eye, assets[s].center, light))
    return views


def poisson_requests(assets, n=2000, rate=2000.0, seed=23, side_lo=128, side_hi=2048):
    """n view requests with exponential inter-arrival times (rate per second),
    area log-uniform in [side_lo^2, side_hi^2], a uniformly chosen scene and a
    seeded camera on the upper sphere around it.  Returns (views, arrivals)."""
    rng = np.random.default_rng(seed)
    views = batch_views(assets, n_views=n, seed=seed + 1, side_lo=side_lo, side_hi=side_hi)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    arrivals -= arrivals[0]
    return views, arrivals


class GpuBatch:
    """Assets resident on one context plus per-view colour/depth textures.
    submit() queues every view (targets, viewport, clears, camera, draw); the
    library renders all of them in one pass of its kernels."""

    def __init__(self, ctx, assets, views, gen_mips=True):
        self.ctx, self.assets, self.views = ctx, assets, views
        self.tex = []
        self.vao = []
        for a in assets:
            ids = []
            for t in a.textures:
                idx = ctx.texture_create(t.width, t.height, 4, wrap=t.wrap, mip_levels=t.levels, pixels=t.images[0])
                if t.levels > 1 and gen_mips:
                    ctx.texture_generate_mips(idx)
                ids.append(idx)
            self.tex.append(ids)
            v = ctx.vao_create()
            for s, arr in a.mesh.items():
                ctx.vao_attach(v, s, ctx.vbo_create(arr))
            self.vao.append(v)
        self.prog = ctx.program_create(N.PROC_BLINN)
        self.targets = []
        for vw in views:
            c = ctx.texture_create(vw.width, vw.height, 4, wrap=N.CLAMP)
            d = ctx.texture_create(vw.width, vw.height, 4, wrap=N.CLAMP)
            self.targets.append((c, d))

    def submit(self, indices=None, clear_bgra=0x00302010):
        ctx = self.ctx
        ctx.program_use(self.prog)
        for i in (range(len(self.views)) if indices is None else indices):
            vw = self.views[i]
            c, d = self.targets[i]
            ctx.target(c)
            ctx.depth_target(d)
            ctx.viewport(vw.width, vw.height)
            ctx.clear_depth(1.0)
            ctx.clear_target(clear_bgra)
            u = vw.uniforms()
            u[9] = np.array([self.tex[vw.scene][0]], np.int32)
            u[10] = np.array([self.tex[vw.scene][1]], np.int32)
            for k, val in u.items():
                ctx.uniform(k, val)
            ctx.draw(self.vao[vw.scene])
        ctx.target(N.DISPLAY)
        ctx.depth_target(N.DEFAULT_DEPTH)
        ctx.viewport(ctx.width, ctx.height)

    def read(self, i):
        c, d = self.targets[i]
        vw = self.views[i]
        col = self.ctx.texture_download(c).reshape(vw.height, vw.width, 4)
        dep = self.ctx.texture_download(d).view(np.float32).reshape(vw.height, vw.width)
        return col, dep

    def view_scene(self, i):
        """the view as a stand-alone Scene (for the oracle or a one-view render)"""
        return view_as_scene(self.assets, self.views[i], "view%d" % i)


def view_as_scene(assets, vw, tag="view"):
    """a view of a resident asset as a stand-alone Scene (textures 0 and 1 are
    the asset's albedo and normal map)"""
    a = assets[vw.scene]
    sc = Scene(vw.width, vw.height)
    sc.name = "%s_%s_%dx%d" % (tag, a.name, vw.width, vw.height)
    sc.meshes["m"] = a.mesh
    sc.textures = list(a.textures)
    sc.uniforms = vw.uniforms()
    sc.draws.append(DrawSpec("m", N.PROC_BLINN))
    return sc


# ---------------------------------------------------------------- scene files
SCENE_MAGIC = 0x43535350   # "PSSC"


def save_scene(scene, path):
    """Write a Scene as a flat little-endian file that the C++ facade's
    psrender program reads (puresoft3d_amd/host/psrender.cpp):

      u32 magic, u32 version=1, i32 width, height, f32 clear_depth, u32 clear_colour
      i32 n_textures; per texture: i32 width, height, layers, wrap, levels, then
          layers * height * width * 4 bytes (BGRA8, layer-major)
      i32 n_meshes; per mesh: i32 n_slots; per slot: i32 slot, comps, count,
          then count * comps float32
      i32 n_uniforms; per uniform: i32 slot, i32 bytes, then the bytes
      i32 n_draws; per draw: i32 mesh, prog, behavior (-1 = default),
          i32 n_uniforms; per uniform as above
    """
    import struct
    names = list(scene.meshes)
    out = [struct.pack("<IIiifI", SCENE_MAGIC, 1, scene.width, scene.height, scene.clear_depth,
                       scene.clear_colour & 0xFFFFFFFF)]

    def uniforms(u):
        parts = [struct.pack("<i", len(u))]
        for k, v in sorted(u.items()):
            b = np.ascontiguousarray(v).tobytes()
            parts += [struct.pack("<ii", k, len(b)), b]
        return parts

    out.append(struct.pack("<i", len(scene.textures)))
    for t in scene.textures:
        out.append(struct.pack("<iiiii", t.width, t.height, t.layers, t.wrap, t.levels))
        out += [im.tobytes() for im in t.images]
    out.append(struct.pack("<i", len(names)))
    for nm in names:
        m = scene.meshes[nm]
        out.append(struct.pack("<i", len(m)))
        for s, a in sorted(m.items()):
            a = np.ascontiguousarray(a, np.float32)
            out += [struct.pack("<iii", s, a.shape[1], a.shape[0]), a.tobytes()]
    out += uniforms(scene.uniforms)
    out.append(struct.pack("<i", len(scene.draws)))
    for d in scene.draws:
        beh = -1 if d.behavior is None else d.behavior
        out.append(struct.pack("<iii", names.index(d.mesh), d.prog, beh))
        out += uniforms(d.uniforms)
    with open(path, "wb") as f:
        f.write(b"".join(out))
