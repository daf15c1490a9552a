"""Seeded micro-scenes for parity tests: random triangles in clip space for each
built-in processor, with the edge cases the reference's rasterizer and depth
test are sensitive to (flat rows, off-screen spans, z at the +-1 limits,
coplanar overlaps for the 1e-4 depth epsilon, degenerate triangles)."""
import numpy as np

from puresoft3d_amd import native as N
from puresoft3d_amd import scenes as S


def _slots_for(prog, pos, rng):
    n = pos.shape[0]
    def f4(w=0.0):
        a = rng.uniform(-1, 1, (n, 4)).astype(np.float32)
        a[:, 3] = w
        return a
    uv = rng.uniform(-0.25, 1.25, (n, 2)).astype(np.float32)
    if prog in (N.PROC_DEF01, N.PROC_LAMBERT):
        return {0: pos, 3: f4(), 4: uv}
    if prog == N.PROC_DEF02:
        col = rng.uniform(0, 255, (n, 4)).astype(np.float32)
        return {0: pos, 1: f4(), 2: col}
    if prog in (N.PROC_DEF03, N.PROC_BLINN, N.PROC_SHADOWED):
        return {0: pos, 1: f4(), 2: f4(), 3: f4(), 4: uv}
    if prog == N.PROC_CLOUD:
        nrm = f4()
        nrm[::7] *= 150.0   # 2 (L . n) 255 beyond 32767: the signed byte pack gives 0
        return {0: pos, 3: nrm, 4: uv}
    if prog == N.PROC_CLOUD_SHADOW:
        return {0: pos, 4: uv}
    return {0: pos}


def random_clip_triangles(rng, n, kind="mixed"):
    """n triangles as (3n, 4) clip positions (w > 0)"""
    P = np.zeros((n, 3, 4), np.float32)
    for t in range(n):
        c = rng.integers(0, 6) if kind == "mixed" else kind
        if c == 0:      # ordinary, partially off screen
            xy = rng.uniform(-1.3, 1.3, (3, 2))
            z = rng.uniform(-0.95, 0.95, 3)
            w = rng.uniform(0.5, 2.0, 3)
        elif c == 1:    # tiny
            ctr = rng.uniform(-1, 1, 2)
            xy = ctr + rng.uniform(-0.02, 0.02, (3, 2))
            z = np.full(3, rng.uniform(-0.9, 0.9))
            w = np.ones(3)
        elif c == 2:    # flat top/bottom on a pixel row
            xy = rng.uniform(-1, 1, (3, 2))
            xy[1, 1] = xy[0, 1]
            z = rng.uniform(-0.9, 0.9, 3)
            w = np.ones(3)
        elif c == 3:    # z touching the limits
            xy = rng.uniform(-1, 1, (3, 2))
            z = rng.choice([-1.0, 1.0, -0.99999, 0.99999, 0.0], 3)
            w = np.ones(3)
        elif c == 4:    # huge, extends far off screen
            xy = rng.uniform(-40, 40, (3, 2))
            z = rng.uniform(-0.5, 0.5, 3)
            w = np.ones(3)
        else:           # degenerate: repeated vertex
            xy = rng.uniform(-1, 1, (3, 2))
            xy[2] = xy[0]
            z = rng.uniform(-0.9, 0.9, 3)
            w = np.ones(3)
        P[t, :, 0] = xy[:, 0] * w
        P[t, :, 1] = xy[:, 1] * w
        P[t, :, 2] = z * w
        P[t, :, 3] = w
    return P.reshape(-1, 4)


def coplanar_pairs(rng, n):
    """pairs of overlapping triangles at nearly equal depth (exercise z - d < -1e-4)"""
    out = []
    for _ in range(n):
        xy = rng.uniform(-0.8, 0.8, (3, 2))
        z = rng.uniform(-0.5, 0.5)
        dz = rng.choice([0.0, 5e-5, -5e-5, 1.5e-4, -1.5e-4, 1e-4])
        for zz in (z, z + dz):
            tri = np.zeros((3, 4), np.float32)
            tri[:, :2] = xy + rng.uniform(-0.05, 0.05, (3, 2))
            tri[:, 2] = zz
            tri[:, 3] = 1.0
            out.append(tri)
    return np.concatenate(out, 0)


def add_depth_map(sc, rng, tex, tex_slot, pv_slot):
    """a random FP32 depth map (some NaN and inf texels) as the scene's next
    texture, named by uniform tex_slot, and in uniform pv_slot a random
    projection that reaches beyond the map on every side"""
    dm = rng.uniform(-1.2, 1.2, (tex, tex + 7)).astype(np.float32)
    dm.ravel()[rng.integers(0, dm.size, 6)] = [np.nan, np.inf, -np.inf, np.nan, 0.0, -0.0]
    img = dm.view(np.uint8).reshape(tex, tex + 7, 4)
    sc.uniforms[tex_slot] = np.array([len(sc.textures)], np.int32)
    sc.textures.append(S.TextureSpec([img], wrap=int(rng.integers(0, 2))))
    sc.uniforms[pv_slot] = S.colmajor(np.eye(4) * 0.8 + rng.uniform(-0.4, 0.4, (4, 4)))


def micro_scene(prog, seed, width=96, height=64, n=120, tex=32):
    rng = np.random.default_rng(seed)
    sc = S.Scene(width, height)
    sc.name = "micro_p%d_s%d_%dx%d" % (prog, seed, width, height)
    pos = np.concatenate([random_clip_triangles(rng, n), coplanar_pairs(rng, n // 6)], 0)
    sc.meshes["m"] = _slots_for(prog, pos, rng)
    I = S.colmajor(np.eye(4))
    if prog == N.PROC_DEF04:
        layers = [S.noise_texture(tex, seed + 10 + L, cells=4) for L in range(6)]
        sc.textures.append(S.TextureSpec(layers, wrap=N.CLAMP))
        R = np.eye(4)
        ang = rng.uniform(0, 2 * np.pi)
        R[0, 0], R[0, 2], R[2, 0], R[2, 2] = np.cos(ang), np.sin(ang), -np.sin(ang), np.cos(ang)
        sc.uniforms[0] = I
        sc.uniforms[1] = S.colmajor(R)
        sc.uniforms[2] = np.array([0], np.int32)
    else:
        M = np.eye(4)
        M[:3, :3] += rng.uniform(-0.05, 0.05, (3, 3))
        sc.uniforms[3] = I
        sc.uniforms[4] = S.colmajor(M)
        sc.uniforms[5] = S.colmajor(np.eye(4) + np.diag(rng.uniform(-0.2, 0.2, 4)))
        sc.uniforms[7] = S.vec4(*rng.uniform(-3, 3, 3), 0.0) if prog != N.PROC_LAMBERT else S.vec4(0.48, 0.6, 0.64)
        sc.uniforms[8] = S.vec4(*rng.uniform(-3, 3, 3), 0.0)
        lv = S.mip_levels(tex, tex) if prog in (N.PROC_BLINN, N.PROC_SHADOWED) else 1
        sc.textures.append(S.TextureSpec([S.noise_texture(tex, seed + 1, cells=4)], wrap=int(rng.integers(0, 2)), levels=lv))
        sc.textures.append(S.TextureSpec([S.normal_map(tex, seed + 2, cells=4)], wrap=N.WRAP, levels=lv))
        sc.uniforms[9] = np.array([0], np.int32)
        sc.uniforms[10] = np.array([1], np.int32)
        sc.uniforms[11] = np.array([int(rng.integers(1, 64))], np.int32)
        if prog == N.PROC_SHADOWED:
            add_depth_map(sc, rng, tex, 13, 12)
        if prog == N.PROC_CLOUD:
            add_depth_map(sc, rng, tex, 15, 16)
    beh = None
    if prog == N.PROC_CLOUD and seed % 2:
        beh = N.DEFAULT_BEHAVIOR | N.ALPHABLEND
    sc.draws.append(S.DrawSpec("m", prog, behavior=beh))
    sc.clear_colour = 0x00203040
    return sc


def behavior_scene(seed, width=96, height=64, n=60):
    """draws of random triangles under every combination of the depth-update,
    depth-test and face-culling behaviours (pipeline.cpp:324-332), DEF02 and
    Blinn-Phong interleaved, so later draws land on the earlier ones' depth"""
    rng = np.random.default_rng(seed)
    sc = micro_scene(N.PROC_BLINN, seed, width, height, n)
    sc.name = "behaviors_s%d_%dx%d" % (seed, width, height)
    sc.draws = []
    for beh in range(8):
        for prog in (N.PROC_DEF02, N.PROC_BLINN):
            name = "m%d_%d" % (beh, prog)
            pos = random_clip_triangles(rng, n // 4)
            sc.meshes[name] = _slots_for(prog, pos, rng)
            sc.draws.append(S.DrawSpec(name, prog, behavior=beh))
    sc.draws.insert(0, S.DrawSpec("m", N.PROC_BLINN))
    return sc


def blend_scene(seed, width=96, height=64, n=60):
    """the two programmes shaded in the raster walk (PSGPU_PROC_CLOUD with and
    without blending, PSGPU_PROC_CLOUD_SHADOW's discard) between Blinn-Phong
    draws, under every combination of the depth behaviours, so blended and
    discarded fragments land on one another's depth in submission order"""
    rng = np.random.default_rng(seed)
    sc = micro_scene(N.PROC_BLINN, seed, width, height, n)
    add_depth_map(sc, rng, 32, 15, 16)
    sc.name = "blends_s%d_%dx%d" % (seed, width, height)
    sc.draws = [S.DrawSpec("m", N.PROC_BLINN)]
    for beh in range(8):
        for prog, extra in ((N.PROC_CLOUD, N.ALPHABLEND), (N.PROC_BLINN, 0), (N.PROC_CLOUD_SHADOW, N.ALPHABLEND * (beh & 1)),
                            (N.PROC_CLOUD, 0)):
            name = "m%d_%d_%d" % (beh, prog, extra)
            pos = random_clip_triangles(rng, n // 4)
            sc.meshes[name] = _slots_for(prog, pos, rng)
            sc.draws.append(S.DrawSpec(name, prog, behavior=beh | extra))
    return sc


def big_triangle_scene(prog, seed, width=512, height=384, n=12):
    """a few triangles spanning most of the screen, overlapping at varying
    depths: spans hundreds of pixels long cross many bin chunks, so every
    chunk after the first starts from the serial accumulators carried
    across the span"""
    rng = np.random.default_rng(seed)
    sc = micro_scene(prog, seed, width, height, n=12)
    sc.name = "big_p%d_s%d_%dx%d" % (prog, seed, width, height)
    P = np.zeros((n, 3, 4), np.float32)
    for t in range(n):
        xy = rng.uniform(-1.6, 1.6, (3, 2))
        z = rng.uniform(-0.9, 0.9, 3)
        w = rng.uniform(0.6, 1.8, 3)
        P[t, :, 0] = xy[:, 0] * w
        P[t, :, 1] = xy[:, 1] * w
        P[t, :, 2] = z * w
        P[t, :, 3] = w
    sc.meshes["m"] = _slots_for(prog, P.reshape(-1, 4), rng)
    sc.draws[0].behavior = N.UPDATE_DEPTH | N.TEST_DEPTH   # both windings drawn
    return sc


def sliver_scene(prog, seed, width=160, height=96, n=1000):
    """triangles lying wholly left or right of the screen whose first row is
    below their lowest vertex: the row range truncates y toward zero, so that
    row evaluates the edges by extrapolation (rasterizer.cpp:65-68,149-163),
    and a long, nearly flat edge carries the span back onto the screen"""
    rng = np.random.default_rng(seed)
    sc = micro_scene(prog, seed, width, height, n=12)
    sc.name = "sliver_p%d_s%d_%dx%d" % (prog, seed, width, height)
    hw, hh = width // 2, height // 2
    P = np.zeros((n, 3, 4), np.float32)
    for t in range(n):
        side = -1.0 if rng.random() < 0.5 else 1.0
        yb = float(rng.integers(2, height - 4)) + rng.uniform(0.5, 0.98)
        xb = -rng.uniform(3, 60) if side < 0 else width + rng.uniform(3, 60)
        dy = rng.uniform(0.05, 0.6)
        far = (xb + side * rng.uniform(100, 400), yb + dy)     # long, flat edge away from the screen
        third = (xb + side * rng.uniform(0, 20), yb + rng.uniform(1.0, 3.0))
        pts = [(xb, yb), far, third]
        if rng.random() < 0.5:
            pts = [pts[0], pts[2], pts[1]]
        for v, (X, Y) in enumerate(pts):
            P[t, v, 0] = (X - hw) / hw
            P[t, v, 1] = (Y - hh) / hh
        P[t, :, 2] = rng.uniform(-0.5, 0.5)
        P[t, :, 3] = 1.0
    sc.meshes["m"] = _slots_for(prog, P.reshape(-1, 4), rng)
    sc.draws[0].behavior = N.UPDATE_DEPTH | N.TEST_DEPTH
    return sc


def objx_scene(which, width=640, height=480, pictures=False):
    """the reference's own assets (src/test/sphere.objx, src/test2/plane.objx,
    kept gzipped in tests/golden) through the OBJX loader.  sphere.objx's one
    mesh has no 'object/' prefix, which the demo loader skips; it is drawn here
    as 'ball/sphere' with a diffuse file, so that it takes the textured
    tangent-space programme."""
    import gzip
    import os
    from puresoft3d_amd import objx as X
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "%s.objx.gz" % which)
    with gzip.open(path, "rb") as f:
        sc = X.read_objx(f.read())
    pdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pictures") if pictures else None
    if which == "sphere":
        # with pictures: the demo's Earth (test/testobjs.cpp:12-15), its diffuse and bump maps
        m = sc.meshes[0]
        sc.meshes = [X.ObjxMesh("ball/" + m.name, m.vertices, m.normals, m.tangents, m.texcoords, m.indices, m.ambient,
                                m.diffuse, m.specular, m.specular_exponent, "earth.jpg" if pictures else "sphere.jpg",
                                bump_file="earth.dot3.png" if pictures else "")]
    return X.load_scene(sc, width, height, name="objx_%s%s_%dx%d" % (which, "_pictures" if pictures else "", width,
                                                                      height),
                        picture_dir=pdir, picture_wrap=N.WRAP if which == "sphere" else N.CLAMP)


def cloud_quad_scene(shadow_depth, normal_scale, red, bg=0x00406080, w=64, h=48):
    """one clip-space quad of PSGPU_PROC_CLOUD under blending: normal (0, 0, 1)
    times normal_scale, a light far along +z (L . n = normal_scale to FP32
    precision), a constant texel with the given red, a constant shadow map"""
    sc = S.Scene(w, h)
    q = np.array([[-0.8, -0.7, 0.0, 1.0], [0.8, -0.7, 0.0, 1.0], [0.8, 0.7, 0.0, 1.0],
                  [-0.8, -0.7, 0.0, 1.0], [0.8, 0.7, 0.0, 1.0], [-0.8, 0.7, 0.0, 1.0]], np.float32)
    nrm = np.tile(np.float32([0.0, 0.0, normal_scale, 0.0]), (6, 1))
    uv = np.clip(q[:, :2] * 0.5 + 0.5, 0, 1).astype(np.float32)
    sc.meshes["q"] = {0: q, 3: nrm, 4: uv}
    tex = np.zeros((4, 4, 4), np.uint8)
    tex[..., 2] = red
    sc.textures.append(S.TextureSpec([tex], wrap=N.CLAMP))
    dm = np.full((16, 16), shadow_depth, np.float32)
    sc.textures.append(S.TextureSpec([dm.view(np.uint8).reshape(16, 16, 4)], wrap=N.CLAMP))
    bias = np.array([[0.5, 0, 0, 0.5], [0, 0.5, 0, 0.5], [0, 0, 1, 0], [0, 0, 0, 1]], np.float64)
    sc.uniforms.update({3: S.colmajor(np.eye(4)), 4: S.colmajor(np.eye(4)), 5: S.colmajor(np.eye(4)), 7: S.vec4(0.0, 0.0, 1e6, 0.0), 9: np.array([0], np.int32),
                        15: np.array([1], np.int32), 16: S.colmajor(bias)})
    sc.draws.append(S.DrawSpec("q", N.PROC_CLOUD, behavior=N.UPDATE_DEPTH | N.TEST_DEPTH | N.ALPHABLEND))
    sc.clear_colour = bg
    return sc


def cloud_shadow_quad(red, w=64, h=48):
    """one clip-space quad of PSGPU_PROC_CLOUD_SHADOW at z = 0.25 over a constant texel"""
    sc = S.Scene(w, h)
    q = np.array([[-0.9, -0.9, 0.25, 1.0], [0.9, -0.9, 0.25, 1.0], [0.9, 0.9, 0.25, 1.0],
                  [-0.9, -0.9, 0.25, 1.0], [0.9, 0.9, 0.25, 1.0], [-0.9, 0.9, 0.25, 1.0]], np.float32)
    sc.meshes["q"] = {0: q, 4: np.clip(q[:, :2] * 0.5 + 0.5, 0, 1).astype(np.float32)}
    tex = np.zeros((4, 4, 4), np.uint8)
    tex[..., 2] = red
    sc.textures.append(S.TextureSpec([tex], wrap=N.CLAMP))
    sc.uniforms.update({3: S.colmajor(np.eye(4)), 4: S.colmajor(np.eye(4)), 9: np.array([0], np.int32)})
    sc.draws.append(S.DrawSpec("q", N.PROC_CLOUD_SHADOW, behavior=N.UPDATE_DEPTH | N.TEST_DEPTH))
    sc.clear_colour = 0x00112233
    return sc


def reference_shader_scene(prog, width=320, height=240, seed=3):
    """a lit sphere through one of the reference's built-in programmes:
    DEF01 (textured), DEF02 (per-vertex colour), DEF03 (textured + bump map)
    with a point light and a camera near it, or DEF04 (a skybox quad filling
    the view, seen through a turned view matrix)"""
    rng = np.random.default_rng(seed)
    sc = S.Scene(width, height)
    sc.name = "refshader_p%d_%dx%d" % (prog, width, height)
    if prog == N.PROC_DEF04:
        a = 1.0
        sc.meshes["sky"] = {0: np.array([[-a, -a, 0.99, 1], [a, -a, 0.99, 1], [a, a, 0.99, 1],
                                         [-a, -a, 0.99, 1], [a, a, 0.99, 1], [-a, a, 0.99, 1]], np.float32)}
        sc.textures.append(S.TextureSpec([S.noise_texture(128, seed + L, cells=8) for L in range(6)], wrap=N.CLAMP))
        V = S.look_at((0, 0, 0), (0.6, 0.3, -1.0))
        sc.uniforms[0] = S.colmajor(S.perspective(np.radians(60.0), width / height, 0.1, 10.0))
        sc.uniforms[1] = S.colmajor(V)
        sc.uniforms[2] = np.array([0], np.int32)
        sc.draws.append(S.DrawSpec("sky", prog))
        return sc
    m = S.uv_sphere(48, 48, 1.0, seed=seed, tangents=True)
    if prog == N.PROC_DEF01:
        sc.meshes["s"] = {0: m[0], 3: m[3], 4: m[4]}
    elif prog == N.PROC_DEF02:
        col = rng.uniform(0, 255, (m[0].shape[0], 4)).astype(np.float32)
        sc.meshes["s"] = {0: m[0], 1: m[3], 2: col}
    else:
        sc.meshes["s"] = m
    P = S.perspective(np.radians(60.0), width / height, 0.1, 10.0)
    V = S.look_at((0.4, 0.3, 2.6), (0, 0, 0))
    R = np.eye(4)
    ang = 0.7
    R[0, 0], R[0, 2], R[2, 0], R[2, 2] = np.cos(ang), np.sin(ang), -np.sin(ang), np.cos(ang)
    sc.uniforms[3] = S.colmajor(P @ V)
    sc.uniforms[4] = S.colmajor(S.translate(0.1, -0.05, 0.0) @ R)
    sc.uniforms[5] = S.colmajor(R)
    sc.uniforms[7] = S.vec4(1.5, 2.0, 2.5, 0.0)
    sc.uniforms[8] = S.vec4(0.4, 0.3, 2.6, 0.0)
    sc.textures.append(S.TextureSpec([S.noise_texture(256, seed)], wrap=N.WRAP))
    sc.textures.append(S.TextureSpec([S.normal_map(256, seed + 1)], wrap=N.WRAP))
    sc.uniforms[9] = np.array([0], np.int32)
    sc.uniforms[10] = np.array([1], np.int32)
    sc.draws.append(S.DrawSpec("s", prog))
    sc.clear_colour = 0x00102030
    return sc


def hazard_triangles(rng, n):
    """triangles the reference never clips (drawvao.cpp:39-63): one vertex
    with w at or below zero (0, -0.5, +-1e-30; the divide gives infinities,
    NaN or a mirrored vertex), or a coordinate that is NaN or infinite"""
    P = random_clip_triangles(rng, n, kind=0).reshape(n, 3, 4)
    for t in range(n):
        v = int(rng.integers(0, 3))
        if t % 2 == 0:
            w = rng.choice([0.0, -0.0, -0.5, 1e-30, -1e-30])
            P[t, v, 3] = w
            if rng.integers(0, 3) == 0:
                P[t, v, 2] = 0.0          # with w = 0: 0 * inf is NaN, which passes the z test
        else:
            P[t, v, int(rng.integers(0, 4))] = rng.choice([np.nan, np.inf, -np.inf])
    return P.reshape(-1, 4)


def hazard_scene(prog, seed, extra_vertices=0, width=96, height=64, n=90):
    """a micro scene whose triangles include the hazards above, ordinary ones
    around them, and a vertex stream of 3k + extra_vertices entries: the
    trailing partial triangle is dropped (vertthrd.cpp:25-29)"""
    sc = micro_scene(prog, seed, width, height, n=n)
    rng = np.random.default_rng(seed + 1000)
    m = sc.meshes["m"]
    pos = np.concatenate([m[0], hazard_triangles(rng, n // 2)], 0)
    rest = _slots_for(prog, pos, rng)
    for k in rest:
        if k != 0:
            rest[k][: m[k].shape[0]] = m[k]
    rest[0] = pos
    if extra_vertices:
        for k in rest:
            tail = rest[k][:extra_vertices].copy()
            rest[k] = np.concatenate([rest[k], tail], 0)
    sc.meshes["m"] = rest
    sc.name = "hazard_p%d_s%d_x%d" % (prog, seed, extra_vertices)
    return sc


def skybox_scene(width=320, height=240, turn=0.6):
    """the demo's skybox (test/testobjs.cpp:185-198): DEF04 over a full-view
    quad, the six purplenebula pictures as layers +X, -X, +Y, -Y, +Z, -Z
    (rt, lf, up, dn, ft, bk), loaded as the reference's loader leaves them,
    seen through a view turned about y and x"""
    import os
    from puresoft3d_amd import objx as X
    pdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pictures")
    faces = [X.load_picture(os.path.join(pdir, "purplenebula_%s.png" % k)) for k in ("rt", "lf", "up", "dn", "ft", "bk")]
    sc = S.Scene(width, height)
    sc.name = "skybox_purplenebula_%dx%d" % (width, height)
    z = 0.999
    sc.meshes["sky"] = {0: np.array([[-1, -1, z, 1], [1, -1, z, 1], [1, 1, z, 1],
                                     [-1, -1, z, 1], [1, 1, z, 1], [-1, 1, z, 1]], np.float32)}
    sc.textures.append(S.TextureSpec(faces, wrap=N.CLAMP))
    c, s_ = np.cos(turn), np.sin(turn)
    Ry = np.array([[c, 0, s_, 0], [0, 1, 0, 0], [-s_, 0, c, 0], [0, 0, 0, 1]])
    c2, s2 = np.cos(0.35), np.sin(0.35)
    Rx = np.array([[1, 0, 0, 0], [0, c2, -s2, 0], [0, s2, c2, 0], [0, 0, 0, 1]])
    sc.uniforms[0] = S.colmajor(np.eye(4))
    sc.uniforms[1] = S.colmajor(Rx @ Ry)
    sc.uniforms[2] = np.array([0], np.int32)
    sc.draws.append(S.DrawSpec("sky", N.PROC_DEF04))
    return sc
