"""Synthetic scenes for the benchmark configurations, built through the renderer API.

Parameters the configurations leave open are fixed here once (seeds, camera, light):
  plain small  : 512x512 RGBA8 + RG32F depth/stencil, 1 sample, 2,000 non-indexed triangles,
                 clip-space x,y ~ U[-1,1], z ~ U[0.1,0.9], w = 1, RGBA ~ U[0,1]^4, seed 0x5A11A001;
                 pass-through vertex shader, Gouraud pixel shader, depth LESS, blend off, cull none.
  textured mesh: 1920x1080 RGBA8 4x MSAA + resolve; 512x512-vertex grid over [-1,1]^2 in the xy
                 plane, z = 0.1 sin(8x) cos(8y), analytic normal facing -z, uv = grid/511, 32-bit
                 indices (522,242 triangles); 2048^2 RGBA8 uniform-noise texture (seed 0x5A11A002)
                 with the full box-filtered mip chain; min/mag linear, mip anisotropic x16, wrap;
                 Lambert with L = normalize(1,1,-1); depth LESS; cull none; camera at
                 (0, 1.5, -1.5) looking at the origin, vertical fov pi/2, near 0.1, far 100 (the
                 grid then covers ~11% of the frame at 45 degrees: ~0.45 pixels per triangle).
                 The "close" camera, (0, 0.8, -0.6) looking at (0, 0.5, 0), fills ~80% of the frame
                 with a flatter view (~4 pixels per triangle, more shading, less anisotropy).
  multipass    : shadow map + shaded, blended and mirrored passes over 200,000 random triangles
                 (seed 0x5A11A003; setup_multipass below has the exact recipe).
Random numbers come from splitmix64; a float in [0,1) is (u >> 40) * 2^-24.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import api as A

SEED_PLAIN = 0x5A11A001
SEED_TEXTURE = 0x5A11A002
MASK64 = (1 << 64) - 1


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n consecutive outputs of splitmix64 started at `seed`."""
    with np.errstate(over="ignore"):
        z = np.arange(1, n + 1, dtype=np.uint64)
        z *= np.uint64(0x9E3779B97F4A7C15)
        z += np.uint64(seed & MASK64)
        t = z >> np.uint64(30)
        z ^= t
        z *= np.uint64(0xBF58476D1CE4E5B9)
        np.right_shift(z, np.uint64(27), out=t)
        z ^= t
        z *= np.uint64(0x94D049BB133111EB)
        np.right_shift(z, np.uint64(31), out=t)
        z ^= t
        return z


def uniform01(seed: int, n: int) -> np.ndarray:
    return (splitmix64(seed, n) >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)


# ---- plain small ------------------------------------------------------------------------
def plain_triangles(n_tris: int, seed: int = SEED_PLAIN) -> np.ndarray:
    """(3*n_tris, 8) float32: x y z w r g b a per vertex."""
    u = uniform01(seed, n_tris * 3 * 7).reshape(n_tris * 3, 7)
    v = np.empty((n_tris * 3, 8), np.float64)
    v[:, 0] = -1.0 + 2.0 * u[:, 0]
    v[:, 1] = -1.0 + 2.0 * u[:, 1]
    v[:, 2] = 0.1 + 0.8 * u[:, 2]
    v[:, 3] = 1.0
    v[:, 4:8] = u[:, 3:7]
    return v.astype(np.float32)


# ---- textured mesh ----------------------------------------------------------------------
def grid_mesh(n: int = 512):
    """Vertices (n*n, 8) float32 [pos.xyz, normal.xyz, uv] and uint32 indices (2*(n-1)^2*3)."""
    t = np.arange(n, dtype=np.float64) / (n - 1)
    c = -1.0 + 2.0 * t
    y, x = np.meshgrid(c, c, indexing="ij")  # row j = y, column i = x
    z = 0.1 * np.sin(8 * x) * np.cos(8 * y)
    dzdx = 0.8 * np.cos(8 * x) * np.cos(8 * y)
    dzdy = -0.8 * np.sin(8 * x) * np.sin(8 * y)
    nrm = np.stack([dzdx, dzdy, -np.ones_like(z)], -1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    vu, uu = np.meshgrid(t, t, indexing="ij")
    verts = np.concatenate([np.stack([x, y, z], -1), nrm, np.stack([uu, vu], -1)], -1).reshape(n * n, 8)
    j, i = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    v00 = (j * n + i).ravel()
    v10 = v00 + 1
    v01 = v00 + n
    v11 = v01 + 1
    tri = np.stack([np.stack([v00, v10, v11], -1), np.stack([v00, v11, v01], -1)], 1).reshape(-1, 3)
    return verts.astype(np.float32), tri.astype(np.uint32).ravel()


def look_at_lh(eye, target, up):
    eye, target, up = (np.asarray(a, np.float64) for a in (eye, target, up))
    za = target - eye
    za /= np.linalg.norm(za)
    xa = np.cross(up, za)
    xa /= np.linalg.norm(xa)
    ya = np.cross(za, xa)
    m = np.eye(4)
    m[0:3, 0] = xa
    m[0:3, 1] = ya
    m[0:3, 2] = za
    m[3, 0:3] = [-xa @ eye, -ya @ eye, -za @ eye]
    return m


def perspective_fov_lh(fovy, aspect, zn, zf):
    ys = 1.0 / math.tan(fovy / 2)
    xs = ys / aspect
    m = np.zeros((4, 4))
    m[0, 0] = xs
    m[1, 1] = ys
    m[2, 2] = zf / (zf - zn)
    m[2, 3] = 1.0
    m[3, 2] = -zn * zf / (zf - zn)
    return m


# Cameras of the textured-mesh scene (eye, target).  Both keep the grid clear of the near plane
# (clip w >= 1.40 and >= 0.43), so no triangle takes the clipper's slow path.
MESH_CAMERAS = {
    "oblique": ((0.0, 1.5, -1.5), (0.0, 0.0, 0.0)),  # the configuration's camera: ~11% of the frame
    "close": ((0.0, 0.8, -0.6), (0.0, 0.5, 0.0)),    # ~80% of the frame
}
MESH_FOVY = math.pi / 2


def mesh_wvp(width: int, height: int, camera: str = "oblique") -> np.ndarray:
    eye, target = MESH_CAMERAS[camera]
    v = look_at_lh(eye, target, (0.0, 1.0, 0.0))
    p = perspective_fov_lh(MESH_FOVY, width / height, 0.1, 100.0)
    return (v @ p).astype(np.float32)  # row-vector convention: clip = [x y z 1] @ wvp


def noise_texture(size: int, seed: int = SEED_TEXTURE) -> np.ndarray:
    """(size, size, 4) uint8 uniform noise."""
    u = splitmix64(seed, size * size) >> np.uint64(32)
    return u.astype(np.uint32).view(np.uint8).reshape(size, size, 4).copy()


LIGHT_DIR = (np.array([1.0, 1.0, -1.0]) / math.sqrt(3.0)).astype(np.float32)


# ---- scene objects bound to one renderer -------------------------------------------------
@dataclass
class Frame:
    renderer: A.Renderer
    color: A.Surface
    ds: A.Surface
    resolved: A.Surface | None
    draw: callable
    n_tris: int
    width: int
    height: int
    samples: int
    keep: list

    def render(self, clear=True, resolve=True):
        r = self.renderer
        if clear:
            r.clear_color(self.color, (0.0, 0.0, 0.0, 0.0))
            r.clear_depth_stencil(self.ds, A.CLEAR_DEPTH | A.CLEAR_STENCIL, 1.0, 0)
        self.draw()
        if resolve and self.resolved is not None:
            r.resolve(self.color, self.resolved)


def _targets(r: A.Renderer, w, h, samples):
    ct = r.create_tex2d(w, h, samples, A.PF_RGBA8)
    dt = r.create_tex2d(w, h, samples, A.PF_RG32F)
    color, ds = ct.surface(0), dt.surface(0)
    resolved = None
    keep = [ct, dt]
    if samples > 1:
        rt = r.create_tex2d(w, h, 1, A.PF_RGBA8)
        resolved = rt.surface(0)
        keep.append(rt)
    r.set_render_targets([color], ds)
    r.set_viewport(0, 0, w, h, 0.0, 1.0)
    return color, ds, resolved, keep


def setup_plain(r: A.Renderer, width=512, height=512, n_tris=2000, samples=1, seed=SEED_PLAIN,
                cull=A.CULL_NONE) -> Frame:
    lib = r.lib
    color, ds, resolved, keep = _targets(r, width, height, samples)
    verts = plain_triangles(n_tris, seed)
    vb = r.create_buffer_from(verts)
    layout = r.create_input_layout([("POSITION", 0, A.FMT_R32G32B32A32_FLOAT, 0, 0),
                                    ("COLOR", 0, A.FMT_R32G32B32A32_FLOAT, 0, 16)], A.VS_PASSTHROUGH_COLOR)
    r.set_input_layout(layout)
    r.set_vertex_buffers(0, [vb], [32])
    r.set_primitive_topology(A.TOPOLOGY_TRIANGLE_LIST)
    r.set_vertex_shader(A.VS_PASSTHROUGH_COLOR)
    r.set_pixel_shader(A.PS_GOURAUD)
    r.set_blend_shader(A.BLEND_OFF)
    rd = A.default_raster_desc(lib)
    rd.cull_mode = cull
    rs = r.create_raster_state(rd)
    r.set_rasterizer_state(rs)
    dss = r.create_depth_stencil_state(A.default_depth_stencil_desc(lib))
    r.set_depth_stencil_state(dss, 0)
    keep += [vb, layout, rs, dss]
    return Frame(r, color, ds, resolved, lambda: r.draw(0, n_tris), n_tris, width, height, samples, keep)


def setup_mesh(r: A.Renderer, width=1920, height=1080, samples=4, grid=512, tex_size=2048,
               max_aniso=16, mip_filter=A.FILTER_ANISOTROPIC, camera="oblique", sasl=False) -> Frame:
    """The textured-mesh frame; sasl=True shades it with SASL_SHADERS["lambert"] instead of
    PS_LAMBERT_TEX (the same lighting with SASL's normalize, v / sqrt(dot(v, v)), in place of
    eflib's normalize3)."""
    lib = r.lib
    color, ds, resolved, keep = _targets(r, width, height, samples)
    verts, idx = grid_mesh(grid)
    vb = r.create_buffer_from(verts)
    ib = r.create_buffer_from(idx)
    layout = r.create_input_layout([("POSITION", 0, A.FMT_R32G32B32_FLOAT, 0, 0),
                                    ("NORMAL", 0, A.FMT_R32G32B32_FLOAT, 0, 12),
                                    ("TEXCOORD", 0, A.FMT_R32G32_FLOAT, 0, 24)], A.VS_WVP_NORMAL_UV)
    r.set_input_layout(layout)
    r.set_vertex_buffers(0, [vb], [32])
    r.set_index_buffer(ib, A.FMT_R32_UINT)
    r.set_primitive_topology(A.TOPOLOGY_TRIANGLE_LIST)
    r.set_vertex_shader(A.VS_WVP_NORMAL_UV, mesh_wvp(width, height, camera).ravel())
    r.set_pixel_shader(A.PS_LAMBERT_TEX, np.array([*LIGHT_DIR, 0.0], np.float32))
    tex = r.create_tex2d(tex_size, tex_size, 1, A.PF_RGBA8)
    tex.surface(0).write(noise_texture(tex_size))
    tex.gen_mipmap(A.FILTER_LINEAR, True)
    sd = A.default_sampler_desc(lib)
    sd.min_filter = A.FILTER_LINEAR
    sd.mag_filter = A.FILTER_LINEAR
    sd.mip_filter = mip_filter
    sd.max_anisotropy = max_aniso
    samp = r.create_sampler(sd, tex)
    r.set_ps_sampler(0, samp)
    if sasl:
        code = lib.compile_shader(SASL_SHADERS["lambert"])
        r.set_pixel_shader_code(code)
        r.set_ps_sampler_named("diffuse", samp)
        r.set_ps_variable("light_dir", [*LIGHT_DIR, 0.0])
        keep.append(code)
    r.set_blend_shader(A.BLEND_OFF)
    rd = A.default_raster_desc(lib)
    rd.cull_mode = A.CULL_NONE
    rs = r.create_raster_state(rd)
    r.set_rasterizer_state(rs)
    dss = r.create_depth_stencil_state(A.default_depth_stencil_desc(lib))
    r.set_depth_stencil_state(dss, 0)
    n_tris = idx.size // 3
    keep += [vb, ib, layout, tex, samp, rs, dss]
    return Frame(r, color, ds, resolved, lambda: r.draw_index(0, n_tris, 0), n_tris, width, height, samples, keep)


# ---- multipass ---------------------------------------------------------------------------
SEED_MULTIPASS = 0x5A11A003


def multipass_triangles(n_tris: int, width: int, height: int, seed: int = SEED_MULTIPASS):
    """Positions (3n, 3) in the camera's clip space (w = 1) and colours (3n, 4).

    Per triangle: centroid ~ U[-1,1]^2, area ~ log-uniform over [1, 4096] pixels of a width x height
    target, an isotropic random shape (three standard-normal offsets scaled to that area), depth
    ~ U[0.1, 0.9] per vertex, RGB ~ U[0,1]^3; 30% of the triangles are translucent with alpha ~
    U[0.2, 0.8] (the rest have alpha 1).  Returns (pos, col, translucent mask per triangle)."""
    u = uniform01(seed, n_tris * 17).reshape(n_tris, 17)
    c = -1.0 + 2.0 * u[:, 0:2]
    area_px = np.exp(np.log(4096.0) * u[:, 2])
    # three isotropic offsets from two uniforms each (Box-Muller)
    r = np.sqrt(-2.0 * np.log(np.maximum(u[:, 3:6], 1e-12)))
    th = 2.0 * math.pi * u[:, 6:9]
    off = np.stack([r * np.cos(th), r * np.sin(th)], -1)  # (n, 3, 2)
    off -= off.mean(1, keepdims=True)
    a0 = 0.5 * np.abs((off[:, 1, 0] - off[:, 0, 0]) * (off[:, 2, 1] - off[:, 0, 1])
                      - (off[:, 2, 0] - off[:, 0, 0]) * (off[:, 1, 1] - off[:, 0, 1]))
    scale = np.sqrt(area_px / np.maximum(a0, 1e-6))
    xy = c[:, None, :] + off * scale[:, None, None] * np.array([2.0 / width, 2.0 / height])
    z = 0.1 + 0.8 * u[:, 9:12]
    pos = np.concatenate([xy, z[..., None]], -1).reshape(-1, 3)
    translucent = u[:, 12] < 0.3
    alpha = np.where(translucent, 0.2 + 0.6 * u[:, 13], 1.0)
    rgba = np.concatenate([u[:, 14:17], alpha[:, None]], 1)
    col = np.repeat(rgba, 3, 0)
    return pos.astype(np.float32), col.astype(np.float32), translucent


def light_matrix() -> np.ndarray:
    """Row-vector orthographic light: rotate about y by 20 and x by 15 degrees around (0, 0, 0.5),
    scale x, y by 0.8 and map the rotated depth to [0.1, 0.9]."""
    ay, ax = math.radians(20.0), math.radians(15.0)
    t0 = np.eye(4); t0[3, 2] = -0.5
    ry = np.array([[math.cos(ay), 0, -math.sin(ay), 0], [0, 1, 0, 0], [math.sin(ay), 0, math.cos(ay), 0], [0, 0, 0, 1]])
    rx = np.array([[1, 0, 0, 0], [0, math.cos(ax), math.sin(ax), 0], [0, -math.sin(ax), math.cos(ax), 0], [0, 0, 0, 1]])
    sc = np.diag([0.8, 0.8, 0.4, 1.0]); sc[3, 2] = 0.5
    return (t0 @ ry @ rx @ sc).astype(np.float32)


def reflect_matrix(plane) -> np.ndarray:
    """eflib mat_reflect (eflib/src/math.cpp:429-456) of the plane (a, b, c, d), row-vector form."""
    n = np.asarray(plane[:3], np.float64)
    ln = np.linalg.norm(n)
    x, y, z = n / ln
    w = plane[3] / ln
    return np.array([[-2 * x * x + 1, -2 * y * x, -2 * z * x, 0], [-2 * x * y, -2 * y * y + 1, -2 * z * y, 0],
                     [-2 * x * z, -2 * y * z, -2 * z * z + 1, 0], [-2 * x * w, -2 * y * w, -2 * z * w, 1]], np.float32)


class MultipassFrame:
    """One frame: shadow pass over all triangles (depth only, from the light) -> opaque shaded pass with a 3x3 PCF
    shadow lookup -> translucent triangles back-to-front (blend on, no depth write) -> mirror quad at
    depth 0.95 marking stencil 1 (StencilMirror's stencil-replace state with depth ALWAYS, so the
    mirror's depth replaces the scene's; colour writes disabled) -> the opaque triangles reflected
    in the mirror plane x = 0.55 where stencil equals 1 -> MSAA resolve."""

    def __init__(self, r: A.Renderer, width=1920, height=1080, samples=4, shadow_size=2048, n_tris=200_000,
                 seed=SEED_MULTIPASS):
        lib = r.lib
        self.r, self.width, self.height, self.samples, self.n_tris = r, width, height, samples, n_tris
        self.color, self.ds, self.resolved, self.keep = _targets(r, width, height, samples)
        self.shadow_tex = r.create_tex2d(shadow_size, shadow_size, 1, A.PF_RG32F)
        self.shadow = self.shadow_tex.surface(0)
        pos, col, tr = multipass_triangles(n_tris, width, height, seed)
        tri_pos = pos.reshape(-1, 3, 3)
        tri_col = col.reshape(-1, 3, 4)
        opaque = ~tr
        # translucent triangles back-to-front: larger mean depth first (stable)
        tidx = np.nonzero(tr)[0]
        tidx = tidx[np.argsort(-tri_pos[tidx, :, 2].mean(1), kind="stable")]
        order = np.concatenate([np.nonzero(opaque)[0], tidx])
        verts = np.concatenate([tri_pos[order].reshape(-1, 3), tri_col[order].reshape(-1, 4)], 1).astype(np.float32)
        self.n_opaque, self.n_trans = int(opaque.sum()), int(tr.sum())
        self.vb = r.create_buffer_from(np.ascontiguousarray(verts))
        self.layout = r.create_input_layout([("POSITION", 0, A.FMT_R32G32B32_FLOAT, 0, 0),
                                             ("COLOR", 0, A.FMT_R32G32B32A32_FLOAT, 0, 12)], A.VS_WVP_COLOR_LIGHT)
        mirror = np.array([[0.2, -0.9, 0.95, 1, 1, 1, 1], [0.9, -0.9, 0.95, 1, 1, 1, 1], [0.9, -0.2, 0.95, 1, 1, 1, 1],
                           [0.2, -0.9, 0.95, 1, 1, 1, 1], [0.9, -0.2, 0.95, 1, 1, 1, 1], [0.2, -0.2, 0.95, 1, 1, 1, 1]],
                          np.float32)
        self.mvb = r.create_buffer_from(mirror)
        self.L = light_matrix()
        self.I = np.eye(4, dtype=np.float32)
        self.M = reflect_matrix((1.0, 0.0, 0.0, -0.55))
        sd = A.default_sampler_desc(lib)
        sd.min_filter = sd.mag_filter = sd.mip_filter = A.FILTER_POINT
        sd.addr_mode_u = sd.addr_mode_v = sd.addr_mode_w = A.ADDRESS_BORDER
        for k in range(4):
            sd.border_color[k] = 1.0
        self.samp = r.create_sampler(sd, self.shadow_tex)
        rd = A.default_raster_desc(lib)
        rd.cull_mode = A.CULL_NONE
        self.rs = r.create_raster_state(rd)
        d = A.default_depth_stencil_desc(lib)
        self.dss_normal = r.create_depth_stencil_state(d)
        dt = A.default_depth_stencil_desc(lib)
        dt.depth_write_mask = 0
        self.dss_translucent = r.create_depth_stencil_state(dt)
        dw = A.default_depth_stencil_desc(lib)  # StencilMirror.cpp:197-201, depth reset to the mirror plane
        dw.stencil_enable = 1
        dw.depth_func = A.CMP_ALWAYS
        for f in (dw.front_face, dw.back_face):
            f.stencil_pass_op = A.SOP_REPLACE
        self.dss_mark = r.create_depth_stencil_state(dw)
        de = A.default_depth_stencil_desc(lib)  # StencilMirror.cpp:203-206
        de.stencil_enable = 1
        for f in (de.front_face, de.back_face):
            f.stencil_func = A.CMP_EQUAL
        self.dss_test = r.create_depth_stencil_state(de)
        self.pcf = np.array([0.3, 0.002, 1.0 / shadow_size, 0.0], np.float32)
        self.shadow_size = shadow_size

    def _bind_geometry(self, vb, wvp, light, first_tri=0):
        r = self.r
        r.set_input_layout(self.layout)
        # a non-indexed draw ignores its start (index_fetcher.cpp:58-62,103-105), so a sub-range of
        # the vertex buffer is selected by the binding's byte offset
        r.set_vertex_buffers(0, [vb], [28], [first_tri * 3 * 28])
        r.set_primitive_topology(A.TOPOLOGY_TRIANGLE_LIST)
        r.set_vertex_shader(A.VS_WVP_COLOR_LIGHT, np.concatenate([wvp.ravel(), light.ravel()]).astype(np.float32))
        r.set_rasterizer_state(self.rs)

    def render(self, clear=True, resolve=True):
        r = self.r
        # (a) shadow map, depth only, from the light
        r.set_render_targets([], self.shadow)
        r.set_viewport(0, 0, self.shadow_size, self.shadow_size, 0.0, 1.0)
        r.clear_depth_stencil(self.shadow, A.CLEAR_DEPTH | A.CLEAR_STENCIL, 1.0, 0)
        self._bind_geometry(self.vb, self.L, self.L)
        r.set_pixel_shader(A.PS_GOURAUD)
        r.set_blend_shader(A.BLEND_OFF)
        r.set_depth_stencil_state(self.dss_normal, 0)
        r.draw(0, self.n_tris)  # every triangle casts a shadow, translucent ones included
        # (b) shaded opaque, then translucent back-to-front
        r.set_render_targets([self.color], self.ds)
        r.set_viewport(0, 0, self.width, self.height, 0.0, 1.0)
        if clear:
            r.clear_color(self.color, (0.0, 0.0, 0.0, 1.0))
            r.clear_depth_stencil(self.ds, A.CLEAR_DEPTH | A.CLEAR_STENCIL, 1.0, 0)
        self._bind_geometry(self.vb, self.I, self.L)
        r.set_pixel_shader(A.PS_SHADOW_PCF, self.pcf)
        r.set_ps_sampler(0, self.samp)
        r.draw(0, self.n_opaque)
        self._bind_geometry(self.vb, self.I, self.L, first_tri=self.n_opaque)
        r.set_blend_shader(A.BLEND_ON)
        r.set_depth_stencil_state(self.dss_translucent, 0)
        r.draw(0, self.n_trans)
        # (c) mirror quad: stencil 1, colour writes disabled
        self._bind_geometry(self.mvb, self.I, self.L)
        r.set_pixel_shader(A.PS_CONSTANT, np.array([1, 1, 1, 1], np.float32))
        r.set_blend_shader(A.BLEND_DISABLED)
        r.set_depth_stencil_state(self.dss_mark, 1)
        r.draw(0, 2)
        # (d) the opaque triangles reflected in the mirror plane x = 0.55, where stencil == 1
        self._bind_geometry(self.vb, self.M, self.L)
        r.set_pixel_shader(A.PS_SHADOW_PCF, self.pcf)
        r.set_blend_shader(A.BLEND_OFF)
        r.set_depth_stencil_state(self.dss_test, 1)
        r.draw(0, self.n_opaque)
        if resolve and self.resolved is not None:
            r.resolve(self.color, self.resolved)


# ---- batched jobs -----------------------------------------------------------------------
# 256 independent jobs (seed 0x5A11A004 + job).  Per job, four draws from the job's splitmix64
# stream pick the resolution from BATCH_RESOLUTIONS, the triangle count log-uniform in
# [1,000, 1,000,000], the resident texture's size from BATCH_TEX_SIZES and the sample count from
# the caller's choices; the triangles follow (plain_triangles' recipe on the same stream).  Each
# triangle with n > 2,000 is shrunk about its centroid by sqrt(2,000 / n), so every job keeps the
# plain scene's depth complexity (~150 layers at 2,000 triangles) instead of growing it to ~76,000
# at 1,000,000.  Shaders and states are the plain scene's; the texture (uniform noise, full mip
# chain) only occupies memory, as the plain shaders sample nothing.
SEED_BATCH = 0x5A11A004
BATCH_RESOLUTIONS = ((256, 256), (512, 512), (1024, 1024), (1920, 1080))
BATCH_TEX_SIZES = (256, 512, 1024, 2048)
BATCH_REF_TRIS = 2000


@dataclass
class JobSpec:
    seed: int
    width: int
    height: int
    n_tris: int
    tex_size: int
    samples: int


def batch_job_specs(n_jobs=256, seed=SEED_BATCH, min_tris=1000, max_tris=1_000_000, samples=(1,)):
    specs = []
    lo, hi = math.log(min_tris), math.log(max_tris)
    for j in range(n_jobs):
        u = uniform01(seed + j, 4)
        w, h = BATCH_RESOLUTIONS[min(int(u[0] * 4), 3)]
        n = int(round(math.exp(lo + u[1] * (hi - lo))))
        tex = BATCH_TEX_SIZES[min(int(u[2] * 4), 3)]
        s = samples[min(int(u[3] * len(samples)), len(samples) - 1)]
        specs.append(JobSpec(seed + j, w, h, n, tex, s))
    return specs


def batch_triangles(spec: JobSpec) -> np.ndarray:
    """(3*n, 8) float32 x y z w r g b a: plain_triangles' recipe after the four spec draws, with
    x, y scaled about each triangle's centroid by min(1, sqrt(2,000 / n))."""
    n = spec.n_tris
    # floats in [0,1) with 24 bits: exact in float32
    u = ((splitmix64(spec.seed, 4 + n * 21)[4:] >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -24))
    u = u.reshape(n * 3, 7)
    v = np.empty((n * 3, 8), np.float32)
    xy = (np.float32(-1.0) + np.float32(2.0) * u[:, 0:2]).reshape(n, 3, 2)
    k = np.float32(min(1.0, math.sqrt(BATCH_REF_TRIS / n)))
    if k < 1.0:
        c = (xy[:, 0:1] + xy[:, 1:2] + xy[:, 2:3]) * np.float32(1.0 / 3.0)
        xy = c + (xy - c) * k
    v[:, 0:2] = xy.reshape(n * 3, 2)
    v[:, 2] = np.float32(0.1) + np.float32(0.8) * u[:, 2]
    v[:, 3] = 1.0
    v[:, 4:8] = u[:, 3:7]
    return v


class BatchedJobs:
    """The jobs' resources on one renderer; render_all() issues every job's commands (clears,
    draw, resolve) one by one, record() the same commands into a batch."""

    def __init__(self, r: A.Renderer, specs, with_textures=True):
        lib = r.lib
        self.r = r
        self.specs = specs
        self.jobs = []
        self.layout = r.create_input_layout([("POSITION", 0, A.FMT_R32G32B32A32_FLOAT, 0, 0),
                                             ("COLOR", 0, A.FMT_R32G32B32A32_FLOAT, 0, 16)], A.VS_PASSTHROUGH_COLOR)
        rd = A.default_raster_desc(lib)
        rd.cull_mode = A.CULL_NONE
        self.rs = r.create_raster_state(rd)
        self.dss = r.create_depth_stencil_state(A.default_depth_stencil_desc(lib))
        from concurrent.futures import ThreadPoolExecutor

        with ThreadPoolExecutor(8) as ex:  # numpy releases the GIL in the generators
            geo = ex.map(batch_triangles, specs)
            texels = ex.map(lambda sp: noise_texture(sp.tex_size, sp.seed) if with_textures else None, specs)
            data = list(zip(geo, texels))
        for sp, (tris, noise) in zip(specs, data):
            ct = r.create_tex2d(sp.width, sp.height, sp.samples, A.PF_RGBA8)
            dt = r.create_tex2d(sp.width, sp.height, sp.samples, A.PF_RG32F)
            rt = r.create_tex2d(sp.width, sp.height, 1, A.PF_RGBA8) if sp.samples > 1 else None
            vb = r.create_buffer_from(tris)
            tex = None
            if with_textures:
                tex = r.create_tex2d(sp.tex_size, sp.tex_size, 1, A.PF_RGBA8)
                tex.surface(0).write(noise)
                tex.gen_mipmap(A.FILTER_LINEAR, True)
            self.jobs.append(dict(ct=ct, dt=dt, rt=rt, color=ct.surface(0), ds=dt.surface(0),
                                  resolved=rt.surface(0) if rt else None, vb=vb, tex=tex))

    def bind_common(self):
        r = self.r
        r.set_input_layout(self.layout)
        r.set_primitive_topology(A.TOPOLOGY_TRIANGLE_LIST)
        r.set_vertex_shader(A.VS_PASSTHROUGH_COLOR)
        r.set_pixel_shader(A.PS_GOURAUD)
        r.set_blend_shader(A.BLEND_OFF)
        r.set_rasterizer_state(self.rs)
        r.set_depth_stencil_state(self.dss, 0)

    def issue(self, j):
        r, sp, jb = self.r, self.specs[j], self.jobs[j]
        r.set_render_targets([jb["color"]], jb["ds"])
        r.set_viewport(0, 0, sp.width, sp.height, 0.0, 1.0)
        r.set_vertex_buffers(0, [jb["vb"]], [32])
        r.clear_color(jb["color"], (0.0, 0.0, 0.0, 0.0))
        r.clear_depth_stencil(jb["ds"], A.CLEAR_DEPTH | A.CLEAR_STENCIL, 1.0, 0)
        r.draw(0, sp.n_tris)
        if jb["resolved"] is not None:
            r.resolve(jb["color"], jb["resolved"])

    def render_all(self):
        self.bind_common()
        for j in range(len(self.jobs)):
            self.issue(j)

    def record(self) -> A.Batch:
        self.r.batch_begin()
        try:
            self.render_all()
        except Exception:
            self.r.batch_end().release()
            raise
        return self.r.batch_end()

    def output(self, j):
        jb = self.jobs[j]
        return (jb["resolved"] or jb["color"]).read()

    def footprint(self, j) -> int:
        """Device bytes the job keeps resident: geometry, targets and texture chain."""
        sp = self.specs[j]
        px = sp.width * sp.height
        tex = sum(max(sp.tex_size >> l, 1) ** 2 for l in range(int(math.log2(sp.tex_size)) + 1)) * 4
        return sp.n_tris * 96 + px * sp.samples * 12 + (px * 4 if sp.samples > 1 else 0) + tex


# ---- the reference's sample shaders --------------------------------------------------------
# One scene per sample shader pair of the reference (ColorizedTriangle, ComplexMesh,
# StandardShadowMap, StencilMirror's cup, TextureAndBlending's box), drawn over the displaced grid
# of the textured-mesh scene (positions, normals, uv) seen by the close camera.  Lights, materials
# and textures are this module's choices; the shaders are the reference's.
SAMPLE_SHADERS = ("colorized", "mesh", "ssm", "cup", "box")


def _grid_layout(r, vs_id, with_uv=True):
    elems = [("POSITION", 0, A.FMT_R32G32B32_FLOAT, 0, 0), ("NORMAL", 0, A.FMT_R32G32B32_FLOAT, 0, 12)]
    if with_uv:
        elems.append(("TEXCOORD", 0, A.FMT_R32G32_FLOAT, 0, 24))
    return r.create_input_layout(elems, vs_id)


def _linear_sampler(r, tex, addr=A.ADDRESS_WRAP):
    sd = A.default_sampler_desc(r.lib)
    sd.min_filter = sd.mag_filter = sd.mip_filter = A.FILTER_LINEAR
    sd.addr_mode_u = sd.addr_mode_v = sd.addr_mode_w = addr
    return r.create_sampler(sd, tex)


class SampleShaderFrame:
    """One frame of a sample shader scene: clear, (ssm: shadow map pass,) draw, (MSAA: resolve)."""

    def __init__(self, r: A.Renderer, name, width=320, height=240, samples=1, fmt=A.PF_RGBA8, grid=48, tex_size=64):
        if name not in SAMPLE_SHADERS:
            raise ValueError(name)
        lib = r.lib
        self.r, self.name, self.width, self.height, self.samples = r, name, width, height, samples
        ct = r.create_tex2d(width, height, samples, fmt)
        dt = r.create_tex2d(width, height, samples, A.PF_RG32F)
        self.color, self.ds = ct.surface(0), dt.surface(0)
        self.keep = [ct, dt]
        self.resolved = None
        if samples > 1:
            rt = r.create_tex2d(width, height, 1, fmt)
            self.resolved = rt.surface(0)
            self.keep.append(rt)
        verts, idx = grid_mesh(grid)
        self.n_tris = idx.size // 3
        self.vb = r.create_buffer_from(verts)
        self.ib = r.create_buffer_from(idx)
        self.tex = r.create_tex2d(tex_size, tex_size, 1, A.PF_RGBA8)
        self.tex.surface(0).write(noise_texture(tex_size, SEED_TEXTURE + 7))
        self.tex.gen_mipmap(A.FILTER_LINEAR, True)
        rd = A.default_raster_desc(lib)
        rd.cull_mode = A.CULL_NONE
        self.rs = r.create_raster_state(rd)
        self.dss = r.create_depth_stencil_state(A.default_depth_stencil_desc(lib))
        eye, target = MESH_CAMERAS["close"]
        self.view = look_at_lh(eye, target, (0.0, 1.0, 0.0))
        self.proj = perspective_fov_lh(MESH_FOVY, width / height, 0.1, 100.0)
        self.wvp = (self.view @ self.proj).astype(np.float32)
        self.clear_rgba = (0.2, 0.2, 0.5, 1.0)  # StandardShadowMap.cpp:245
        self.pre = None
        f32 = np.float32
        if name == "colorized":  # ColorizedTriangle: three lights in front of the surface
            lights = np.array([[0.6, 0.6, -0.5, 1.0], [-0.7, 0.1, -0.4, 1.0], [0.0, -0.6, -0.6, 1.0]], f32)
            self.vs = (A.VS_COLORIZED, np.concatenate([self.wvp.ravel(), lights.ravel()]).astype(f32))
            self.ps = (A.PS_COLORIZED, None)
            self.layout = _grid_layout(r, A.VS_COLORIZED, with_uv=False)
            self.samplers = {}
            self.blend = A.BLEND_OFF
        elif name == "mesh":  # ComplexMesh: eye-space Blinn-Phong, light and eye in eye space
            consts = np.zeros(40, f32)
            consts[0:16] = self.view.astype(f32).ravel()
            consts[16:32] = self.proj.astype(f32).ravel()
            consts[32:35] = (2.0, 2.0, -1.0)
            consts[36:39] = (0.0, 0.0, 0.0)
            self.vs = (A.VS_MESH, consts)
            self.ps = (A.PS_MESH, None)
            self.layout = _grid_layout(r, A.VS_MESH, with_uv=False)
            self.samplers = {}
            self.blend = A.BLEND_OFF
        elif name == "cup":  # StencilMirror's cup: textured Phong, object-space eye and light
            consts = np.zeros(24, f32)
            consts[0:16] = self.wvp.ravel()
            consts[16:20] = (*MESH_CAMERAS["close"][0], 1.0)
            consts[20:24] = (1.5, 1.5, -1.5, 1.0)
            self.vs = (A.VS_CUP, consts)
            mat = np.array([0.1, 0.1, 0.1, 1.0, 0.7, 0.6, 0.5, 1.0, 0.4, 0.4, 0.4, 1.0], f32)
            self.ps = (A.PS_CUP, mat)
            self.layout = _grid_layout(r, A.VS_CUP)
            self.samplers = {0: _linear_sampler(r, self.tex)}
            self.blend = A.BLEND_OFF
        elif name == "box":  # TextureAndBlending's box: tex2d(attr1) at alpha 0.5 blended over the clear
            self.vs = (A.VS_BOX, self.wvp.ravel())
            self.ps = (A.PS_BOX, None)
            self.layout = r.create_input_layout([("POSITION", 0, A.FMT_R32G32B32_FLOAT, 0, 0),
                                                 ("TEXCOORD", 0, A.FMT_R32G32_FLOAT, 0, 24)], A.VS_BOX)
            self.samplers = {0: _linear_sampler(r, self.tex)}
            self.blend = A.BLEND_ON
        else:  # ssm: StandardShadowMap -- a depth-only pass from the light, then the ESM shader
            light_eye = np.array([1.2, 1.6, -1.4])
            lv = look_at_lh(light_eye, (0.0, 0.3, 0.0), (0.0, 1.0, 0.0))
            lp = perspective_fov_lh(math.pi / 2.5, 1.0, 0.1, 10.0)
            self.light_wvp = (lv @ lp).astype(f32)
            self.sm_size = 256
            smt = r.create_tex2d(self.sm_size, self.sm_size, 1, A.PF_RG32F)
            self.shadow = smt.surface(0)
            self.keep.append(smt)
            sd = A.default_sampler_desc(lib)  # StandardShadowMap.cpp:173-182
            sd.min_filter = sd.mag_filter = sd.mip_filter = A.FILTER_POINT
            sd.addr_mode_u = sd.addr_mode_v = sd.addr_mode_w = A.ADDRESS_BORDER
            for i, c in enumerate((1.0, 0.0, 0.0, 0.0)):
                sd.border_color[i] = c
            self.sm_sampler = r.create_sampler(sd, smt)
            consts = np.zeros(40, f32)
            consts[0:4] = (*MESH_CAMERAS["close"][0], 1.0)
            consts[4:20] = self.wvp.ravel()
            consts[20:24] = (*light_eye, 1.0)
            consts[24:40] = self.light_wvp.ravel()
            self.vs = (A.VS_SSM_DRAW, consts)
            mat = np.zeros(13, f32)  # StandardShadowMap.cpp:256-259
            mat[0:12] = (0.1, 0.1, 0.1, 0.1, 0.8, 0.8, 0.8, 0.1, 0.4, 0.4, 0.4, 0.1)
            mat[12] = np.array([32], np.int32).view(f32)[0]
            self.ps = (A.PS_SSM, mat)
            self.layout = _grid_layout(r, A.VS_SSM_DRAW)
            self.samplers = {0: self.sm_sampler, 1: _linear_sampler(r, self.tex, A.ADDRESS_CLAMP)}
            self.blend = A.BLEND_OFF
            self.pre = self._shadow_pass

    def _bind(self, vs, ps):
        r = self.r
        r.set_input_layout(self.layout)
        r.set_vertex_buffers(0, [self.vb], [32])
        r.set_index_buffer(self.ib, A.FMT_R32_UINT)
        r.set_primitive_topology(A.TOPOLOGY_TRIANGLE_LIST)
        r.set_vertex_shader(*vs)
        r.set_pixel_shader(*ps)
        r.set_rasterizer_state(self.rs)
        r.set_depth_stencil_state(self.dss, 0)

    def _shadow_pass(self):
        r = self.r
        r.set_render_targets([], self.shadow)
        r.set_viewport(0, 0, self.sm_size, self.sm_size, 0.0, 1.0)
        r.clear_depth_stencil(self.shadow, A.CLEAR_DEPTH | A.CLEAR_STENCIL, 1.0, 0)
        consts = np.array(self.vs[1], np.float32).copy()
        consts[4:20] = self.light_wvp.ravel()  # the draw shader seen from the light (GenSM.savs)
        self._bind((A.VS_SSM_DRAW, consts), (A.PS_CONSTANT, np.zeros(4, np.float32)))
        r.set_blend_shader(A.BLEND_OFF)
        r.draw_index(0, self.n_tris, 0)

    def render(self, clear=True, resolve=True):
        r = self.r
        if self.pre is not None:
            self.pre()
        r.set_render_targets([self.color], self.ds)
        r.set_viewport(0, 0, self.width, self.height, 0.0, 1.0)
        if clear:
            r.clear_color(self.color, self.clear_rgba)
            r.clear_depth_stencil(self.ds, A.CLEAR_DEPTH | A.CLEAR_STENCIL, 1.0, 0)
        self._bind(self.vs, self.ps)
        for slot in range(4):
            r.set_ps_sampler(slot, self.samplers.get(slot))
        r.set_blend_shader(self.blend)
        r.draw_index(0, self.n_tris, 0)
        if resolve and self.resolved is not None:
            r.resolve(self.color, self.resolved)


# ---------------------------------------------------------------- SASL pixel shaders
# Pixel shaders written in SASL for draws through set_pixel_shader_code, modelled on the
# reference's SASL sample shaders: TextureAndBlending/AnisotropicFilter's box and plane shaders
# (tex2D with a fixed alpha), Triangles' constant colour, VertexTextureFetch's grey from one
# attribute; plus a textured Lambert shader, a procedural shader with loops and library math, and
# a shader writing its derivatives.
SASL_SHADERS = {
    "box": """
sampler samp;
float4 box_main(float4 p : TEXCOORD(0), float2 uv : TEXCOORD(1)) : COLOR {
  float4 c = tex2D(samp, uv);
  c.a = 0.5f;
  return c;
}
""",
    "plane": """
sampler samp;
float4 plane_main(float2 uv : TEXCOORD1) : COLOR {
  float4 c = tex2D(samp, uv);
  c.w = 1.0f;
  return c;
}
""",
    "constant": """
float4 color;
float4 const_main() : COLOR { return color; }
""",
    "grey": """
float4 grey_main(float h : TEXCOORD0) : COLOR { return float4(h, h, h, 1.0f); }
""",
    "lambert": """
sampler diffuse;
float4 light_dir;
struct VOut { float3 n : TEXCOORD(0); float2 uv : TEXCOORD(1); };
float4 lambert_main(VOut v) : COLOR {
  float3 n = normalize(v.n);
  float k = saturate(dot(n, light_dir.xyz));
  float4 t = tex2D(diffuse, v.uv);
  return float4(t.rgb * k, 1.0);
}
""",
    "procedural": """
float4 procedural_main(float3 n : TEXCOORD(0), float2 uv : TEXCOORD(1)) : COLOR {
  float2 p = uv * 8.0;
  float checker = fmod(floor(p.x) + floor(p.y), 2.0);
  float x = frac(p.x);
  float acc = 0.0;
  int i = 0;
  while (i < 6) {
    acc = acc + pow(x, 2.0) * 0.15;
    x = x * 1.5;
    if (x > 1.0) { x = x - 1.0; }
    i = i + 1;
  }
  float e = exp(-length(n) * 0.5) + sin(p.y) * 0.1;
  float band = 0.0;
  do { band = band + 0.25; } while (band < frac(p.y));
  return float4(checker * 0.8 + 0.1, saturate(acc), e, band);
}
""",
    "derivatives": """
float4 dd_main(float3 n : TEXCOORD(0), float2 uv : TEXCOORD(1)) : COLOR {
  float2 dx = ddx(uv);
  float2 dy = ddy(uv);
  if (uv.x > 0.5) { dx = ddx(uv * uv); }
  return float4(dx * 64.0, dy * 64.0);
}
""",
    # derivatives of texture samples and a branch on one: every pixel of a quad must sample
    "tex_derivatives": """
sampler diffuse;
float4 td_main(float3 n : TEXCOORD(0), float2 uv : TEXCOORD(1)) : COLOR {
  float4 t = tex2D(diffuse, uv);
  float a = ddx(t.x) * 4.0 + 0.5;
  float b = ddy(t.y) * 4.0 + 0.5;
  if (t.z > 0.5) { a = 1.0 - a; }
  return float4(a, b, t.z, 1.0);
}
""",
}
# the vertex shader each one runs behind (its attributes are the pixel shader's inputs)
SASL_VS = {"box": "box", "plane": "box", "constant": "uv", "grey": "uv", "lambert": "uv", "procedural": "uv",
           "derivatives": "uv", "tex_derivatives": "uv"}


class SaslFrame:
    """One frame of the grid mesh drawn with a SASL pixel shader: clear, draw, (MSAA: resolve)."""

    def __init__(self, r: A.Renderer, name, width=320, height=240, samples=1, fmt=A.PF_RGBA8, grid=48, tex_size=64,
                 aniso=False, blend=A.BLEND_OFF):
        lib = r.lib
        self.r, self.name, self.width, self.height, self.samples = r, name, width, height, samples
        ct = r.create_tex2d(width, height, samples, fmt)
        dt = r.create_tex2d(width, height, samples, A.PF_RG32F)
        self.color, self.ds = ct.surface(0), dt.surface(0)
        self.keep = [ct, dt]
        self.resolved = None
        if samples > 1:
            rt = r.create_tex2d(width, height, 1, fmt)
            self.resolved = rt.surface(0)
            self.keep.append(rt)
        verts, idx = grid_mesh(grid)
        self.n_tris = idx.size // 3
        self.vb = r.create_buffer_from(verts)
        self.ib = r.create_buffer_from(idx)
        self.tex = r.create_tex2d(tex_size, tex_size, 1, A.PF_RGBA8)
        self.tex.surface(0).write(noise_texture(tex_size, SEED_TEXTURE + 11))
        self.tex.gen_mipmap(A.FILTER_LINEAR, True)
        sd = A.default_sampler_desc(lib)
        sd.min_filter = sd.mag_filter = sd.mip_filter = A.FILTER_LINEAR
        if aniso:
            sd.mip_filter = A.FILTER_ANISOTROPIC
            sd.max_anisotropy = 16
        sd.addr_mode_u = sd.addr_mode_v = sd.addr_mode_w = A.ADDRESS_WRAP
        self.sampler = r.create_sampler(sd, self.tex)
        rd = A.default_raster_desc(lib)
        rd.cull_mode = A.CULL_NONE
        self.rs = r.create_raster_state(rd)
        self.dss = r.create_depth_stencil_state(A.default_depth_stencil_desc(lib))
        eye, target = MESH_CAMERAS["close"]
        view = look_at_lh(eye, target, (0.0, 1.0, 0.0))
        proj = perspective_fov_lh(MESH_FOVY, width / height, 0.1, 100.0)
        self.wvp = (view @ proj).astype(np.float32)
        if SASL_VS[name] == "box":
            self.vs = (A.VS_BOX, self.wvp.ravel())
            self.layout = r.create_input_layout([("POSITION", 0, A.FMT_R32G32B32_FLOAT, 0, 0),
                                                 ("TEXCOORD", 0, A.FMT_R32G32_FLOAT, 0, 24)], A.VS_BOX)
        else:
            self.vs = (A.VS_WVP_NORMAL_UV, self.wvp.ravel())
            self.layout = _grid_layout(r, A.VS_WVP_NORMAL_UV)
        self.code = lib.compile_shader(SASL_SHADERS[name])
        self.blend = blend
        self.clear_rgba = (0.2, 0.2, 0.5, 1.0)

    def render(self, clear=True, resolve=True):
        r = self.r
        r.set_render_targets([self.color], self.ds)
        r.set_viewport(0, 0, self.width, self.height, 0.0, 1.0)
        if clear:
            r.clear_color(self.color, self.clear_rgba)
            r.clear_depth_stencil(self.ds, A.CLEAR_DEPTH | A.CLEAR_STENCIL, 1.0, 0)
        r.set_input_layout(self.layout)
        r.set_vertex_buffers(0, [self.vb], [32])
        r.set_index_buffer(self.ib, A.FMT_R32_UINT)
        r.set_primitive_topology(A.TOPOLOGY_TRIANGLE_LIST)
        r.set_vertex_shader(*self.vs)
        r.set_pixel_shader(A.PS_CONSTANT, np.array([1.0, 0.0, 1.0, 1.0], np.float32))  # overridden by the code
        r.set_pixel_shader_code(self.code)
        for name in ("samp", "diffuse"):
            r.set_ps_sampler_named(name, self.sampler)
        r.set_ps_variable("color", [0.25, 0.5, 0.75, 1.0])
        ld = np.array([0.3, 0.8, -0.5], np.float64)
        r.set_ps_variable("light_dir", [*(ld / np.linalg.norm(ld)), 0.0])
        r.set_rasterizer_state(self.rs)
        r.set_depth_stencil_state(self.dss, 0)
        r.set_blend_shader(self.blend)
        r.draw_index(0, self.n_tris, 0)
        if resolve and self.resolved is not None:
            r.resolve(self.color, self.resolved)
