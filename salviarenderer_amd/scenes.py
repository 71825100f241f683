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
                 (0, 0.8, -0.6) looking at (0, 0.5, 0), vertical fov pi/2, near 0.1, far 100.
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
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & MASK64) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


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


# Camera of the textured-mesh scene: the grid covers ~79% of a 16:9 frame and stays clear of
# the near plane (clip w >= 0.43), so no triangle takes the clipper's slow path.
MESH_EYE = (0.0, 0.8, -0.6)
MESH_TARGET = (0.0, 0.5, 0.0)
MESH_FOVY = math.pi / 2


def mesh_wvp(width: int, height: int) -> np.ndarray:
    v = look_at_lh(MESH_EYE, MESH_TARGET, (0.0, 1.0, 0.0))
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
               max_aniso=16, mip_filter=A.FILTER_ANISOTROPIC) -> Frame:
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
    r.set_vertex_shader(A.VS_WVP_NORMAL_UV, mesh_wvp(width, height).ravel())
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
