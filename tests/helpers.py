"""Small scene builders shared by the tests (through the renderer API)."""
import numpy as np

from salviarenderer_amd import api as A


class Target:
    def __init__(self, r, w, h, samples=1, fmt=A.PF_RGBA8):
        self.r = r
        self.ct = r.create_tex2d(w, h, samples, fmt)
        self.dt = r.create_tex2d(w, h, samples, A.PF_RG32F)
        self.color = self.ct.surface(0)
        self.ds = self.dt.surface(0)
        self.w, self.h, self.samples = w, h, samples
        r.set_render_targets([self.color], self.ds)
        r.set_viewport(0, 0, w, h, 0.0, 1.0)

    def clear(self, rgba=(0, 0, 0, 0), depth=1.0, stencil=0):
        self.r.clear_color(self.color, rgba)
        self.r.clear_depth_stencil(self.ds, A.CLEAR_DEPTH | A.CLEAR_STENCIL, depth, stencil)


def draw_colored(r, verts_xyzw, colors, *, cull=A.CULL_NONE, depth_func=A.CMP_LESS, depth_enable=True,
                 depth_write=True, ps=A.PS_GOURAUD, ps_consts=None, blend=A.BLEND_OFF, stencil=None):
    """Non-indexed triangle list; verts_xyzw: (3n, 4) clip space, colors: (3n, 4)."""
    lib = r.lib
    v = np.concatenate([np.asarray(verts_xyzw, np.float32), np.asarray(colors, np.float32)], 1)
    vb = r.create_buffer_from(np.ascontiguousarray(v))
    layout = r.create_input_layout([("POSITION", 0, A.FMT_R32G32B32A32_FLOAT, 0, 0),
                                    ("COLOR", 0, A.FMT_R32G32B32A32_FLOAT, 0, 16)], A.VS_PASSTHROUGH_COLOR)
    r.set_input_layout(layout)
    r.set_vertex_buffers(0, [vb], [32])
    r.set_primitive_topology(A.TOPOLOGY_TRIANGLE_LIST)
    r.set_vertex_shader(A.VS_PASSTHROUGH_COLOR)
    r.set_pixel_shader(ps, ps_consts)
    r.set_blend_shader(blend)
    rd = A.default_raster_desc(lib)
    rd.cull_mode = cull
    rs = r.create_raster_state(rd)
    r.set_rasterizer_state(rs)
    dd = A.default_depth_stencil_desc(lib)
    dd.depth_func = depth_func
    dd.depth_enable = int(depth_enable)
    dd.depth_write_mask = int(depth_write)
    ref = 0
    if stencil is not None:
        dd.stencil_enable = 1
        for face in (dd.front_face, dd.back_face):
            face.stencil_func = stencil["func"]
            face.stencil_pass_op = stencil.get("pass_op", A.SOP_KEEP)
        dd.stencil_read_mask = stencil.get("read_mask", 0xFF)
        dd.stencil_write_mask = stencil.get("write_mask", 0xFF)
        ref = stencil.get("ref", 0)
    dss = r.create_depth_stencil_state(dd)
    r.set_depth_stencil_state(dss, ref)
    r.draw(0, len(v) // 3)
    r.flush()
    return [vb, layout, rs, dss]


def screen_to_clip(x, y, w, h):
    """Inverse of the viewport transform for a (0, 0, w, h) viewport and clip w = 1."""
    return (np.asarray(x, np.float64) - w / 2) / (w / 2), -((np.asarray(y, np.float64) - h / 2) / (h / 2))
