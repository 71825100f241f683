"""ctypes binding of the srx C ABI (include/srx.h).

`Lib(path)` binds one shared library that exports the ABI; `Renderer` and the resource
classes mirror salvia::core::renderer (salvia/include/salvia/core/renderer.h:42-131) and
its resources, method for method, so a script written against the reference reads the
same here.  Every call checks its result code and raises `SrxError` on failure.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

# ---- enums (include/srx.h) -------------------------------------------------------------
OK, FAILED, OUT_OF_MEMORY, INVALID_PARAMETER, E_HIP, E_UNSUPPORTED = 0, 1, 2, 3, 16, 17
LANG_VERTEX_SHADER, LANG_PIXEL_SHADER = 2, 3
PF_RGBA32F, PF_RGB32F, PF_BGRA8, PF_RGBA8, PF_R32F, PF_RG32F, PF_R32I = range(7)
PF_SIZE = {PF_RGBA32F: 16, PF_RGB32F: 12, PF_BGRA8: 4, PF_RGBA8: 4, PF_R32F: 4, PF_RG32F: 8, PF_R32I: 4}
FMT_R32G32B32A32_FLOAT, FMT_R32G32B32_FLOAT, FMT_R32G32_FLOAT, FMT_R32_FLOAT = 2, 6, 16, 41
FMT_R32_UINT, FMT_R16_UINT = 42, 57
TOPOLOGY_TRIANGLE_LIST, TOPOLOGY_TRIANGLE_STRIP = 2, 4
CULL_NONE, CULL_FRONT, CULL_BACK = 0, 1, 2
ADDRESS_WRAP, ADDRESS_MIRROR, ADDRESS_CLAMP, ADDRESS_BORDER = 0, 1, 2, 3
FILTER_POINT, FILTER_LINEAR, FILTER_ANISOTROPIC = 0, 1, 2
MIP_LO, MIP_MI, MIP_HI = 0, 1, 2
CMP_NEVER, CMP_LESS, CMP_EQUAL, CMP_LESS_EQUAL, CMP_GREATER, CMP_NOT_EQUAL, CMP_GREATER_EQUAL, CMP_ALWAYS = range(8)
SOP_KEEP, SOP_ZERO, SOP_REPLACE, SOP_INCR_SAT, SOP_DECR_SAT, SOP_INVERT, SOP_INCR_WRAP, SOP_DECR_WRAP = range(1, 9)
CLEAR_DEPTH, CLEAR_STENCIL = 1, 2
QUERY_PIPELINE_STATISTICS, QUERY_INTERNAL_STATISTICS, QUERY_PIPELINE_PROFILES = 3, 5, 6
VS_PASSTHROUGH_COLOR, VS_WVP_NORMAL_UV, VS_WVP_COLOR, VS_WVP_COLOR_LIGHT = 1, 2, 3, 4
VS_COLORIZED, VS_MESH, VS_SSM_DRAW, VS_CUP, VS_BOX = 5, 6, 7, 8, 9
PS_GOURAUD, PS_LAMBERT_TEX, PS_TEX_CPP, PS_CONSTANT, PS_SHADOW_PCF = 1, 2, 3, 4, 5
PS_COLORIZED, PS_MESH, PS_SSM, PS_CUP, PS_BOX = 6, 7, 8, 9, 10
BLEND_OFF, BLEND_ON, BLEND_DISABLED = 0, 1, 2
RENDERER_ASYNC, RENDERER_BENCHMARK = 0, 1
ASYNC_ERROR, ASYNC_TIMEOUT, ASYNC_READY = 0, 1, 2
MAP_READ, MAP_WRITE, MAP_READ_WRITE, MAP_WRITE_DISCARD, MAP_WRITE_NO_OVERWRITE = 1, 2, 3, 4, 5


class SrxError(RuntimeError):
    def __init__(self, fn, code, detail):
        super().__init__(f"{fn} failed with code {code}: {detail}")
        self.code = code


# ---- descriptor structs ----------------------------------------------------------------
class Viewport(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("x", "y", "w", "h", "minz", "maxz")]


class RasterDesc(C.Structure):
    _fields_ = [("fill_mode", C.c_int32), ("cull_mode", C.c_int32), ("front_ccw", C.c_int32),
                ("depth_bias", C.c_int32), ("depth_bias_clamp", C.c_float),
                ("slope_scaled_depth_bias", C.c_float), ("depth_clip_enable", C.c_int32),
                ("scissor_enable", C.c_int32), ("multisample_enable", C.c_int32),
                ("anti_aliased_line_enable", C.c_int32)]


class DepthStencilOpDesc(C.Structure):
    _fields_ = [("stencil_fail_op", C.c_int32), ("stencil_depth_fail_op", C.c_int32),
                ("stencil_pass_op", C.c_int32), ("stencil_func", C.c_int32)]


class DepthStencilDesc(C.Structure):
    _fields_ = [("depth_enable", C.c_int32), ("depth_write_mask", C.c_int32), ("depth_func", C.c_int32),
                ("stencil_enable", C.c_int32), ("stencil_read_mask", C.c_uint32),
                ("stencil_write_mask", C.c_uint32), ("front_face", DepthStencilOpDesc),
                ("back_face", DepthStencilOpDesc)]


class SamplerDesc(C.Structure):
    _fields_ = [("min_filter", C.c_int32), ("mag_filter", C.c_int32), ("mip_filter", C.c_int32),
                ("mip_qual", C.c_int32), ("addr_mode_u", C.c_int32), ("addr_mode_v", C.c_int32),
                ("addr_mode_w", C.c_int32), ("mip_lod_bias", C.c_float), ("max_anisotropy", C.c_uint32),
                ("comparison_func", C.c_int32), ("border_color", C.c_float * 4), ("min_lod", C.c_float),
                ("max_lod", C.c_float)]


class InputElementDesc(C.Structure):
    _fields_ = [("semantic_name", C.c_char_p), ("semantic_index", C.c_uint32), ("data_format", C.c_uint32),
                ("input_slot", C.c_uint32), ("aligned_byte_offset", C.c_uint32), ("slot_class", C.c_uint32),
                ("instance_data_step_rate", C.c_uint32)]


class MappedResource(C.Structure):
    _fields_ = [("data", C.c_void_p), ("row_pitch", C.c_uint32), ("depth_pitch", C.c_uint32)]


class PipelineStatistics(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "ia_vertices", "ia_primitives", "vs_invocations", "gs_invocations", "gs_primitives",
        "cinvocations", "cprimitives", "ps_invocations", "hs_invocations", "ds_invocations", "cs_invocations")]


class InternalStatistics(C.Structure):
    _fields_ = [("backend_input_pixels", C.c_uint64)]


class BatchInfo(C.Structure):
    _fields_ = [("jobs", C.c_uint32), ("clears", C.c_uint32), ("resolves", C.c_uint32), ("tile", C.c_uint32),
                ("bin_items", C.c_uint32), ("vis_threads", C.c_uint32), ("vis_launches", C.c_uint32),
                ("pad", C.c_uint32), ("primitives", C.c_uint64), ("records", C.c_uint64), ("device_bytes", C.c_uint64)]


class BatchTuning(C.Structure):
    _fields_ = [("tile", C.c_int32), ("bin_items", C.c_int32), ("vis_threads", C.c_int32), ("ms", C.c_float)]


class ServerDesc(C.Structure):
    _fields_ = [("pool_bytes", C.c_uint64), ("arena_bytes", C.c_uint64), ("staging_bytes", C.c_uint64),
                ("max_batch_jobs", C.c_uint32), ("streams", C.c_uint32), ("bin_items", C.c_int32),
                ("vis_threads", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("mesh_key", C.c_uint64), ("mesh_data", C.c_void_p), ("n_tris", C.c_uint32),
                ("texture_size", C.c_uint32), ("texture_key", C.c_uint64), ("texture_data", C.c_void_p),
                ("width", C.c_uint32), ("height", C.c_uint32), ("samples", C.c_uint32), ("want_depth", C.c_uint32),
                ("clear_color", C.c_float * 4), ("arrival_ns", C.c_int64)]


class Completion(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("status", C.c_int32), ("batch_jobs", C.c_uint32),
                ("arrival_ns", C.c_int64), ("admitted_ns", C.c_int64), ("done_ns", C.c_int64),
                ("color", C.c_void_p), ("color_bytes", C.c_uint64), ("depth", C.c_void_p), ("depth_bytes", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("batches", "jobs", "retries", "uploads", "evictions", "max_batch_jobs",
                                          "pool_bytes", "resident_meshes", "resident_textures", "arena_peak_bytes",
                                          "queued", "running")]


class ShaderInfo(C.Structure):
    _fields_ = [("input_count", C.c_uint32), ("output_count", C.c_uint32), ("input_components", C.c_uint32 * 8),
                ("output_targets", C.c_uint32 * 8), ("output_components", C.c_uint32 * 8),
                ("uniform_count", C.c_uint32), ("sampler_count", C.c_uint32), ("instruction_count", C.c_uint32),
                ("register_count", C.c_uint32)]


_H = C.c_void_p
_SZ = C.c_size_t
_I32 = C.c_int32
_U32 = C.c_uint32
_P = C.c_void_p

# name -> (restype, argtypes); every function of include/srx.h
PROTOTYPES = {
    "srx_abi_version": (_U32, []),
    "srx_backend_name": (C.c_char_p, []),
    "srx_last_error": (_I32, []),
    "srx_last_error_string": (C.c_char_p, []),
    "srx_retain": (None, [_H]),
    "srx_release": (None, [_H]),
    "srx_create_renderer": (_I32, [_U32, C.POINTER(_H)]),
    "srx_create_buffer": (_I32, [_H, _SZ, C.POINTER(_H)]),
    "srx_create_tex2d": (_I32, [_H, _SZ, _SZ, _SZ, _I32, C.POINTER(_H)]),
    "srx_create_sampler": (_I32, [_H, C.POINTER(SamplerDesc), _H, C.POINTER(_H)]),
    "srx_create_query": (_I32, [_H, _U32, C.POINTER(_H)]),
    "srx_create_input_layout": (_I32, [_H, C.POINTER(InputElementDesc), _SZ, _I32, C.POINTER(_H)]),
    "srx_create_raster_state": (_I32, [_H, C.POINTER(RasterDesc), C.POINTER(_H)]),
    "srx_create_depth_stencil_state": (_I32, [_H, C.POINTER(DepthStencilDesc), C.POINTER(_H)]),
    "srx_default_raster_desc": (None, [C.POINTER(RasterDesc)]),
    "srx_default_depth_stencil_desc": (None, [C.POINTER(DepthStencilDesc)]),
    "srx_default_sampler_desc": (None, [C.POINTER(SamplerDesc)]),
    "srx_buffer_size": (_SZ, [_H]),
    "srx_buffer_write": (_I32, [_H, _SZ, _P, _SZ]),
    "srx_buffer_read": (_I32, [_H, _SZ, _P, _SZ]),
    "srx_buffer_transfer": (_I32, [_H, _SZ, _P, _SZ, _SZ, _SZ, _SZ]),
    "srx_texture_surface": (_I32, [_H, _SZ, C.POINTER(_H)]),
    "srx_texture_gen_mipmap": (_I32, [_H, _I32, _I32]),
    "srx_texture_max_lod": (_SZ, [_H]),
    "srx_texture_min_lod": (_SZ, [_H]),
    "srx_surface_width": (_SZ, [_H]),
    "srx_surface_height": (_SZ, [_H]),
    "srx_surface_sample_count": (_SZ, [_H]),
    "srx_surface_format": (_I32, [_H]),
    "srx_surface_size_bytes": (_SZ, [_H]),
    "srx_surface_write": (_I32, [_H, _P, _SZ]),
    "srx_surface_read": (_I32, [_H, _P, _SZ]),
    "srx_surface_resolve": (_I32, [_H, _H, _H]),
    "srx_set_vertex_buffers": (_I32, [_H, _SZ, _SZ, C.POINTER(_H), C.POINTER(_SZ), C.POINTER(_SZ)]),
    "srx_set_index_buffer": (_I32, [_H, _H, _U32]),
    "srx_set_input_layout": (_I32, [_H, _H]),
    "srx_set_primitive_topology": (_I32, [_H, _U32]),
    "srx_set_vertex_shader": (_I32, [_H, _I32, _P, _SZ]),
    "srx_set_pixel_shader": (_I32, [_H, _I32, _P, _SZ]),
    "srx_set_ps_sampler": (_I32, [_H, _U32, _H]),
    "srx_set_blend_shader": (_I32, [_H, _I32]),
    "srx_compile_shader": (_I32, [C.c_char_p, _SZ, _I32, C.POINTER(_H)]),
    "srx_shader_reflect": (_I32, [_H, C.POINTER(ShaderInfo)]),
    "srx_set_pixel_shader_code": (_I32, [_H, _H]),
    "srx_get_pixel_shader_code": (_I32, [_H, C.POINTER(_H)]),
    "srx_set_ps_variable": (_I32, [_H, C.c_char_p, _P, _SZ]),
    "srx_set_ps_sampler_named": (_I32, [_H, C.c_char_p, _H]),
    "srx_ps_execute_quads": (_I32, [_H, _H, _P, _P, _SZ]),
    "srx_create_swap_chain": (_I32, [_H, _SZ, _SZ, _SZ, _I32, C.POINTER(_H)]),
    "srx_swap_chain_surface": (_I32, [_H, C.POINTER(_H)]),
    "srx_swap_chain_present": (_I32, [_H]),
    "srx_swap_chain_presented": (_I32, [_H, C.POINTER(_H)]),
    "srx_set_rasterizer_state": (_I32, [_H, _H]),
    "srx_set_depth_stencil_state": (_I32, [_H, _H, _I32]),
    "srx_set_render_targets": (_I32, [_H, _SZ, C.POINTER(_H), _H]),
    "srx_set_viewport": (_I32, [_H, C.POINTER(Viewport)]),
    "srx_get_viewport": (_I32, [_H, C.POINTER(Viewport)]),
    "srx_set_vs_sampler": (_I32, [_H, _U32, _H]),
    "srx_get_index_buffer": (_I32, [_H, C.POINTER(_H)]),
    "srx_get_index_format": (_U32, [_H]),
    "srx_get_primitive_topology": (_U32, [_H]),
    "srx_get_vertex_shader": (_I32, [_H]),
    "srx_get_rasterizer_state": (_I32, [_H, C.POINTER(_H)]),
    "srx_get_pixel_shader": (_I32, [_H]),
    "srx_get_blend_shader": (_I32, [_H]),
    "srx_get_depth_stencil_state": (_I32, [_H, C.POINTER(_H), C.POINTER(_I32)]),
    "srx_raster_state_desc": (_I32, [_H, C.POINTER(RasterDesc)]),
    "srx_depth_stencil_state_desc": (_I32, [_H, C.POINTER(DepthStencilDesc)]),
    "srx_map_buffer": (_I32, [_H, C.POINTER(MappedResource), _H, _I32]),
    "srx_map_surface": (_I32, [_H, C.POINTER(MappedResource), _H, _I32]),
    "srx_unmap": (_I32, [_H]),
    "srx_begin": (_I32, [_H, _H]),
    "srx_end": (_I32, [_H, _H]),
    "srx_get_data": (_I32, [_H, _H, _P, _SZ, _I32]),
    "srx_draw": (_I32, [_H, _SZ, _SZ]),
    "srx_draw_index": (_I32, [_H, _SZ, _SZ, _I32]),
    "srx_draw_noalloc": (_I32, [_H, _SZ, _SZ]),
    "srx_clear_color": (_I32, [_H, _H, C.POINTER(C.c_float)]),
    "srx_clear_depth_stencil": (_I32, [_H, _H, _U32, C.c_float, _U32]),
    "srx_flush": (_I32, [_H]),
    "srx_batch_begin": (_I32, [_H]),
    "srx_batch_end": (_I32, [_H, C.POINTER(_H)]),
    "srx_batch_launch": (_I32, [_H]),
    "srx_batch_status": (_I32, [_H, C.POINTER(_U32)]),
    "srx_batch_configure": (_I32, [_H, _I32, _I32, _I32]),
    "srx_batch_autotune": (_I32, [_H, _U32, C.POINTER(BatchTuning), _SZ, C.POINTER(_SZ)]),
    "srx_batch_get_info": (_I32, [_H, C.POINTER(BatchInfo)]),
    "srx_default_server_desc": (None, [C.POINTER(ServerDesc)]),
    "srx_create_server": (_I32, [_H, C.POINTER(ServerDesc), C.POINTER(_H)]),
    "srx_server_put_mesh": (_I32, [_H, C.c_uint64, _P, _U32]),
    "srx_server_put_texture": (_I32, [_H, C.c_uint64, _P, _U32]),
    "srx_server_submit": (_I32, [_H, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "srx_server_poll": (_I32, [_H, C.POINTER(Completion), _SZ, C.POINTER(_SZ), _I32]),
    "srx_server_release": (_I32, [_H, C.c_uint64]),
    "srx_server_now": (C.c_int64, [_H]),
    "srx_server_get_stats": (_I32, [_H, C.POINTER(ServerStats)]),
}


class Lib:
    """One loaded library exporting the srx ABI."""

    def __init__(self, path: str):
        if not os.path.exists(path):
            raise FileNotFoundError(f"srx library not built: {path}")
        self.path = path
        self.dll = C.CDLL(path, mode=os.RTLD_LOCAL | os.RTLD_NOW)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(self.dll, name)  # AttributeError if a symbol is missing
            fn.restype = res
            fn.argtypes = args
        self.backend = self.dll.srx_backend_name().decode()

    def check(self, name, code):
        if code != OK:
            detail = self.dll.srx_last_error_string()
            raise SrxError(name, code, detail.decode() if detail else "")
        return code

    def call(self, name, *args):
        return self.check(name, getattr(self.dll, name)(*args))

    def renderer(self, kind=RENDERER_BENCHMARK) -> "Renderer":
        return Renderer(self, kind)

    def compile_shader(self, code: str, lang=LANG_PIXEL_SHADER) -> "ShaderCode":
        """compile(code, lang) (salvia/include/salvia/core/renderer.h:136-147)."""
        src = code.encode()
        h = _H()
        self.call("srx_compile_shader", src, len(src), lang, C.byref(h))
        return ShaderCode(self, h.value)


class _Obj:
    def __init__(self, lib: Lib, handle):
        self.lib = lib
        self.h = handle

    def release(self):
        if self.h:
            self.lib.dll.srx_release(self.h)
            self.h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class SwapChain(_Obj):
    """A swap chain without a window: back buffer, present() = flush + resolve."""

    def surface(self) -> "Surface":
        h = _H()
        self.lib.call("srx_swap_chain_surface", self.h, C.byref(h))
        return Surface(self.lib, h.value)

    def present(self):
        self.lib.call("srx_swap_chain_present", self.h)

    def presented(self) -> "Surface":
        h = _H()
        self.lib.call("srx_swap_chain_presented", self.h, C.byref(h))
        return Surface(self.lib, h.value)


class ShaderCode(_Obj):
    """A compiled SASL pixel shader (shader_object)."""

    def reflect(self) -> ShaderInfo:
        info = ShaderInfo()
        self.lib.call("srx_shader_reflect", self.h, C.byref(info))
        return info


class Buffer(_Obj):
    def write(self, data, offset=0):
        arr = np.ascontiguousarray(data)
        self.lib.call("srx_buffer_write", self.h, offset, arr.ctypes.data, arr.nbytes)

    def transfer(self, data, offset, stride_dest, stride_src, size, count):
        """buffer::transfer: count rows of size bytes, stride_src apart in data, stride_dest apart here."""
        arr = np.ascontiguousarray(data)
        self.lib.call("srx_buffer_transfer", self.h, offset, arr.ctypes.data, stride_dest, stride_src, size, count)

    def read(self, size=None, offset=0) -> bytes:
        n = self.lib.dll.srx_buffer_size(self.h) - offset if size is None else size
        out = np.empty(n, np.uint8)
        self.lib.call("srx_buffer_read", self.h, offset, out.ctypes.data, n)
        return out.tobytes()


class Surface(_Obj):
    @property
    def width(self):
        return self.lib.dll.srx_surface_width(self.h)

    @property
    def height(self):
        return self.lib.dll.srx_surface_height(self.h)

    @property
    def samples(self):
        return self.lib.dll.srx_surface_sample_count(self.h)

    @property
    def format(self):
        return self.lib.dll.srx_surface_format(self.h)

    def nbytes(self):
        return self.lib.dll.srx_surface_size_bytes(self.h)

    def read(self) -> np.ndarray:
        """Raw texels as uint8 array shaped (h, w, samples, elem_bytes)."""
        n = self.nbytes()
        out = np.empty(n, np.uint8)
        self.lib.call("srx_surface_read", self.h, out.ctypes.data, n)
        return out.reshape(self.height, self.width, self.samples, PF_SIZE[self.format])

    def write(self, data):
        arr = np.ascontiguousarray(data)
        if arr.nbytes != self.nbytes():
            raise ValueError("surface write size mismatch")
        self.lib.call("srx_surface_write", self.h, arr.ctypes.data, arr.nbytes)


class Texture(_Obj):
    def surface(self, level=0) -> Surface:
        h = C.c_void_p()
        self.lib.call("srx_texture_surface", self.h, level, C.byref(h))
        return Surface(self.lib, h)

    def gen_mipmap(self, filt=FILTER_LINEAR, auto_gen=True):
        self.lib.call("srx_texture_gen_mipmap", self.h, filt, int(auto_gen))

    @property
    def max_lod(self):
        return self.lib.dll.srx_texture_max_lod(self.h)

    @property
    def min_lod(self):
        return self.lib.dll.srx_texture_min_lod(self.h)


class Query(_Obj):
    def __init__(self, lib, handle, qid):
        super().__init__(lib, handle)
        self.id = qid


class Server(_Obj):
    """Request server: continuous batching of plain-pipeline render requests over resident meshes."""

    def put_mesh(self, key, verts):
        v = np.ascontiguousarray(verts, np.float32)
        self.lib.call("srx_server_put_mesh", self.h, key, v.ctypes.data, v.size // 24)

    def put_texture(self, key, rgba8):
        t = np.ascontiguousarray(rgba8, np.uint8)
        self.lib.call("srx_server_put_texture", self.h, key, t.ctypes.data, t.shape[0])

    def submit(self, mesh_key, n_tris, width, height, samples=1, *, mesh=None, texture_key=0, texture=None,
               clear=(0.0, 0.0, 0.0, 0.0), want_depth=False, arrival_ns=None) -> int:
        q = Request()
        q.mesh_key = mesh_key
        keep = []
        if mesh is not None:
            m = np.ascontiguousarray(mesh, np.float32)
            keep.append(m)
            q.mesh_data = m.ctypes.data
        q.n_tris = n_tris
        q.texture_key = texture_key
        if texture is not None:
            t = np.ascontiguousarray(texture, np.uint8)
            keep.append(t)
            q.texture_data = t.ctypes.data
            q.texture_size = t.shape[0]
        q.width, q.height, q.samples, q.want_depth = width, height, samples, int(want_depth)
        for i in range(4):
            q.clear_color[i] = clear[i]
        q.arrival_ns = self.now() if arrival_ns is None else arrival_ns
        tk = C.c_uint64()
        self.lib.call("srx_server_submit", self.h, C.byref(q), C.byref(tk))
        return tk.value

    def poll(self, capacity=256, wait_ms=0, copy=True):
        """Finished requests as dicts; with copy, their images as numpy arrays (then released)."""
        arr = (Completion * capacity)()
        n = C.c_size_t()
        self.lib.call("srx_server_poll", self.h, arr, capacity, C.byref(n), wait_ms)
        out = []
        for c in arr[: n.value]:
            d = {"ticket": c.ticket, "status": c.status, "batch_jobs": c.batch_jobs, "arrival_ns": c.arrival_ns,
                 "admitted_ns": c.admitted_ns, "done_ns": c.done_ns}
            if copy:
                if c.color:
                    d["color"] = np.frombuffer(C.string_at(c.color, c.color_bytes), np.uint8).copy()
                if c.depth:
                    d["depth"] = np.frombuffer(C.string_at(c.depth, c.depth_bytes), np.uint8).copy()
                self.release(c.ticket)
            out.append(d)
        return out

    def release(self, ticket):
        self.lib.call("srx_server_release", self.h, ticket)

    def now(self) -> int:
        return int(self.lib.dll.srx_server_now(self.h))

    def stats(self) -> ServerStats:
        st = ServerStats()
        self.lib.call("srx_server_get_stats", self.h, C.byref(st))
        return st


def default_server_desc(lib: Lib) -> ServerDesc:
    d = ServerDesc()
    lib.dll.srx_default_server_desc(C.byref(d))
    return d


class Batch(_Obj):
    """Recorded independent render jobs replayed as one launch per pipeline stage."""

    def launch(self):
        self.lib.call("srx_batch_launch", self.h)

    def status(self) -> int:
        st = C.c_uint32()
        self.lib.call("srx_batch_status", self.h, C.byref(st))
        return st.value

    def configure(self, tile, bin_items, vis_threads):
        self.lib.call("srx_batch_configure", self.h, tile, bin_items, vis_threads)

    def autotune(self, reps=3):
        """Times every (tile, bin_items, vis_threads) the jobs allow; the batch keeps the fastest."""
        res = (BatchTuning * 64)()
        n = C.c_size_t()
        self.lib.call("srx_batch_autotune", self.h, reps, res, 64, C.byref(n))
        return [(r.tile, r.bin_items, r.vis_threads, r.ms) for r in res[: min(n.value, 64)]]

    def info(self) -> BatchInfo:
        d = BatchInfo()
        self.lib.call("srx_batch_get_info", self.h, C.byref(d))
        return d


def default_raster_desc(lib: Lib) -> RasterDesc:
    d = RasterDesc()
    lib.dll.srx_default_raster_desc(C.byref(d))
    return d


def default_depth_stencil_desc(lib: Lib) -> DepthStencilDesc:
    d = DepthStencilDesc()
    lib.dll.srx_default_depth_stencil_desc(C.byref(d))
    return d


def default_sampler_desc(lib: Lib) -> SamplerDesc:
    d = SamplerDesc()
    lib.dll.srx_default_sampler_desc(C.byref(d))
    return d


class Renderer(_Obj):
    """salvia::core::renderer over the C ABI."""

    def __init__(self, lib: Lib, kind=RENDERER_BENCHMARK):
        h = C.c_void_p()
        lib.call("srx_create_renderer", kind, C.byref(h))
        super().__init__(lib, h)
        self._keep = {}

    # creators
    def create_buffer(self, size) -> Buffer:
        h = C.c_void_p()
        self.lib.call("srx_create_buffer", self.h, size, C.byref(h))
        return Buffer(self.lib, h)

    def create_buffer_from(self, arr) -> Buffer:
        arr = np.ascontiguousarray(arr)
        b = self.create_buffer(arr.nbytes)
        b.write(arr)
        return b

    def create_tex2d(self, w, h, samples, fmt) -> Texture:
        hh = C.c_void_p()
        self.lib.call("srx_create_tex2d", self.h, w, h, samples, fmt, C.byref(hh))
        return Texture(self.lib, hh)

    def create_sampler(self, desc: SamplerDesc, tex: Texture):
        h = C.c_void_p()
        self.lib.call("srx_create_sampler", self.h, C.byref(desc), tex.h, C.byref(h))
        return _Obj(self.lib, h)

    def create_query(self, qid) -> Query:
        h = C.c_void_p()
        self.lib.call("srx_create_query", self.h, qid, C.byref(h))
        return Query(self.lib, h, qid)

    def create_input_layout(self, elems, vs_id):
        arr = (InputElementDesc * len(elems))()
        names = []
        for i, (sem, idx, fmt, slot, off) in enumerate(elems):
            names.append(sem.encode())
            arr[i] = InputElementDesc(names[-1], idx, fmt, slot, off, 0, 0)
        h = C.c_void_p()
        self.lib.call("srx_create_input_layout", self.h, arr, len(elems), vs_id, C.byref(h))
        return _Obj(self.lib, h)

    def create_raster_state(self, desc: RasterDesc):
        h = C.c_void_p()
        self.lib.call("srx_create_raster_state", self.h, C.byref(desc), C.byref(h))
        return _Obj(self.lib, h)

    def create_depth_stencil_state(self, desc: DepthStencilDesc):
        h = C.c_void_p()
        self.lib.call("srx_create_depth_stencil_state", self.h, C.byref(desc), C.byref(h))
        return _Obj(self.lib, h)

    # state
    def set_vertex_buffers(self, start, buffers, strides, offsets=None):
        n = len(buffers)
        hb = (C.c_void_p * n)(*[b.h for b in buffers])
        st = (C.c_size_t * n)(*strides)
        of = (C.c_size_t * n)(*(offsets or [0] * n))
        self._keep["vb"] = list(buffers)
        self.lib.call("srx_set_vertex_buffers", self.h, start, n, hb, st, of)

    def set_index_buffer(self, buf: Buffer, fmt):
        self._keep["ib"] = buf
        self.lib.call("srx_set_index_buffer", self.h, buf.h, fmt)

    def set_input_layout(self, layout):
        self._keep["layout"] = layout
        self.lib.call("srx_set_input_layout", self.h, layout.h)

    def set_primitive_topology(self, topo):
        self.lib.call("srx_set_primitive_topology", self.h, topo)

    def set_vertex_shader(self, vs_id, consts=None):
        c = np.ascontiguousarray(consts if consts is not None else np.zeros(0), np.float32)
        self.lib.call("srx_set_vertex_shader", self.h, vs_id, c.ctypes.data if c.size else None, c.nbytes)

    def set_pixel_shader(self, ps_id, consts=None):
        c = np.ascontiguousarray(consts if consts is not None else np.zeros(0), np.float32)
        self.lib.call("srx_set_pixel_shader", self.h, ps_id, c.ctypes.data if c.size else None, c.nbytes)

    def set_ps_sampler(self, slot, samp):
        self._keep[f"samp{slot}"] = samp
        self.lib.call("srx_set_ps_sampler", self.h, slot, samp.h if samp is not None else None)

    def create_swap_chain(self, w, h, samples=1, fmt=PF_RGBA8) -> SwapChain:
        out = _H()
        self.lib.call("srx_create_swap_chain", self.h, w, h, samples, fmt, C.byref(out))
        return SwapChain(self.lib, out.value)

    def set_pixel_shader_code(self, code):
        self._keep["ps_code"] = code
        self.lib.call("srx_set_pixel_shader_code", self.h, code.h if code is not None else None)

    def get_pixel_shader_code(self):
        h = _H()
        self.lib.call("srx_get_pixel_shader_code", self.h, C.byref(h))
        return ShaderCode(self.lib, h.value) if h.value else None

    def set_ps_variable(self, name, value, dtype=np.float32):
        a = np.ascontiguousarray(np.atleast_1d(np.asarray(value, dtype=dtype)))
        self.lib.call("srx_set_ps_variable", self.h, name.encode(), a.ctypes.data, a.nbytes)

    def set_ps_sampler_named(self, name, samp):
        self._keep[f"psnamed:{name}"] = samp
        self.lib.call("srx_set_ps_sampler_named", self.h, name.encode(), samp.h if samp is not None else None)

    def ps_execute_quads(self, code, inputs):
        """Runs `code` on quads: inputs (nquads, 4, input_count, 4) float32 -> (nquads, 4, output_count, 4)."""
        info = code.reflect()
        x = np.ascontiguousarray(inputs, np.float32)
        nq = x.shape[0] if x.ndim else 0
        if x.size and x.shape[1:] != (4, info.input_count, 4):
            raise ValueError(f"inputs must be (nquads, 4, {info.input_count}, 4)")
        out = np.zeros((nq, 4, info.output_count, 4), np.float32)
        self.lib.call("srx_ps_execute_quads", self.h, code.h, x.ctypes.data if x.size else None,
                      out.ctypes.data if out.size else None, nq)
        return out

    def set_blend_shader(self, bid):
        self.lib.call("srx_set_blend_shader", self.h, bid)

    def set_rasterizer_state(self, rs):
        self._keep["rs"] = rs
        self.lib.call("srx_set_rasterizer_state", self.h, rs.h)

    def set_depth_stencil_state(self, dss, stencil_ref=0):
        self._keep["dss"] = dss
        self.lib.call("srx_set_depth_stencil_state", self.h, dss.h, stencil_ref)

    def set_render_targets(self, color_targets, ds_target):
        n = len(color_targets)
        arr = (C.c_void_p * max(n, 1))(*[c.h for c in color_targets])
        self._keep["rt"] = (list(color_targets), ds_target)
        self.lib.call("srx_set_render_targets", self.h, n, arr if n else None, ds_target.h if ds_target else None)

    def set_viewport(self, x, y, w, h, minz=0.0, maxz=1.0):
        vp = Viewport(x, y, w, h, minz, maxz)
        self.lib.call("srx_set_viewport", self.h, C.byref(vp))

    # operations
    def begin(self, q: Query):
        self.lib.call("srx_begin", self.h, q.h)

    def end(self, q: Query):
        self.lib.call("srx_end", self.h, q.h)

    def get_data(self, q: Query, do_not_wait=False):
        if q.id == QUERY_PIPELINE_STATISTICS:
            d = PipelineStatistics()
        elif q.id == QUERY_INTERNAL_STATISTICS:
            d = InternalStatistics()
        else:
            d = (C.c_uint64 * 10)()
        st = self.lib.dll.srx_get_data(self.h, q.h, C.byref(d), C.sizeof(d), int(do_not_wait))
        return st, d

    # map / unmap (renderer.h:61-63)
    def map(self, res, mode) -> MappedResource:
        """Maps a Buffer or Surface; returns the mapped resource (data, row_pitch, depth_pitch)."""
        m = MappedResource()
        fn = "srx_map_buffer" if isinstance(res, Buffer) else "srx_map_surface"
        self.lib.call(fn, self.h, C.byref(m), res.h, mode)
        self._keep["mapped"] = res
        return m

    def unmap(self):
        self.lib.call("srx_unmap", self.h)
        self._keep.pop("mapped", None)

    def set_vs_sampler(self, slot, samp):
        self._keep[f"vs_sampler{slot}"] = samp
        self.lib.call("srx_set_vs_sampler", self.h, slot, samp.h if samp else None)

    # state get (renderer.h:104-113)
    def get_index_buffer(self):
        h = C.c_void_p()
        self.lib.call("srx_get_index_buffer", self.h, C.byref(h))
        return Buffer(self.lib, h) if h.value else None

    def get_index_format(self):
        return self.lib.dll.srx_get_index_format(self.h)

    def get_primitive_topology(self):
        return self.lib.dll.srx_get_primitive_topology(self.h)

    def get_vertex_shader(self):
        return self.lib.dll.srx_get_vertex_shader(self.h)

    def get_pixel_shader(self):
        return self.lib.dll.srx_get_pixel_shader(self.h)

    def get_blend_shader(self):
        return self.lib.dll.srx_get_blend_shader(self.h)

    def get_rasterizer_state(self) -> RasterDesc:
        h = C.c_void_p()
        self.lib.call("srx_get_rasterizer_state", self.h, C.byref(h))
        d = RasterDesc()
        try:
            self.lib.call("srx_raster_state_desc", h, C.byref(d))
        finally:
            self.lib.dll.srx_release(h)
        return d

    def get_depth_stencil_state(self):
        """(DepthStencilDesc, stencil_ref) of the bound state."""
        h, ref = C.c_void_p(), C.c_int32()
        self.lib.call("srx_get_depth_stencil_state", self.h, C.byref(h), C.byref(ref))
        d = DepthStencilDesc()
        try:
            self.lib.call("srx_depth_stencil_state_desc", h, C.byref(d))
        finally:
            self.lib.dll.srx_release(h)
        return d, ref.value

    def get_viewport(self) -> Viewport:
        vp = Viewport()
        self.lib.call("srx_get_viewport", self.h, C.byref(vp))
        return vp

    def draw(self, start, primcnt):
        self.lib.call("srx_draw", self.h, start, primcnt)

    def draw_noalloc(self, start, primcnt):
        """draw() with every scratch buffer in the library's fixed per-GPU region (no allocation)."""
        self.lib.call("srx_draw_noalloc", self.h, start, primcnt)

    def draw_index(self, start, primcnt, basevert=0):
        self.lib.call("srx_draw_index", self.h, start, primcnt, basevert)

    def clear_color(self, surf: Surface, rgba):
        c = (C.c_float * 4)(*rgba)
        self.lib.call("srx_clear_color", self.h, surf.h, c)

    def clear_depth_stencil(self, surf: Surface, flags, depth, stencil):
        self.lib.call("srx_clear_depth_stencil", self.h, surf.h, flags, depth, stencil)

    def resolve(self, src: Surface, dst: Surface):
        self.lib.call("srx_surface_resolve", self.h, src.h, dst.h)

    def flush(self):
        self.lib.call("srx_flush", self.h)

    # batches: commands between batch_begin and batch_end are recorded, not executed
    def batch_begin(self):
        self.lib.call("srx_batch_begin", self.h)

    def create_server(self, desc: ServerDesc = None) -> Server:
        d = desc if desc is not None else default_server_desc(self.lib)
        h = C.c_void_p()
        self.lib.call("srx_create_server", self.h, C.byref(d), C.byref(h))
        return Server(self.lib, h)

    def batch_end(self) -> Batch:
        h = C.c_void_p()
        self.lib.call("srx_batch_end", self.h, C.byref(h))
        return Batch(self.lib, h)
