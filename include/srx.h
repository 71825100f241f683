/*
 * srx.h -- C ABI of the Salvia rasterization pipeline on MI355X (gfx950).
 *
 * Every entry point replaces one method of the abstract C++ class
 * salvia::core::renderer (salvia/include/salvia/core/renderer.h:42-131) or of a
 * resource class it hands out, with the same argument order and meaning.
 * Objects are opaque, reference-counted handles (the reference uses
 * std::shared_ptr for every object, EFLIB_DECLARE_CLASS_SHARED_PTR).
 * Descriptor structs are plain C with the reference's field order.
 *
 * Two libraries export this ABI:
 *   libsrx_hip.so  -- the product: HIP kernels on one MI355X.
 *   libsrx_cpu.so  -- the CPU restatement used as test oracle (oracle/).
 *
 * Threading contract (renderer.h, async_renderer.cpp:17-110): one API thread per
 * renderer; commands are queued in submission order on the renderer's stream and
 * srx_flush() waits for all of them.
 */
#ifndef SRX_H_
#define SRX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SRX_ABI_VERSION 1

/* ---- result codes: numbering of salvia::result (salvia/common/constants.h:9) ---- */
enum srx_result {
  SRX_OK = 0,
  SRX_FAILED = 1,
  SRX_OUT_OF_MEMORY = 2,
  SRX_INVALID_PARAMETER = 3,
  SRX_E_HIP = 16,         /* device error; srx_last_error() carries the hipError_t */
  SRX_E_UNSUPPORTED = 17  /* valid in the reference but not implemented here */
};

/* async_status (constants.h:178) */
enum srx_async_status { SRX_ASYNC_ERROR = 0, SRX_ASYNC_TIMEOUT = 1, SRX_ASYNC_READY = 2 };

/* pixel_format codes (salvia/common/colors_convertors.h:40-47) */
enum srx_pixel_format {
  SRX_PF_RGBA32F = 0,
  SRX_PF_RGB32F = 1,
  SRX_PF_BGRA8 = 2,
  SRX_PF_RGBA8 = 3,
  SRX_PF_R32F = 4,
  SRX_PF_RG32F = 5, /* the only depth/stencil format: depth fp32 in .r, stencil u32 in .g */
  SRX_PF_R32I = 6
};

/* salvia::format subset used by vertex and index data (salvia/common/format.h) */
enum srx_format {
  SRX_FMT_R32G32B32A32_FLOAT = 2,
  SRX_FMT_R32G32B32A32_SINT = 4,
  SRX_FMT_R32G32B32_FLOAT = 6,
  SRX_FMT_R32G32_FLOAT = 16,
  SRX_FMT_R32_FLOAT = 41,
  SRX_FMT_R32_UINT = 42,
  SRX_FMT_R16_UINT = 57
};

/* primitive_topology (constants.h:48-60) */
enum srx_topology {
  SRX_TOPOLOGY_LINE_LIST = 0,
  SRX_TOPOLOGY_LINE_STRIP = 1,
  SRX_TOPOLOGY_TRIANGLE_LIST = 2,
  SRX_TOPOLOGY_TRIANGLE_FAN = 3,
  SRX_TOPOLOGY_TRIANGLE_STRIP = 4
};

enum srx_cull_mode { SRX_CULL_NONE = 0, SRX_CULL_FRONT = 1, SRX_CULL_BACK = 2 };
enum srx_fill_mode { SRX_FILL_WIREFRAME = 0, SRX_FILL_SOLID = 1 };
enum srx_address_mode { SRX_ADDRESS_WRAP = 0, SRX_ADDRESS_MIRROR = 1, SRX_ADDRESS_CLAMP = 2, SRX_ADDRESS_BORDER = 3 };
enum srx_filter { SRX_FILTER_POINT = 0, SRX_FILTER_LINEAR = 1, SRX_FILTER_ANISOTROPIC = 2 };
enum srx_mip_quality { SRX_MIP_LO = 0, SRX_MIP_MI = 1, SRX_MIP_HI = 2 };
enum srx_compare {
  SRX_CMP_NEVER = 0, SRX_CMP_LESS = 1, SRX_CMP_EQUAL = 2, SRX_CMP_LESS_EQUAL = 3,
  SRX_CMP_GREATER = 4, SRX_CMP_NOT_EQUAL = 5, SRX_CMP_GREATER_EQUAL = 6, SRX_CMP_ALWAYS = 7
};
enum srx_stencil_op {
  SRX_SOP_KEEP = 1, SRX_SOP_ZERO = 2, SRX_SOP_REPLACE = 3, SRX_SOP_INCR_SAT = 4,
  SRX_SOP_DECR_SAT = 5, SRX_SOP_INVERT = 6, SRX_SOP_INCR_WRAP = 7, SRX_SOP_DECR_WRAP = 8
};
enum srx_clear_flag { SRX_CLEAR_DEPTH = 1, SRX_CLEAR_STENCIL = 2 };
/* map_mode (constants.h:11-18) */
enum srx_map_mode {
  SRX_MAP_READ = 1, SRX_MAP_WRITE = 2, SRX_MAP_READ_WRITE = 3, SRX_MAP_WRITE_DISCARD = 4, SRX_MAP_WRITE_NO_OVERWRITE = 5
};

/* async_object_ids (constants.h:166-175) */
enum srx_query_id { SRX_QUERY_PIPELINE_STATISTICS = 3, SRX_QUERY_INTERNAL_STATISTICS = 5, SRX_QUERY_PIPELINE_PROFILES = 6 };

/* vs_output attribute modifiers (shader_regs.h:55-61) */
enum srx_attr_modifier {
  SRX_AM_LINEAR = 1, SRX_AM_CENTROID = 2, SRX_AM_NOINTERPOLATION = 4, SRX_AM_NOPERSPECTIVE = 8, SRX_AM_SAMPLE = 16
};

/* renderer_capacity.h:7-15 */
#define SRX_MAX_VS_INPUT_ATTRS 8
#define SRX_MAX_VS_OUTPUT_ATTRS 5 /* shader.cpp:45-77 initialises only ops 0..5 */
#define SRX_MAX_INPUT_SLOTS 32
#define SRX_MAX_SAMPLE_COUNT 16
#define SRX_MAX_RENDER_TARGET_SIZE 8192
#define SRX_MAX_SAMPLER_SLOTS 4

/* ---- descriptor structs (POD, reference field order) ---- */

/* viewport (salvia/core/viewport.h:5-12) */
typedef struct srx_viewport { float x, y, w, h, minz, maxz; } srx_viewport;

/* raster_desc (salvia/core/raster_state.h:17-40) */
typedef struct srx_raster_desc {
  int32_t fill_mode;   /* srx_fill_mode, default SOLID */
  int32_t cull_mode;   /* srx_cull_mode, default BACK */
  int32_t front_ccw;   /* default 0 */
  int32_t depth_bias;
  float depth_bias_clamp;
  float slope_scaled_depth_bias;
  int32_t depth_clip_enable;
  int32_t scissor_enable;
  int32_t multisample_enable;
  int32_t anti_aliased_line_enable;
} srx_raster_desc;

/* depth_stencil_op_desc (salvia/core/framebuffer.h:20-31) */
typedef struct srx_depth_stencil_op_desc {
  int32_t stencil_fail_op;
  int32_t stencil_depth_fail_op;
  int32_t stencil_pass_op;
  int32_t stencil_func;
} srx_depth_stencil_op_desc;

/* depth_stencil_desc (salvia/core/framebuffer.h:33-50) */
typedef struct srx_depth_stencil_desc {
  int32_t depth_enable;     /* default 1 */
  int32_t depth_write_mask; /* default 1 */
  int32_t depth_func;       /* default LESS */
  int32_t stencil_enable;   /* default 0 */
  uint32_t stencil_read_mask;  /* default 0xFF */
  uint32_t stencil_write_mask; /* default 0xFF */
  srx_depth_stencil_op_desc front_face;
  srx_depth_stencil_op_desc back_face;
} srx_depth_stencil_desc;

/* sampler_desc (salvia/resource/sampler.h:16-45) */
typedef struct srx_sampler_desc {
  int32_t min_filter, mag_filter, mip_filter; /* default POINT */
  int32_t mip_qual;                           /* default MI */
  int32_t addr_mode_u, addr_mode_v, addr_mode_w; /* default WRAP */
  float mip_lod_bias;
  uint32_t max_anisotropy;
  int32_t comparison_func;
  float border_color[4];
  float min_lod, max_lod;
} srx_sampler_desc;

/* input_element_desc (salvia/resource/input_layout.h:20-52) */
typedef struct srx_input_element_desc {
  const char* semantic_name; /* POSITION, NORMAL, TEXCOORD, COLOR ... (case-insensitive) */
  uint32_t semantic_index;
  uint32_t data_format;      /* srx_format */
  uint32_t input_slot;
  uint32_t aligned_byte_offset;
  uint32_t slot_class;       /* 0 = per vertex */
  uint32_t instance_data_step_rate;
} srx_input_element_desc;

/* pipeline_statistics + internal_statistics (salvia/core/async_object.h:90-175) */
typedef struct srx_pipeline_statistics {
  uint64_t ia_vertices, ia_primitives, vs_invocations, gs_invocations, gs_primitives;
  uint64_t cinvocations, cprimitives, ps_invocations, hs_invocations, ds_invocations, cs_invocations;
} srx_pipeline_statistics;
typedef struct srx_internal_statistics { uint64_t backend_input_pixels; } srx_internal_statistics;

/* ---- built-in shaders ----
 * The reference runs shaders as host C++ virtuals (cpp_vertex_shader::shader_prog,
 * shader.h:159-215) that device code cannot call.  The sample shaders are provided as
 * device functions selected by id; their constant blobs are listed next to each id.
 */
enum srx_vertex_shader_id {
  /* in: POSITION0 (r0), COLOR0 (r1).  out: pos = r0; attr0 = r1. consts: none */
  SRX_VS_PASSTHROUGH_COLOR = 1,
  /* in: POSITION0, NORMAL0, TEXCOORD0.  out: pos = mul(float4(P,1), wvp); attr0 = N; attr1 = UV.
     consts: float wvp[16] (row-major, row-vector convention of eflib::transform). */
  SRX_VS_WVP_NORMAL_UV = 2,
  /* in: POSITION0, COLOR0.  out: pos = mul(P, wvp); attr0 = COLOR0. consts: float wvp[16] */
  SRX_VS_WVP_COLOR = 3,
  /* in: POSITION0, COLOR0.  out: pos = mul(P, wvp); attr0 = COLOR0; attr1 = mul(P, light_wvp), the
     light's clip-space position for a shadow-map lookup (StandardShadowMap.cpp draw_cpp_vs).
     consts: float wvp[16], float light_wvp[16] */
  SRX_VS_WVP_COLOR_LIGHT = 4,
  /* ColorizedTriangle's vs_main (samples/ColorizedTriangle/ColorizedTriangle.cpp:29-53).
     in: POSITION0, NORMAL0 (float4).  out: pos = mul(pos, wvp); attr0 = norm;
     attr1..3 = lightPos0..2 - pos.  consts: float wvp[16], lightPos0[4], lightPos1[4], lightPos2[4] */
  SRX_VS_COLORIZED = 5,
  /* ComplexMesh's vs_mesh (samples/ComplexMesh/ComplexMesh.cpp:185-222).  in: POSITION0, NORMAL0.
     pos_es = transform(pos, wv); out: pos = transform(pos_es, proj); attr0 = (light - pos_es.xyz, 1);
     attr1 = (eye - pos_es.xyz, 1); attr2 = transform33(normal, wv).
     consts: float wv[16], proj[16], light_pos[4] (xyz used), eye_pos[4] (xyz used) */
  SRX_VS_MESH = 6,
  /* StandardShadowMap's Draw.savs (resources/ssm/Draw.savs).  in: POSITION0, NORMAL0, TEXCOORD0.
     out: pos = mul(pos, cameraWvp); attr0 = tex; attr1 = norm; attr2 = lightPos - pos;
     attr3 = cameraPos - pos; attr4 = mul(pos, lightWvp).
     consts: float cameraPos[4], cameraWvp[16], lightPos[4], lightWvp[16] */
  SRX_VS_SSM_DRAW = 7,
  /* StencilMirror's cup_vs_code (samples/StencilMirror/StencilMirror.cpp:33-57).  in: POSITION0,
     NORMAL0, TEXCOORD0.  out: pos = mul(pos, wvp); attr0 = tex; attr1 = norm; attr2 = lightPos - pos;
     attr3 = eyePos - pos.  consts: float wvp[16], eyePos[4], lightPos[4] */
  SRX_VS_CUP = 8,
  /* TextureAndBlending's vs_box (samples/TextureAndBlending/TextureAndBlending.cpp:60-86).
     in: POSITION0, TEXCOORD0.  out: pos = transform(pos, wvp); attr0 = pos; attr1 = tex.
     consts: float wvp[16] */
  SRX_VS_BOX = 9
};
enum srx_pixel_shader_id {
  /* color0 = attr0 (Gouraud). consts: none */
  SRX_PS_GOURAUD = 1,
  /* n = normalize3(attr0.xyz); k = clamp(dot(n, L), 0, 1); t = tex2D(s0, attr1.xy) with the
     SASL per-row/per-column derivatives and sample_2d_grad (sampler_api.cpp:13-18);
     color0 = (t.r*k, t.g*k, t.b*k, 1).  consts: float L[4] (xyz used, already normalised) */
  SRX_PS_LAMBERT_TEX = 2,
  /* color0 = tex2d(s0, attr0) with the cpp_pixel_shader coarse derivatives and the cached
     per-quad LOD (cpp_pixel_shader.cpp:22-31); alpha forced to 1 (AnisotropicFilter.cpp:146-152).
     consts: none */
  SRX_PS_TEX_CPP = 3,
  /* color0 = constant. consts: float c[4] */
  SRX_PS_CONSTANT = 4,
  /* Shadowed vertex colour with a 3x3 percentage-closer lookup: l = attr1.xyz / attr1.w, shadow-map
     centre ((l.x + 1) * 0.5, 1 - (l.y + 1) * 0.5) and the 3x3 tap layout of StandardShadowMap.cpp:89-100
     with spacing consts[2]; each tap is tex2dlod(s0, tap, lod 0).r and counts as lit when
     l.z - consts[1] <= tap; lit = count / 9; k = consts[0] + (1 - consts[0]) * lit;
     color0 = (attr0.r * k, attr0.g * k, attr0.b * k, attr0.a).
     consts: float ambient, bias, spacing, unused */
  SRX_PS_SHADOW_PCF = 5,
  /* ColorizedTriangle's ps (ColorizedTriangle.cpp:55-92): three point lights with 1/d^2 falloff,
     attr0 = normal, attr1..3 = light vectors; colour clamped to [0, 1], alpha 1.  consts: none */
  SRX_PS_COLORIZED = 6,
  /* ComplexMesh's ps_mesh (ComplexMesh.cpp:227-246): n.l * (0.8 + 3.5 pow(max(n.h, 0), 5) * 0.4)
     with l, e, n = normalize3(attr0, attr1, attr2).xyz.  consts: none */
  SRX_PS_MESH = 7,
  /* StandardShadowMap's draw_cpp_ps (StandardShadowMap.cpp:62-142): exponential shadow map with a
     3x3 Gaussian of tex2dlod(s0) taps around attr4's projection (sampler 0 = depth map; none bound:
     occlusion 0), texture tex2d(s1, attr0) (none bound: white), diffuse + pow(specular, shininess).
     consts: float ambient[4], diffuse[4], specular[4]; int32 shininess (its bits in float slot 12) */
  SRX_PS_SSM = 8,
  /* StencilMirror's cup_ps (StencilMirror.cpp:66-110): tex2d(s0, attr0) (none bound: white) times
     ambient + diffuse * clamp(l.n) + specular * clamp(reflect(l, n).e); alpha 1.
     consts: float ambient[4], diffuse[4], specular[4] */
  SRX_PS_CUP = 9,
  /* TextureAndBlending's ps_box (TextureAndBlending.cpp:96-114): tex2d(s0, attr1) with alpha 0.5.
     consts: none */
  SRX_PS_BOX = 10
};
enum srx_blend_shader_id {
  SRX_BLEND_OFF = 0,      /* ts_blend_off: target = src (TextureAndBlending.cpp:182-193) */
  SRX_BLEND_ON = 1,       /* ts_blend_on: target = lerp(target, src, src.a) (:168-180) */
  SRX_BLEND_DISABLED = 2  /* color_disabled_bs: no colour write (StencilMirror.cpp:156-166) */
};

/* ---- opaque handles ---- */
typedef struct srx_object_t* srx_handle; /* any handle below */
typedef struct srx_renderer_t* srx_renderer;
typedef struct srx_buffer_t* srx_buffer;
typedef struct srx_texture_t* srx_texture;
typedef struct srx_surface_t* srx_surface;
typedef struct srx_sampler_t* srx_sampler;
typedef struct srx_input_layout_t* srx_input_layout;
typedef struct srx_raster_state_t* srx_raster_state;
typedef struct srx_ds_state_t* srx_ds_state;
typedef struct srx_query_t* srx_query;
typedef struct srx_shader_code_t* srx_shader_code;
typedef struct srx_swap_chain_t* srx_swap_chain;

/* ---- library ---- */
uint32_t srx_abi_version(void);
/* "hip" or "cpu" */
const char* srx_backend_name(void);
/* Last error detail of this thread (hipError_t for SRX_E_HIP, else 0) and its message. */
int32_t srx_last_error(void);
const char* srx_last_error_string(void);
/* Reference counting for every handle type (shared_ptr semantics). */
void srx_retain(void* handle);
void srx_release(void* handle);

/* ---- renderer creation: create_software_renderer / create_benchmark_renderer (renderer.cpp:29-39) ---- */
#define SRX_RENDERER_ASYNC 0u     /* create_software_renderer */
#define SRX_RENDERER_BENCHMARK 1u /* create_benchmark_renderer (sync) */
int32_t srx_create_renderer(uint32_t kind, srx_renderer* out);

/* ---- creators (renderer.h:46-60) ---- */
int32_t srx_create_buffer(srx_renderer r, size_t size, srx_buffer* out);
int32_t srx_create_tex2d(srx_renderer r, size_t width, size_t height, size_t num_samples,
                         int32_t fmt, srx_texture* out);
int32_t srx_create_sampler(srx_renderer r, const srx_sampler_desc* desc, srx_texture tex,
                           srx_sampler* out);
int32_t srx_create_query(srx_renderer r, uint32_t id, srx_query* out);
/* create_input_layout(elem_descs, count, cpp_vertex_shader) -- the shader is named by id */
int32_t srx_create_input_layout(srx_renderer r, const srx_input_element_desc* descs, size_t count,
                                int32_t vs_id, srx_input_layout* out);
/* raster_state(raster_desc) (raster_state.cpp:23-33) */
int32_t srx_create_raster_state(srx_renderer r, const srx_raster_desc* desc, srx_raster_state* out);
/* depth_stencil_state(depth_stencil_desc) (framebuffer.cpp:168-230) */
int32_t srx_create_depth_stencil_state(srx_renderer r, const srx_depth_stencil_desc* desc,
                                       srx_ds_state* out);
/* Default-initialised descriptors (the reference's constructors). */
void srx_default_raster_desc(srx_raster_desc* d);
void srx_default_depth_stencil_desc(srx_depth_stencil_desc* d);
void srx_default_sampler_desc(srx_sampler_desc* d);

/* ---- resources: buffer / texture / surface (buffer.h:15-35, texture.h:17-102, surface.h:22-92) ---- */
size_t srx_buffer_size(srx_buffer b);
/* map(write) + memcpy + unmap (resource_manager.cpp:7-113) */
int32_t srx_buffer_write(srx_buffer b, size_t offset, const void* data, size_t size);
/* map(read): waits for queued work that writes the buffer */
int32_t srx_buffer_read(srx_buffer b, size_t offset, void* data, size_t size);
/* buffer::transfer (buffer.cpp:42-62): `count` rows of `size` bytes from data (rows stride_src apart)
   to the buffer at offset (rows stride_dest apart); SRX_INVALID_PARAMETER when the last row ends past
   the buffer (the reference asserts and copies nothing). */
int32_t srx_buffer_transfer(srx_buffer b, size_t offset, const void* data, size_t stride_dest, size_t stride_src,
                            size_t size, size_t count);
/* texture::subresource(level) (texture.h:47-52): retained surface handle */
int32_t srx_texture_surface(srx_texture t, size_t level, srx_surface* out);
/* texture_2d::gen_mipmap(filter, auto_gen) (texture2d.cpp:25-36) */
int32_t srx_texture_gen_mipmap(srx_texture t, int32_t filter, int32_t auto_gen);
size_t srx_texture_max_lod(srx_texture t);
size_t srx_texture_min_lod(srx_texture t);
size_t srx_surface_width(srx_surface s);
size_t srx_surface_height(srx_surface s);
size_t srx_surface_sample_count(srx_surface s);
int32_t srx_surface_format(srx_surface s);
/* Linear layout ((y*w + x)*samples + s) * elem_size (surface.cpp:277-295); size = pitch*h */
size_t srx_surface_size_bytes(srx_surface s);
int32_t srx_surface_write(srx_surface s, const void* data, size_t size);
int32_t srx_surface_read(srx_surface s, void* data, size_t size);
/* surface::resolve(target) (surface.cpp:123-140); queued like a command */
int32_t srx_surface_resolve(srx_renderer r, srx_surface src, srx_surface dst);

/* ---- state set (renderer.h:66-103) ---- */
int32_t srx_set_vertex_buffers(srx_renderer r, size_t start_slot, size_t count,
                               const srx_buffer* buffers, const size_t* strides,
                               const size_t* offsets);
int32_t srx_set_index_buffer(srx_renderer r, srx_buffer b, uint32_t index_fmt);
int32_t srx_set_input_layout(srx_renderer r, srx_input_layout layout);
int32_t srx_set_primitive_topology(srx_renderer r, uint32_t topology);
/* set_vertex_shader(cpp_vertex_shader) with its constants (declare_constant/set_constant) */
int32_t srx_set_vertex_shader(srx_renderer r, int32_t vs_id, const void* consts, size_t size);
int32_t srx_set_pixel_shader(srx_renderer r, int32_t ps_id, const void* consts, size_t size);
int32_t srx_set_ps_sampler(srx_renderer r, uint32_t slot, srx_sampler s);
int32_t srx_set_blend_shader(srx_renderer r, int32_t blend_id);
int32_t srx_set_rasterizer_state(srx_renderer r, srx_raster_state rs);
int32_t srx_set_depth_stencil_state(srx_renderer r, srx_ds_state dss, int32_t stencil_ref);
int32_t srx_set_render_targets(srx_renderer r, size_t color_target_count,
                               const srx_surface* color_targets, srx_surface ds_target);
int32_t srx_set_viewport(srx_renderer r, const srx_viewport* vp);
/* set_vs_sampler(name, sampler) (renderer.h:80) by slot.  No built-in vertex shader samples a
   texture; the binding is kept (and validated) for the renderer state. */
int32_t srx_set_vs_sampler(srx_renderer r, uint32_t slot, srx_sampler s);

/* ---- state get (renderer.h:104-113; renderer_impl.cpp:61-157) ---- */
/* get_index_buffer: retained handle (release it), NULL when none is bound */
int32_t srx_get_index_buffer(srx_renderer r, srx_buffer* out);
uint32_t srx_get_index_format(srx_renderer r);       /* default SRX_FMT_R16_UINT */
uint32_t srx_get_primitive_topology(srx_renderer r); /* default SRX_TOPOLOGY_TRIANGLE_LIST */
int32_t srx_get_vertex_shader(srx_renderer r);       /* srx_vertex_shader_id, 0 when none */
/* get_rasterizer_state: retained handle; a fresh renderer holds the default state (cull back) */
int32_t srx_get_rasterizer_state(srx_renderer r, srx_raster_state* out);
int32_t srx_get_pixel_shader(srx_renderer r);        /* srx_pixel_shader_id, 0 when none */
int32_t srx_get_blend_shader(srx_renderer r);        /* srx_blend_shader_id */
int32_t srx_get_viewport(srx_renderer r, srx_viewport* vp);
/* renderer_impl::get_depth_stencil_state / get_stencil_ref: retained handle and the reference value */
int32_t srx_get_depth_stencil_state(srx_renderer r, srx_ds_state* out, int32_t* stencil_ref);
/* The descriptors behind state objects (raster_state::get_desc, depth_stencil_state::get_desc). */
int32_t srx_raster_state_desc(srx_raster_state rs, srx_raster_desc* d);
int32_t srx_depth_stencil_state_desc(srx_ds_state ds, srx_depth_stencil_desc* d);

/* ---- map / unmap (renderer.h:61-63; resource_manager.cpp:7-113) ----
 * One resource is mapped per renderer at a time (a second map fails).  Every mode waits for the
 * renderer's queued work and points data at host memory holding the resource's current bytes;
 * srx_unmap writes them back for every mode but SRX_MAP_READ.  Buffers: row_pitch = depth_pitch =
 * size (buffer.cpp:25); surfaces: row_pitch = width * samples * texel bytes, depth_pitch =
 * row_pitch * height (surface.cpp:111-112).  data stays valid until srx_unmap. */
typedef struct srx_mapped_resource {
  void* data;
  uint32_t row_pitch;
  uint32_t depth_pitch;
} srx_mapped_resource;
int32_t srx_map_buffer(srx_renderer r, srx_mapped_resource* mapped, srx_buffer b, int32_t mode);
int32_t srx_map_surface(srx_renderer r, srx_mapped_resource* mapped, srx_surface s, int32_t mode);
int32_t srx_unmap(srx_renderer r);

/* ---- SASL pixel shaders (renderer.h:82-86, 136-147) ----
 * srx_compile_shader is compile(code, lang) (renderer.cpp): it accepts the pixel-shader subset the
 * reference's SIMD code generator implements (sasl/src/codegen/cg_simd.cpp): struct inputs/outputs
 * or parameters with TEXCOORD(n) inputs and COLOR(n)/SV_Target(n) outputs, float/int uniforms and
 * samplers, if/else, while, do-while, for, break, + - * / % (ints: + - *), comparisons, unary
 * minus, swizzles and write masks, vector constructors, and the intrinsics abs acos asin atan
 * ceil clamp cos cosh cross ddx ddy distance dot exp exp2 floor fmod frac ldexp length lerp log
 * log10 log2 max min normalize pow rcp reflect round rsqrt saturate sin sinh sqrt tan tanh
 * tex2D tex2Dgrad trunc.  Constructs the reference leaves unimplemented in pixel shaders (&&, ||,
 * ?:, casts, continue, tex2Dlod/tex2Dbias/tex2Dproj) and matrices, arrays and helper functions
 * return SRX_E_UNSUPPORTED; syntax and type errors SRX_INVALID_PARAMETER; srx_last_error_string()
 * names the line.  Limits: 256 instructions, 32 registers, 32 literals and uniforms, 4 samplers,
 * 8 inputs and outputs, nesting depth 8.  A quad stops after 65536 loop iterations.
 *
 * Pixel-shader input i (in TEXCOORD index order) is vertex-shader output attribute i
 * (shader_unit.cpp:106-145 with a C++ vertex shader).  Colour target 0 receives COLOR(0) (zeros
 * when the shader has none; components the output type lacks are zero).  A shader set with
 * srx_set_pixel_shader_code shades instead of the srx_set_pixel_shader id (rasterizer.cpp:1402-1407)
 * until it is cleared with NULL.  Variables (srx_set_ps_variable) and samplers
 * (srx_set_ps_sampler_named) are kept by name on the renderer and applied at each draw
 * (render_core.cpp:80-93); names the shader lacks are ignored, unset variables read zero, unbound
 * samplers sample zero.  Draws whose colour such a shader writes run the in-order raster kernel. */
enum srx_shader_language { SRX_LANG_VERTEX_SHADER = 2, SRX_LANG_PIXEL_SHADER = 3 };
typedef struct srx_shader_info {
  uint32_t input_count, output_count;
  uint32_t input_components[8];
  uint32_t output_targets[8], output_components[8]; /* outputs in COLOR index order */
  uint32_t uniform_count, sampler_count, instruction_count, register_count;
} srx_shader_info;
int32_t srx_compile_shader(const char* code, size_t length, int32_t lang, srx_shader_code* out);
/* shader_object::get_reflection (shader_reflection layouts) */
int32_t srx_shader_reflect(srx_shader_code code, srx_shader_info* info);
int32_t srx_set_pixel_shader_code(srx_renderer r, srx_shader_code code);
/* get_pixel_shader_code: retained handle, NULL when none is set */
int32_t srx_get_pixel_shader_code(srx_renderer r, srx_shader_code* out);
int32_t srx_set_ps_variable(srx_renderer r, const char* name, const void* data, size_t size);
/* set_ps_sampler(name, sampler) (renderer.h:83); NULL unbinds */
int32_t srx_set_ps_sampler_named(srx_renderer r, const char* name, srx_sampler s);
/* pixel_shader_unit::update + execute (shader_unit.cpp:106-176) on caller-supplied quads with the
 * renderer's variables and named samplers: in holds nquads * 4 pixels (0 1 / 2 3) * input_count
 * float4, out receives nquads * 4 * output_count float4.  Waits for the renderer's queued work. */
int32_t srx_ps_execute_quads(srx_renderer r, srx_shader_code code, const float* in, float* out, size_t nquads);

/* ---- render operations (renderer.h:117-129) ---- */
/* begin/end do not wait for the GPU: they queue the counters' read-back on the renderer's stream. */
int32_t srx_begin(srx_renderer r, srx_query q);
int32_t srx_end(srx_renderer r, srx_query q);
/* async_object::get (async_object.h:62-81): SRX_ASYNC_ERROR while the query runs or before its
   first end; with do_not_wait, SRX_ASYNC_TIMEOUT until the GPU has finished the work queued before
   srx_end, else it waits; then SRX_ASYNC_READY with
   data: srx_pipeline_statistics / srx_internal_statistics / uint64_t[8] pipeline profiles in ns
   (gather_vtx, vtx_proc, clipping, compact_clip, vp_trans, tri_dispatch, ras, profiled draws);
   a uint64_t[10] buffer also receives ras split into visibility (8) and deferred shading (9). */
int32_t srx_get_data(srx_renderer r, srx_query q, void* data, size_t size, int32_t do_not_wait);
int32_t srx_draw(srx_renderer r, size_t startpos, size_t primcnt);
int32_t srx_draw_index(srx_renderer r, size_t startpos, size_t primcnt, int32_t basevert);
int32_t srx_clear_color(srx_renderer r, srx_surface color_target, const float rgba[4]);
int32_t srx_clear_depth_stencil(srx_renderer r, srx_surface ds_target, uint32_t flags, float depth,
                                uint32_t stencil);
int32_t srx_flush(srx_renderer r);

/* ---- swap chain without a window (salvia/include/salvia/ext/swap_chain/swap_chain.h:16-21) ----
 * salviax_create_swap_chain_and_renderer's swap chain for an existing renderer, with no display
 * (present_impl's copy to a D3D11/GL window is out of scope): a back buffer of width x height x
 * samples in fmt, plus a single-sample surface of the same size when samples > 1
 * (swap_chain_impl.cpp:14-35).  srx_swap_chain_present is present() (:41-49): it flushes the
 * renderer and resolves the back buffer into the presented surface, and returns when that image is
 * complete; srx_swap_chain_presented returns that surface (the back buffer itself for 1 sample).
 * Both getters return retained handles. */
int32_t srx_create_swap_chain(srx_renderer r, size_t width, size_t height, size_t samples, int32_t fmt,
                              srx_swap_chain* out);
int32_t srx_swap_chain_surface(srx_swap_chain sc, srx_surface* out);
int32_t srx_swap_chain_present(srx_swap_chain sc);
int32_t srx_swap_chain_presented(srx_swap_chain sc, srx_surface* out);

/* ---- a draw that allocates nothing ----
 * No counterpart in the reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops and harnesses that forbid hipMalloc): srx_draw with every buffer
 * beyond its inputs and outputs carved from one fixed region of SRX_NOALLOC_SCRATCH_BYTES per GPU,
 * which the library holds until the process ends.  Call it with the renderer's device current.
 * A draw whose scratch does not fit returns SRX_OUT_OF_MEMORY and writes no target (about 100k
 * primitives over a 1024x1024 1-sample target fit; the binned-record count decides).  It is not
 * profiled by a pipeline_profiles query and cannot be recorded into a batch.  Every call on one
 * GPU shares the region: calls from renderers on different streams must not overlap (one
 * renderer's calls are ordered by its stream). */
#define SRX_NOALLOC_SCRATCH_BYTES (64u << 20)
int32_t srx_draw_noalloc(srx_renderer r, size_t startpos, size_t primcnt);

/* ---- batches of independent render jobs ----
 * No counterpart in the reference, whose renderer executes one command at a time
 * (render_core.cpp:30-61).  Between srx_batch_begin and srx_batch_end the renderer records
 * srx_clear_color, srx_clear_depth_stencil (both planes), srx_draw / srx_draw_index and
 * srx_surface_resolve instead of executing them; srx_batch_launch replays all of them as one
 * launch per pipeline stage (every job's workgroups in one grid) from a HIP graph: clears first,
 * then the draws, then the resolves.  Each job's output is the output of the same commands
 * executed one by one.  Requirements, checked while recording (SRX_E_UNSUPPORTED otherwise, also
 * returned by srx_batch_end): every draw is opaque (blend off or disabled) with an
 * order-independent depth/stencil test (early-Z, or a stencil that keeps its value), no query is
 * running, and no surface is a target of two
 * draws, written after it is read or sampled while written.  Resources stay referenced by the
 * batch; their contents may change between launches (vertex positions included: a launch whose
 * records outgrow the prepared capacity stays in bounds and sets status bit 0). */
typedef struct srx_batch_t* srx_batch;
typedef struct srx_batch_info {
  uint32_t jobs, clears, resolves;
  uint32_t tile;         /* GPU tile side (8, 16, 32) */
  uint32_t bin_items;    /* records binned per workgroup (2048, 4096, 8192, 16384) */
  uint32_t vis_threads;  /* lanes per GPU tile in the visibility kernel (64, 128, 256) */
  uint32_t vis_launches; /* visibility launches (one per shader/sample-count variant) */
  uint32_t pad;
  uint64_t primitives, records, device_bytes;
} srx_batch_info;
typedef struct srx_batch_tuning {
  int32_t tile, bin_items, vis_threads;
  float ms; /* mean time of one launch */
} srx_batch_tuning;
int32_t srx_batch_begin(srx_renderer r);
/* Sizes the batch with one counting pass over its geometry (synchronous) and captures its graph. */
int32_t srx_batch_end(srx_renderer r, srx_batch* out);
int32_t srx_batch_launch(srx_batch b); /* queued on the renderer's stream */
/* Waits for the renderer's stream; status bit 0: records beyond the prepared capacity were
 * dropped, bit 1: clipping overflow.  Zero after every complete launch. */
int32_t srx_batch_status(srx_batch b, uint32_t* status);
/* GPU tile side 8, 16 or 32 (8 and 32: single-sample jobs, and at most 8192 tiles per job), records
 * per binning workgroup 2048, 4096, 8192 or 16384, lanes per tile 64, 128 or 256.  A new tile size
 * repeats the counting pass once. */
int32_t srx_batch_configure(srx_batch b, int32_t tile, int32_t bin_items, int32_t vis_threads);
/* Times `reps` launches of every (tile, bin_items, vis_threads) the batch's jobs allow, reports
 * them in results (up to capacity; *count = configurations tried) and leaves the batch
 * configured with the fastest. */
int32_t srx_batch_autotune(srx_batch b, uint32_t reps, srx_batch_tuning* results, size_t capacity, size_t* count);
int32_t srx_batch_get_info(srx_batch b, srx_batch_info* info);

/* ---- request server ----
 * No counterpart in the reference.  A server renders requests -- a resident mesh of non-indexed
 * triangles (per vertex: position xyzw, colour rgba, 8 floats) drawn with the pass-through vertex
 * shader, the Gouraud pixel shader, cull none and depth LESS into a fresh target of the requested
 * size and sample count (cleared to clear_color and depth 1.0), then resolved -- and returns the
 * resolved RGBA8 image (and, on request, the RG32F depth/stencil target), byte for byte what the
 * renderer API produces for the same commands.  A dispatcher thread admits arrived requests in
 * arrival order by their device footprint (targets, geometry scratch, bins) into batches of up to
 * max_batch_jobs, and launches a new batch on a free stream while earlier ones still run.  Meshes
 * and textures are kept resident under an LRU cap (pool_bytes); a request names them by key and
 * may carry their data for a miss.  Textures only occupy memory: the Gouraud shader samples
 * nothing.  All entry points are thread-safe. */
typedef struct srx_server_t* srx_server;
typedef struct srx_server_desc {
  uint64_t pool_bytes;     /* LRU cap of resident meshes and textures */
  uint64_t arena_bytes;    /* device memory of the batches in flight */
  uint64_t staging_bytes;  /* pinned host memory: batch tables and returned images */
  uint32_t max_batch_jobs; /* requests per batch */
  uint32_t streams;        /* batches in flight (1..8) */
  int32_t bin_items;       /* as srx_batch_configure */
  int32_t vis_threads;
} srx_server_desc;
typedef struct srx_request {
  uint64_t mesh_key;
  const void* mesh_data; /* 3 * n_tris vertices, or NULL when the mesh is resident */
  uint32_t n_tris;
  uint32_t texture_size;  /* RGBA8 texture_size^2 level 0, chain built on the device */
  uint64_t texture_key;   /* 0: none */
  const void* texture_data;
  uint32_t width, height, samples, want_depth;
  float clear_color[4];
  int64_t arrival_ns; /* server clock (srx_server_now); not admitted before */
} srx_request;
typedef struct srx_completion {
  uint64_t ticket;
  int32_t status;      /* SRX_OK or why the request failed */
  uint32_t batch_jobs; /* requests in the batch that carried it */
  int64_t arrival_ns, admitted_ns, done_ns;
  const void* color; /* resolved RGBA8, width*height*4 bytes, valid until srx_server_release */
  uint64_t color_bytes;
  const void* depth; /* RG32F per sample when want_depth */
  uint64_t depth_bytes;
} srx_completion;
typedef struct srx_server_stats {
  uint64_t batches, jobs, retries, uploads, evictions, max_batch_jobs;
  uint64_t pool_bytes, resident_meshes, resident_textures, arena_peak_bytes, queued, running;
} srx_server_stats;
void srx_default_server_desc(srx_server_desc* d); /* 200 GiB pool, 64 GiB arena, 2 GiB staging */
int32_t srx_create_server(srx_renderer r, const srx_server_desc* d, srx_server* out);
int32_t srx_server_put_mesh(srx_server s, uint64_t key, const void* vertices, uint32_t n_tris);
int32_t srx_server_put_texture(srx_server s, uint64_t key, const void* rgba8, uint32_t size);
int32_t srx_server_submit(srx_server s, const srx_request* req, uint64_t* ticket);
/* Up to capacity finished requests (waits up to wait_ms for the first); each stays valid, and keeps
 * its staging, until srx_server_release. */
int32_t srx_server_poll(srx_server s, srx_completion* out, size_t capacity, size_t* count, int32_t wait_ms);
int32_t srx_server_release(srx_server s, uint64_t ticket);
int64_t srx_server_now(srx_server s);
int32_t srx_server_get_stats(srx_server s, srx_server_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* SRX_H_ */
