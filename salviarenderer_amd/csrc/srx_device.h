// Device-side types and arithmetic of the Salvia rasterization pipeline for gfx950.
//
// Every floating-point recipe here reproduces the reference's x86 evaluation order operation
// by operation (one IEEE fp32 op per SSE lane op).  The library is compiled with
// -ffp-contract=off, IEEE division/sqrt and fp32 denormals preserved, so each expression below
// rounds exactly like the reference's SSE code.
#pragma once

#ifndef __HIPCC_RTC__  // run-time compilation provides the HIP builtins and integer types itself
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif
#ifndef INT32_MIN
#define INT32_MIN (-2147483647 - 1)
#endif

namespace srx {

constexpr int MAX_ATTRS = 5;  // shader.cpp:45-77 initialises vs_output ops for 0..5 attributes
constexpr int MAX_REGS = MAX_ATTRS + 1;
constexpr int MAX_LEVELS = 16;
constexpr int MAX_SAMPLERS = 4;
constexpr int MAX_INPUTS = 8;
constexpr int VS_CONSTS = 48;  // floats of a vertex shader's constant blob
constexpr int TILE = 64;  // reference tile size (rasterizer.cpp:32); baked into the arithmetic

enum { PF_RGBA32F = 0, PF_RGB32F = 1, PF_BGRA8 = 2, PF_RGBA8 = 3, PF_R32F = 4, PF_RG32F = 5, PF_R32I = 6 };

struct surface_desc {
  uint8_t* data;
  int32_t w, h, samples, pf;
};

struct sampler_desc_dev {
  int32_t min_filter, mag_filter, mip_filter, mip_qual, addr_u, addr_v;
  float lod_bias;
  uint32_t max_aniso;
  float border[4];
  int32_t nlevels, max_lod, min_lod;
  uint32_t size[3];
  // 1: anisotropic mip filter (max_aniso > 1), linear min/mag, wrap/wrap, every level single-sample
  // RGBA8 -- the sampler's common case, which has a compact inline path
  int32_t fast_aniso;
  surface_desc levels[MAX_LEVELS];
};

struct input_elem_dev {
  const uint8_t* base;  // buffer + slot offset + element offset
  uint64_t avail;       // bytes readable from base
  uint32_t stride;
  uint32_t fmt;
  int32_t reg;
  float default_w;
};

// vs_output attribute modifiers (shader_regs.h)
enum { AM_NOINTERPOLATION = 4, AM_NOPERSPECTIVE = 8 };

// the sampler table of a draw (one descriptor per slot), in device memory
struct ps_samplers {
  sampler_desc_dev s[MAX_SAMPLERS];
};

struct sasl_program;
// ps_id of a draw whose pixel shader is a compiled SASL program (set_pixel_shader_code)
constexpr int32_t PS_SASL = 0x100;

// Everything a draw's kernels read; passed by value as a kernel argument.
struct draw_params {
  // vertex stage
  int32_t vs_id, nin, nattr;
  uint32_t attr_mod[MAX_ATTRS];
  float vs_consts[VS_CONSTS];
  input_elem_dev in[MAX_INPUTS];
  uint32_t nverts;  // vertex-shader outputs stored (one more slot holds an all-zero input)
  // index fetch
  const uint8_t* ib;
  uint64_t ib_size;
  uint32_t istride, istart_bytes, indexed, topo;
  int32_t basevert;
  uint32_t prim_count;
  // raster state
  int32_t cull_mode, front_ccw;
  float vp[6];         // viewport x y w h minz maxz
  float target_w, target_h;
  int32_t samples;
  float sp[4][2];
  int32_t tiles_x, tiles_y;   // 64-pixel reference tiles
  int32_t gt, gtiles_x, gtiles_y;  // GPU tile size (16/32/64) and counts
  // framebuffer
  surface_desc color;  // colour target 0 (data == nullptr: none)
  surface_desc ds;     // RG32F (data == nullptr: none)
  int32_t depth_enable, depth_func, read_depth, write_depth;
  int32_t stencil_enable;
  uint32_t stencil_ref, read_mask, write_mask;
  int32_t st_func[2], st_pass_op[2];  // [0] front face, [1] back face
  int32_t early_z, blend_id;
  int32_t count_stats;  // a statistics query is active: accumulate the device-side counters
  int32_t vis_stencil;  // stencil on with KEEP on pass for both faces: the stencil test is a fixed mask
  // full clears of the targets that the order-independent visibility pass applies as it writes them
  // (it then reads nothing from the depth/stencil target, and samples it leaves take the clear values)
  int32_t clear_ds, clear_color;
  float clear_depth;
  uint32_t clear_stencil;
  float clear_rgba[4];
  // pixel shader
  int32_t ps_id;
  float ps_consts[16];
  int32_t has_sampler[MAX_SAMPLERS];
  const sasl_program* sasl;  // ps_id == PS_SASL: the compiled SASL program (device memory)
};

// ------------------------------------------------------------------ scalar helpers
// eflib/include/eflib/math/math.h:130-158 (double rounding of a float argument)
__device__ __forceinline__ int fast_roundi(double d) { return static_cast<int>(floor(d + 0.5)); }
__device__ __forceinline__ int fast_floori(double d) { return fast_roundi(d - static_cast<double>(0.5f - 1.5e-8)); }
__device__ __forceinline__ int fast_ceili(double d) { return fast_roundi(d + static_cast<double>(0.5f - 1.5e-8)); }

// math.h:89-105
__device__ __forceinline__ float fast_log2(float val) {
  int32_t x = __float_as_int(val);
  int log_2 = ((x >> 23) & 255) - 128;
  x &= ~(255 << 23);
  x += 127 << 23;
  float f = __int_as_float(x);
  f = ((-1.0f / 3) * f + 2) * f - 2.0f / 3;
  return f + static_cast<float>(log_2);
}
// math.h:111-124
__device__ __forceinline__ float fast_round(float val) {
  uint32_t n = __float_as_uint(val);
  float bias = __uint_as_float(((23u + 127u) << 23) + (n & 0x80000000u));
  float f = val + bias;
  return f - bias;
}
__device__ __forceinline__ float fast_floor(float val) {
  float f = fast_round(val);
  return (f > val) ? f - 1 : f;
}
__device__ __forceinline__ bool equal_eps(float a, float b) { return fabsf(a - b) <= 1.1920928955078125e-07f; }

// x86 cvttps2dq: truncate; NaN and out of range give INT_MIN
__device__ __forceinline__ int32_t cvtt(float f) {
  if (!(f >= -2147483648.0f && f < 2147483648.0f)) return INT32_MIN;
  return static_cast<int32_t>(f);
}
__device__ __forceinline__ float sse_max(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float sse_min(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float std_min(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float std_max(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float std_clamp(float v, float lo, float hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }
__device__ __forceinline__ int std_clampi(int v, int lo, int hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }
__device__ __forceinline__ float sse_floor(float f) {
  float ip = static_cast<float>(cvtt(f));
  return (ip > f) ? ip - 1.0f : ip;
}

// colors.h:182-193: clamp(x*255) with maxps/minps then cvtps2dq (round half to even)
__device__ __forceinline__ uint32_t unorm8(float x) {
  float m = x * 255.0f;
  m = sse_max(m, 0.0f);
  m = sse_min(m, 255.0f);
  return static_cast<uint32_t>(__float2int_rn(m));
}

// ------------------------------------------------------------------ texel formats
__host__ __device__ __forceinline__ int pf_size(int pf) {
  switch (pf) {
  case PF_RGBA32F: return 16;
  case PF_RGB32F: return 12;
  case PF_RG32F: return 8;
  default: return 4;
  }
}

__device__ __forceinline__ float4 load_texel(int pf, const uint8_t* p) {
  const float inv = 1.0f / 255;
  switch (pf) {
  case PF_RGBA8: {
    uint32_t u = *reinterpret_cast<const uint32_t*>(p);
    return make_float4((u & 0xFF) * inv, ((u >> 8) & 0xFF) * inv, ((u >> 16) & 0xFF) * inv, (u >> 24) * inv);
  }
  case PF_BGRA8: {
    uint32_t u = *reinterpret_cast<const uint32_t*>(p);
    return make_float4(((u >> 16) & 0xFF) * inv, ((u >> 8) & 0xFF) * inv, (u & 0xFF) * inv, (u >> 24) * inv);
  }
  case PF_RGBA32F: return *reinterpret_cast<const float4*>(p);
  case PF_RGB32F: {
    const float* f = reinterpret_cast<const float*>(p);
    return make_float4(f[0], f[1], f[2], 1.0f);
  }
  case PF_R32F: return make_float4(*reinterpret_cast<const float*>(p), 0.0f, 0.0f, 0.0f);
  case PF_RG32F: {
    float2 f = *reinterpret_cast<const float2*>(p);
    return make_float4(f.x, f.y, 0.0f, 0.0f);
  }
  case PF_R32I: return make_float4(static_cast<float>(*reinterpret_cast<const int32_t*>(p)), 0.0f, 0.0f, 0.0f);
  }
  return make_float4(0, 0, 0, 0);
}

__device__ __forceinline__ void store_texel(int pf, uint8_t* p, float4 c) {
  switch (pf) {
  case PF_RGBA8:
    *reinterpret_cast<uint32_t*>(p) = unorm8(c.x) | (unorm8(c.y) << 8) | (unorm8(c.z) << 16) | (unorm8(c.w) << 24);
    return;
  case PF_BGRA8:
    *reinterpret_cast<uint32_t*>(p) = unorm8(c.z) | (unorm8(c.y) << 8) | (unorm8(c.x) << 16) | (unorm8(c.w) << 24);
    return;
  case PF_RGBA32F: *reinterpret_cast<float4*>(p) = c; return;
  case PF_RGB32F: {
    float* f = reinterpret_cast<float*>(p);
    f[0] = c.x; f[1] = c.y; f[2] = c.z;
    return;
  }
  case PF_R32F: *reinterpret_cast<float*>(p) = c.x; return;
  case PF_RG32F: *reinterpret_cast<float2*>(p) = make_float2(c.x, c.y); return;
  case PF_R32I: *reinterpret_cast<int32_t*>(p) = static_cast<int32_t>(c.x + 0.5f); return;
  }
}

// The 16 bytes of a fill with texel value c: the texel's words repeated (texel sizes 4, 8, 16).
__device__ __forceinline__ uint4 fill_words(int pf, float4 c) {
  switch (pf) {
  case PF_RGBA8: {
    const uint32_t u = unorm8(c.x) | (unorm8(c.y) << 8) | (unorm8(c.z) << 16) | (unorm8(c.w) << 24);
    return make_uint4(u, u, u, u);
  }
  case PF_BGRA8: {
    const uint32_t u = unorm8(c.z) | (unorm8(c.y) << 8) | (unorm8(c.x) << 16) | (unorm8(c.w) << 24);
    return make_uint4(u, u, u, u);
  }
  case PF_RGBA32F: return make_uint4(__float_as_uint(c.x), __float_as_uint(c.y), __float_as_uint(c.z), __float_as_uint(c.w));
  case PF_RG32F: return make_uint4(__float_as_uint(c.x), __float_as_uint(c.y), __float_as_uint(c.x), __float_as_uint(c.y));
  case PF_R32I: {
    const uint32_t u = static_cast<uint32_t>(static_cast<int32_t>(c.x + 0.5f));
    return make_uint4(u, u, u, u);
  }
  default: {  // R32F
    const uint32_t u = __float_as_uint(c.x);
    return make_uint4(u, u, u, u);
  }
  }
}

__device__ __forceinline__ float4 lerp4(float4 a, float4 b, float t) {
  return make_float4(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t, a.w + (b.w - a.w) * t);
}

}  // namespace srx
