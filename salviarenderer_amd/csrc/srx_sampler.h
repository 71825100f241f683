// Texture sampling for gfx950: the reference sampler (salvia/src/resource/sampler.cpp) with its
// SSE addressers evaluated lane by lane, so addressing, filtering, LOD and anisotropic probes
// round exactly as on the CPU.
#pragma once

#include "srx_device.h"

namespace srx {

enum { ADDR_WRAP = 0, ADDR_MIRROR = 1, ADDR_CLAMP = 2, ADDR_BORDER = 3 };
enum { FILTER_POINT = 0, FILTER_LINEAR = 1, FILTER_ANISO = 2 };
enum { MIP_LO = 0, MIP_MI = 1, MIP_HI = 2 };

extern __constant__ float c_ewa[256];

struct coord2 {
  int x0, y0, x1, y1;
  float tx, ty;
};

// Texel address; coordinates are clamped to the surface, which changes nothing for the in-range
// coordinates every addresser produces and keeps NaN/huge texture coordinates (where the reference
// reads out of bounds) inside the allocation.
__device__ __forceinline__ const uint8_t* texel_addr(const surface_desc& s, int x, int y) {
  x = min(max(x, 0), s.w - 1);
  y = min(max(y, 0), s.h - 1);
  return s.data + (static_cast<size_t>(y) * s.w + x) * static_cast<size_t>(s.samples) * pf_size(s.pf);
}

// bilinear blend of four texels (colors.h:372-488 per format): corners 0 1 on the upper row, 2 3 below
__device__ __forceinline__ float4 lerp_2d4(const surface_desc& s, const int* xs, const int* ys, float tx, float ty) {
  const uint8_t* a0 = texel_addr(s, xs[0], ys[0]);
  const uint8_t* a1 = texel_addr(s, xs[1], ys[1]);
  const uint8_t* a2 = texel_addr(s, xs[2], ys[2]);
  const uint8_t* a3 = texel_addr(s, xs[3], ys[3]);
  switch (s.pf) {
  case PF_RGBA8:
  case PF_BGRA8: {
    uint32_t u[4] = {*reinterpret_cast<const uint32_t*>(a0), *reinterpret_cast<const uint32_t*>(a1),
                     *reinterpret_cast<const uint32_t*>(a2), *reinterpret_cast<const uint32_t*>(a3)};
    float ch[4][4];
    for (int k = 0; k < 4; ++k) {
      float b0 = static_cast<float>(u[k] & 0xFF), b1 = static_cast<float>((u[k] >> 8) & 0xFF);
      float b2 = static_cast<float>((u[k] >> 16) & 0xFF), b3 = static_cast<float>(u[k] >> 24);
      if (s.pf == PF_RGBA8) {
        ch[k][0] = b0; ch[k][1] = b1; ch[k][2] = b2; ch[k][3] = b3;
      } else {
        ch[k][0] = b2; ch[k][1] = b1; ch[k][2] = b0; ch[k][3] = b3;
      }
    }
    float r[4];
    for (int q = 0; q < 4; ++q) {
      float c01 = ch[0][q] + (ch[1][q] - ch[0][q]) * tx;
      float c23 = ch[2][q] + (ch[3][q] - ch[2][q]) * tx;
      r[q] = (c01 + (c23 - c01) * ty) * (1.0f / 255);
    }
    return make_float4(r[0], r[1], r[2], r[3]);
  }
  case PF_RGBA32F: {
    float4 c0 = load_texel(s.pf, a0), c1 = load_texel(s.pf, a1), c2 = load_texel(s.pf, a2), c3 = load_texel(s.pf, a3);
    return lerp4(lerp4(c0, c1, tx), lerp4(c2, c3, tx), ty);
  }
  case PF_R32F: {
    float c0 = *reinterpret_cast<const float*>(a0), c1 = *reinterpret_cast<const float*>(a1);
    float c2 = *reinterpret_cast<const float*>(a2), c3 = *reinterpret_cast<const float*>(a3);
    float c01 = c0 + (c1 - c0) * tx, c23 = c2 + (c3 - c2) * tx;
    return make_float4(c01 + (c23 - c01) * ty, 0.0f, 0.0f, 0.0f);
  }
  case PF_RG32F: {
    // the .g channel lerps from the .r base (colors.h:469-478)
    float2 c0 = *reinterpret_cast<const float2*>(a0), c1 = *reinterpret_cast<const float2*>(a1);
    float2 c2 = *reinterpret_cast<const float2*>(a2), c3 = *reinterpret_cast<const float2*>(a3);
    float c01r = c0.x + (c1.x - c0.x) * tx, c01g = c0.x + (c1.y - c0.y) * tx;
    float c23r = c2.x + (c3.x - c2.x) * tx, c23g = c2.x + (c3.y - c2.y) * tx;
    return make_float4(c01r + (c23r - c01r) * ty, c01r + (c23g - c01g) * ty, 0.0f, 0.0f);
  }
  case PF_RGB32F: {
    const float* f0 = reinterpret_cast<const float*>(a0);
    const float* f1 = reinterpret_cast<const float*>(a1);
    const float* f2 = reinterpret_cast<const float*>(a2);
    const float* f3 = reinterpret_cast<const float*>(a3);
    float c01[3], c23[3];
    for (int q = 0; q < 3; ++q) {
      c01[q] = f0[0] + (f1[q] - f0[q]) * tx;
      c23[q] = f2[0] + (f3[q] - f2[q]) * tx;
    }
    return make_float4(c01[0] + (c23[0] - c01[0]) * ty, c01[0] + (c23[1] - c01[1]) * ty,
                       c01[0] + (c23[2] - c01[2]) * ty, 1.0f);
  }
  case PF_R32I: {
    int32_t c0 = *reinterpret_cast<const int32_t*>(a0), c1 = *reinterpret_cast<const int32_t*>(a1);
    int32_t c2 = *reinterpret_cast<const int32_t*>(a2), c3 = *reinterpret_cast<const int32_t*>(a3);
    float c01 = c0 + (c1 - c0) * tx, c23 = c2 + (c3 - c2) * tx;
    return make_float4(c01 + (c23 - c01) * ty, 0.0f, 0.0f, 0.0f);
  }
  }
  return make_float4(0, 0, 0, 0);
}
// bilinear blend of the 2x2 footprint
__device__ __forceinline__ float4 lerp_2d(const surface_desc& s, const coord2& c) {
  const int xs[4] = {c.x0, c.x1, c.x0, c.x1}, ys[4] = {c.y0, c.y0, c.y1, c.y1};
  return lerp_2d4(s, xs, ys, c.tx, c.ty);
}

// m / size, correctly rounded.  For a power-of-two size the quotient is m * 2^-k exactly (both are the
// exact value m * 2^-k rounded once), so the division becomes one multiply.
__device__ __forceinline__ float div_size(float m, int size) {
  if ((size & (size - 1)) == 0) return m * __int_as_float((127 - (__ffs(size) - 1)) << 23);
  return m / static_cast<float>(size);
}

// ---- addressers, same mode on both axes (sampler.cpp SSE branches) ----
__device__ __forceinline__ void wrap_point_axis(float c, int size, int& out) {
  float sz = static_cast<float>(size);
  float f = c - static_cast<float>(cvtt(c));
  f = sz * f;
  float ip = sse_floor(f);
  float m = ip + sz * 8192.0f;
  float dv = static_cast<float>(cvtt(div_size(m, size)));
  out = cvtt(m - dv * sz);
}
__device__ __forceinline__ void wrap_linear_axis(float c, int size, int& lo, int& hi, float& frac) {
  float sz = static_cast<float>(size);
  float f = c - static_cast<float>(cvtt(c));
  f = sz * f;
  f = f - 0.5f;
  float ip = sse_floor(f);
  frac = f - ip;
  float m0 = ip + 0.0f;
  m0 = m0 + sz * 8192.0f;
  float m1 = ip + 1.0f;
  m1 = m1 + sz * 8192.0f;
  lo = cvtt(m0 - static_cast<float>(cvtt(div_size(m0, size))) * sz);
  hi = cvtt(m1 - static_cast<float>(cvtt(div_size(m1, size))) * sz);
}
__device__ __forceinline__ float mirror_fold(float c, int size) {
  int sel = fast_floori(c);
  return ((sel & 1) ? static_cast<float>(1 + sel) - c : c - static_cast<float>(sel)) * static_cast<float>(size);
}
__device__ __forceinline__ void mirror_point_axis(float c, int size, int& out) {
  float ip = sse_floor(mirror_fold(c, size));
  out = cvtt(sse_min(sse_max(ip, 0.0f), static_cast<float>(size) - 1.0f));
}
__device__ __forceinline__ void mirror_linear_axis(float c, int size, int& lo, int& hi, float& frac) {
  float o = mirror_fold(c, size) - 0.5f;
  float ip = sse_floor(o);
  frac = o - ip;
  float szm1 = static_cast<float>(size) - 1.0f;
  lo = cvtt(sse_min(sse_max(ip + 0.0f, 0.0f), szm1));
  hi = cvtt(sse_min(sse_max(ip + 1.0f, 0.0f), szm1));
}
__device__ __forceinline__ void clamp_point_axis(float c, int size, int& out) {
  float sz = static_cast<float>(size);
  float f = c * sz;
  f = sse_max(f, 0.5f);
  f = sse_min(f, sz - 0.5f);
  float ip = sse_floor(f);
  out = cvtt(sse_min(sse_max(ip, 0.0f), sz - 1.0f));
}
__device__ __forceinline__ void clamp_linear_axis(float c, int size, int& lo, int& hi, float& frac) {
  float sz = static_cast<float>(size);
  float f = c * sz;
  f = sse_max(f, 0.5f);
  f = sse_min(f, sz - 0.5f);
  f = f - 0.5f;
  float ip = sse_floor(f);
  frac = f - ip;
  lo = cvtt(sse_min(sse_max(ip + 0.0f, 0.0f), sz - 1.0f));
  hi = cvtt(sse_min(sse_max(ip + 1.0f, 0.0f), sz - 1.0f));
}
__device__ __forceinline__ void border_point_axis(float c, int size, int& out) {
  float sz = static_cast<float>(size);
  float f = c * sz;
  f = sse_max(f, -0.5f);
  f = sse_min(f, sz - (-0.5f));
  int ip = cvtt(sse_floor(f));
  out = ip >= size ? -1 : ip;
}

// sampler.cpp:347-390: the bilinear footprint's indices, -1 at or beyond the size (and below 0)
__device__ __forceinline__ void border_linear_axis(float c, int size, int& lo, int& hi, float& frac) {
  float sz = static_cast<float>(size);
  float f = c * sz;
  f = sse_max(f, -0.5f);
  f = sse_min(f, sz - (-0.5f));
  f = f + (-0.5f);
  float ip = sse_floor(f);
  frac = f - ip;
  int i = cvtt(ip);
  lo = i >= size ? -1 : i;
  hi = i + 1 >= size ? -1 : i + 1;
}

// ---- 1-D addressers for mixed modes (sampler.cpp:20-25,127-130,205-210,300-304,394-412) ----
__device__ __forceinline__ float do_coordf(int mode, float coord, int size) {
  float sz = static_cast<float>(size);
  switch (mode) {
  case ADDR_WRAP: return (coord - fast_floor(coord)) * sz - 0.5f;
  case ADDR_MIRROR: {
    int sel = fast_floori(coord);
    return ((sel & 1) ? static_cast<float>(1 + sel) - coord : coord - static_cast<float>(sel)) * sz - 0.5f;
  }
  case ADDR_CLAMP: return std_clamp(coord * sz, 0.5f, sz - 0.5f) - 0.5f;
  default: return std_clamp(coord * sz, -0.5f, sz + 0.5f) - 0.5f;
  }
}
__device__ __forceinline__ int point_1d(int mode, int coord, int size) {
  switch (mode) {
  case ADDR_WRAP: return (size * 8192 + coord) % size;
  case ADDR_MIRROR:
  case ADDR_CLAMP: return std_clampi(coord, 0, size - 1);
  default: return coord >= size ? -1 : coord;
  }
}

__device__ __forceinline__ float4 filter_point(const surface_desc& s, const sampler_desc_dev& p, float x, float y) {
  int ix = 0, iy = 0;
  if (p.addr_u == p.addr_v) {
    switch (p.addr_u) {
    case ADDR_WRAP: wrap_point_axis(x, s.w, ix); wrap_point_axis(y, s.h, iy); break;
    case ADDR_MIRROR: mirror_point_axis(x, s.w, ix); mirror_point_axis(y, s.h, iy); break;
    case ADDR_CLAMP: clamp_point_axis(x, s.w, ix); clamp_point_axis(y, s.h, iy); break;
    default: border_point_axis(x, s.w, ix); border_point_axis(y, s.h, iy); break;
    }
    if (0 <= ix && ix < s.w && 0 <= iy && iy < s.h) return load_texel(s.pf, texel_addr(s, ix, iy));
    return make_float4(p.border[0], p.border[1], p.border[2], p.border[3]);
  }
  ix = point_1d(p.addr_u, fast_floori(do_coordf(p.addr_u, x, s.w) + 0.5f), s.w);
  iy = point_1d(p.addr_v, fast_floori(do_coordf(p.addr_v, y, s.h) + 0.5f), s.h);
  if (ix < 0 || iy < 0) return make_float4(p.border[0], p.border[1], p.border[2], p.border[3]);
  return load_texel(s.pf, texel_addr(s, ix, iy));
}

// Border addressing hands index -1 to the bilinear fetch (sampler.cpp:347-390, 436-447), which
// reads texel offset ((y*w + x)*samples + s)*elem in size_t arithmetic (surface.cpp:277-295): x = -1
// is the last texel of row y - 1, while y = -1, or x = -1 on row 0, lies before the surface.  The
// first is followed; the second (undefined in the reference) reads the nearest texel instead
// (texel_addr clamps), which gives the reference's result whenever that corner's weight is zero.
__device__ __forceinline__ float4 filter_linear(const surface_desc& s, const sampler_desc_dev& p, float x, float y) {
  coord2 c;
  if (p.addr_u == p.addr_v) {
    switch (p.addr_u) {
    case ADDR_WRAP: wrap_linear_axis(x, s.w, c.x0, c.x1, c.tx); wrap_linear_axis(y, s.h, c.y0, c.y1, c.ty); break;
    case ADDR_MIRROR: mirror_linear_axis(x, s.w, c.x0, c.x1, c.tx); mirror_linear_axis(y, s.h, c.y0, c.y1, c.ty); break;
    case ADDR_CLAMP: clamp_linear_axis(x, s.w, c.x0, c.x1, c.tx); clamp_linear_axis(y, s.h, c.y0, c.y1, c.ty); break;
    default: border_linear_axis(x, s.w, c.x0, c.x1, c.tx); border_linear_axis(y, s.h, c.y0, c.y1, c.ty); break;
    }
  } else {
    float ox = do_coordf(p.addr_u, x, s.w);
    int ipx = fast_floori(ox);
    c.x0 = point_1d(p.addr_u, ipx, s.w);
    c.x1 = point_1d(p.addr_u, ipx + 1, s.w);
    c.tx = ox - static_cast<float>(ipx);
    float oy = do_coordf(p.addr_v, y, s.h);
    int ipy = fast_floori(oy);
    c.y0 = point_1d(p.addr_v, ipy, s.h);
    c.y1 = point_1d(p.addr_v, ipy + 1, s.h);
    c.ty = oy - static_cast<float>(ipy);
  }
  if (p.addr_u == ADDR_BORDER || p.addr_v == ADDR_BORDER) {
    int xs[4] = {c.x0, c.x1, c.x0, c.x1}, ys[4] = {c.y0, c.y0, c.y1, c.y1};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (xs[k] < 0) {
        xs[k] = s.w - 1;
        ys[k] -= 1;
      }
    }
    return lerp_2d4(s, xs, ys, c.tx, c.ty);
  }
  return lerp_2d(s, c);
}

// Bilinear wrap/wrap RGBA8 footprint: the common case of sample_surface, without the mode and
// format dispatch (same operations as wrap_linear_axis + lerp_2d).
__device__ __forceinline__ float4 bilinear_wrap_rgba8(const surface_desc& s, float x, float y) {
  coord2 c;
  wrap_linear_axis(x, s.w, c.x0, c.x1, c.tx);
  wrap_linear_axis(y, s.h, c.y0, c.y1, c.ty);
  // single-sample RGBA8 level: texel (x, y) is word y*w + x (levels stay far below 2^32 bytes)
  const uint32_t* t = reinterpret_cast<const uint32_t*>(s.data);
  const uint32_t x0 = static_cast<uint32_t>(min(max(c.x0, 0), s.w - 1)), x1 = static_cast<uint32_t>(min(max(c.x1, 0), s.w - 1));
  const uint32_t r0 = static_cast<uint32_t>(min(max(c.y0, 0), s.h - 1)) * static_cast<uint32_t>(s.w);
  const uint32_t r1 = static_cast<uint32_t>(min(max(c.y1, 0), s.h - 1)) * static_cast<uint32_t>(s.w);
  const uint32_t u0 = t[r0 + x0], u1 = t[r0 + x1], u2 = t[r1 + x0], u3 = t[r1 + x1];
  float r[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int sh = q * 8;
    float c0 = static_cast<float>((u0 >> sh) & 0xFF), c1 = static_cast<float>((u1 >> sh) & 0xFF);
    float c2 = static_cast<float>((u2 >> sh) & 0xFF), c3 = static_cast<float>((u3 >> sh) & 0xFF);
    float c01 = c0 + (c1 - c0) * c.tx;
    float c23 = c2 + (c3 - c2) * c.tx;
    r[q] = (c01 + (c23 - c01) * c.ty) * (1.0f / 255);
  }
  return make_float4(r[0], r[1], r[2], r[3]);
}

__device__ __forceinline__ float4 sample_surface(const sampler_desc_dev& p, int level, int filter, float x, float y) {
  const surface_desc& s = p.levels[level];
  return filter == FILTER_LINEAR ? filter_linear(s, p, x, y) : filter_point(s, p, x, y);
}

__device__ __forceinline__ float len2(float x, float y) {
  float t = 0.0f;
  t += x * x;
  t += y * y;
  return sqrtf(t);
}

// sampler.cpp:521-603
__device__ inline float calc_lod(const sampler_desc_dev& p, float dxu, float dxv, float dyu, float dyv, float bias) {
  float sw = static_cast<float>(p.size[0]), sh = static_cast<float>(p.size[1]), sd = static_cast<float>(p.size[2]);
  if (p.mip_qual == MIP_LO) {
    float m0 = std_max(fabsf(dxu), fabsf(dyu)) * sw;
    float m1 = std_max(fabsf(dxv), fabsf(dyv)) * sh;
    float m2 = std_max(fabsf(0.0f), fabsf(0.0f)) * sd;
    float rho = std_max(std_max(m0, m1), m2);
    return fast_log2(rho) + bias;
  }
  float x0 = dxu * sw, x1 = dxv * sh, x2 = 0.0f * sd;
  float y0 = dyu * sw, y1 = dyv * sh, y2 = 0.0f * sd;
  float rho;
  if (p.mip_qual == MIP_HI) {
    float A = x0 * x0 + y0 * y0;
    float B = -2.0f * (x0 * x1 + y0 * y1);
    float C = x1 * x1 + y1 * y1;
    float F = A * C - B * B * 0.25f;
    float invF = 1.0f / F;
    A *= invF;
    B *= invF;
    C *= invF;
    float AsubC = A - C;
    float R = sqrtf(AsubC * AsubC + B * B);
    rho = sqrtf(2.0f / (A + C - R));
  } else {
    float tx = 0.0f;
    tx += x0 * x0;
    tx += x1 * x1;
    tx += x2 * x2;
    float ty = 0.0f;
    ty += y0 * y0;
    ty += y1 * y1;
    ty += y2 * y2;
    rho = std_max(sqrtf(tx), sqrtf(ty));
  }
  if (rho == 0.0f) rho = 0.000001f;
  return fast_log2(rho) + bias;
}

struct aniso_info {
  float lod, probe_count, weight_D, du, dv;
};

// sampler.cpp:875-958
__device__ inline aniso_info calc_aniso(const sampler_desc_dev& p, float dxu, float dxv, float dyu, float dyv, float bias) {
  float sw = static_cast<float>(p.size[0]), sh = static_cast<float>(p.size[1]);
  float x0 = dxu * sw, x1 = dxv * sh, y0 = dyu * sw, y1 = dyv * sh;
  // z and w of the texture-space derivatives are exact zeros (0 * depth, 0 * 0)
  float ddx_len = len2(x0, x1), ddy_len = len2(y0, y1);
  float diag0 = len2(x0 - y0, x1 - y1), diag1 = len2(x0 + y0, x1 + y1);
  float minor = std_min(std_min(diag0, diag1), std_min(ddx_len, ddy_len));
  if (minor == 0.0f) minor = 0.000001f;
  float la0, la1, long_len;
  if (ddx_len > ddy_len) {
    long_len = ddx_len;
    la0 = x0;
    la1 = x1;
  } else {
    long_len = ddy_len;
    la0 = y0;
    la1 = y1;
  }
  float probe = (2.0f * long_len / minor) - 1.0f;
  float rounded = fast_round(probe);
  rounded = std_min(static_cast<float>(p.max_aniso), rounded);
  if (rounded < probe) minor = 2.0f * long_len / (rounded + 1.0f);
  aniso_info a;
  a.lod = fast_log2(minor) + bias;
  a.probe_count = rounded;
  if (rounded <= 1.0f) {
    a.du = a.dv = 0.0f;
    a.weight_D = 0.0f;
  } else {
    float r = minor / long_len;
    float k = 1.0f / (rounded - 1.0f);
    float d0 = ((la0 * (1.0f - r)) * 2.0f) * k;
    float d1 = ((la1 * (1.0f - r)) * 2.0f) * k;
    float zz = ((0.0f * (1.0f - r)) * 2.0f) * k;  // z and w lanes of delta_uv
    float lsq = 0.0f;
    lsq += d0 * d0;
    lsq += d1 * d1;
    lsq += zz * zz;
    lsq += zz * zz;
    a.weight_D = static_cast<float>(256) * lsq * 0.25f / (long_len * long_len);
    a.du = div_size(d0, static_cast<int>(p.size[0]));
    a.dv = div_size(d1, static_cast<int>(p.size[1]));
  }
  return a;
}

// anisotropic probes whose texel loads are in flight together
#ifndef SRX_ANISO_BATCH
#define SRX_ANISO_BATCH 4
#endif
// The probes of the anisotropic filter on one bilinear/wrap RGBA8 level (sampler.cpp:730-760),
// SRX_ANISO_BATCH per step: their texel loads are issued together (every address is clamped into
// the level, so probes past the count load harmlessly and are masked out), then the probes are
// accumulated in order, which keeps the reference's summation order.  (x, y) is the first probe.
__device__ __forceinline__ float4 aniso_probes_rgba8(const surface_desc& lv, float x, float y, int n, const aniso_info* af) {
  float cr = 0.0f, cg = 0.0f, cb = 0.0f, ca = 0.0f, w_sum = 0.0f;
  for (int k = 0; k < n; k += SRX_ANISO_BATCH) {
    float4 c[SRX_ANISO_BATCH];
#pragma unroll
    for (int j = 0; j < SRX_ANISO_BATCH; ++j) {
      c[j] = bilinear_wrap_rgba8(lv, x, y);
      x += af->du;
      y += af->dv;
    }
#pragma unroll
    for (int j = 0; j < SRX_ANISO_BATCH; ++j) {
      const int i = 2 * (k + j) - n + 1;
      const bool live = k + j < n;
      float wf = static_cast<float>(i * i) * af->weight_D;
      int wi = (wf >= 0.0f) ? static_cast<int>(sse_min(wf, 255.0f)) : 0;
      float w = c_ewa[live ? wi : 0];
      cr = live ? cr + c[j].x * w : cr;
      cg = live ? cg + c[j].y * w : cg;
      cb = live ? cb + c[j].z * w : cb;
      ca = live ? ca + c[j].w * w : ca;
      w_sum = live ? w_sum + w : w_sum;
    }
  }
  float inv = 1 / w_sum;
  return make_float4(cr * inv, cg * inv, cb * inv, ca * inv);
}

// sampler.cpp:680-764
__device__ __noinline__ float4 sample_impl(const sampler_desc_dev& p, float u, float v, float miplevel, const aniso_info* af) {
  bool is_mag = (p.mip_filter == FILTER_POINT) ? (miplevel < 0.5f) : (miplevel < 0.0f);
  if (is_mag) return sample_surface(p, p.max_lod, p.mag_filter, u, v);
  if (p.mip_filter == FILTER_POINT) {
    int ml = fast_floori(0.5 + static_cast<double>(miplevel));
    ml = std_clampi(ml, p.max_lod, p.min_lod);
    return sample_surface(p, ml, p.min_filter, u, v);
  }
  if (p.mip_filter == FILTER_LINEAR || af == nullptr) {
    int lo = fast_floori(miplevel);
    int hi = lo + 1;
    float frac = miplevel - static_cast<float>(lo);
    int lo_c = std_clampi(lo, p.max_lod, p.min_lod);
    int hi_c = std_clampi(hi, p.max_lod, p.min_lod);
    float4 c0 = sample_surface(p, lo_c, p.min_filter, u, v);
    float4 c1 = sample_surface(p, hi_c, p.min_filter, u, v);
    return lerp4(c0, c1, frac);
  }
  float start = -0.5f * (af->probe_count - 1.0f);
  float x = u + af->du * start;
  float y = v + af->dv * start;
  int lo = fast_roundi(miplevel);
  // size_t clamp of the reference: negative levels wrap to huge values and clamp to min_lod
  uint64_t lo_s = static_cast<uint64_t>(static_cast<int64_t>(lo));
  int lo_c = lo_s < static_cast<uint64_t>(p.max_lod) ? p.max_lod
             : (static_cast<uint64_t>(p.min_lod) < lo_s ? p.min_lod : static_cast<int>(lo_s));
  float cr = 0.0f, cg = 0.0f, cb = 0.0f, ca = 0.0f, w_sum = 0.0f;
  int n = static_cast<int>(af->probe_count);
  const surface_desc& lv = p.levels[lo_c];
  const bool fast = p.min_filter == FILTER_LINEAR && p.addr_u == ADDR_WRAP && p.addr_v == ADDR_WRAP && lv.pf == PF_RGBA8 &&
                    lv.samples == 1;
  if (fast) return aniso_probes_rgba8(lv, x, y, n, af);
  for (int i = -n + 1; i < n; i += 2) {
    float4 c0 = sample_surface(p, lo_c, p.min_filter, x, y);
    float wf = static_cast<float>(i * i) * af->weight_D;
    int wi = (wf >= 0.0f) ? static_cast<int>(sse_min(wf, 255.0f)) : 0;  // the reference reads past its table there
    float w = c_ewa[wi];
    cr = cr + c0.x * w;
    cg = cg + c0.y * w;
    cb = cb + c0.z * w;
    ca = ca + c0.w * w;
    w_sum += w;
    x += af->du;
    y += af->dv;
  }
  float inv = 1 / w_sum;
  return make_float4(cr * inv, cg * inv, cb * inv, ca * inv);
}

// sample_impl at lod 0 (tex2dlod with a zero lod): a point mip filter magnifies there (0 < 0.5), and a
// point mag filter is one texel of the largest level -- inline; anything else takes sample_impl.
__device__ __forceinline__ float4 sample_lod0(const sampler_desc_dev& p, float u, float v) {
  if (p.mip_filter == FILTER_POINT && p.mag_filter == FILTER_POINT) return filter_point(p.levels[p.max_lod], p, u, v);
  return sample_impl(p, u, v, 0.0f, nullptr);
}

// sample_2d_grad for a sampler with fast_aniso set: the same decisions as sample_2d_grad and
// sample_impl (magnification below lod 0, else the probes on the rounded level), inline and compact.
__device__ __forceinline__ float4 sample_2d_grad_fast(const sampler_desc_dev& p, float u, float v, float dxu, float dxv,
                                                      float dyu, float dyv, float bias) {
  const aniso_info af = calc_aniso(p, dxu, dxv, dyu, dyv, bias);
  const float miplevel = af.lod;
  if (miplevel < 0.0f) return bilinear_wrap_rgba8(p.levels[p.max_lod], u, v);
  const float start = -0.5f * (af.probe_count - 1.0f);
  const int lo = fast_roundi(miplevel);
  const uint64_t lo_s = static_cast<uint64_t>(static_cast<int64_t>(lo));
  const int lo_c = lo_s < static_cast<uint64_t>(p.max_lod) ? p.max_lod
                   : (static_cast<uint64_t>(p.min_lod) < lo_s ? p.min_lod : static_cast<int>(lo_s));
  return aniso_probes_rgba8(p.levels[lo_c], u + af.du * start, v + af.dv * start, static_cast<int>(af.probe_count), &af);
}

// sample_2d_grad (sampler.cpp:854-873)
__device__ __noinline__ float4 sample_2d_grad(const sampler_desc_dev& p, float u, float v, float dxu, float dxv, float dyu,
                                        float dyv, float bias) {
  if (p.mip_filter == FILTER_ANISO && p.max_aniso > 1) {
    aniso_info af = calc_aniso(p, dxu, dxv, dyu, dyv, bias);
    return sample_impl(p, u, v, af.lod, &af);
  }
  float lod = calc_lod(p, dxu, dxv, dyu, dyv, bias);
  aniso_info dummy{};
  return sample_impl(p, u, v, lod, &dummy);
}

}  // namespace srx
