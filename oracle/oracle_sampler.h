// ORACLE -- test infrastructure only. Nothing in the product links, includes or calls this.
//
// Texture sampler of salvia/src/resource/sampler.cpp restated on scalars, lane by lane of the
// reference's SSE addressers.  Every function cites the lines it follows.
#pragma once

#include "oracle_math.h"

#include <vector>

namespace srxo {

enum { ADDR_WRAP = 0, ADDR_MIRROR = 1, ADDR_CLAMP = 2, ADDR_BORDER = 3 };
enum { FILTER_POINT = 0, FILTER_LINEAR = 1, FILTER_ANISO = 2 };
enum { MIP_LO = 0, MIP_MI = 1, MIP_HI = 2 };

struct surf_view {
  int w, h, samples, pf;
  const uint8_t* data;
  // coordinates clamped to the surface: identical for every in-range coordinate, and keeps the
  // NaN/huge-coordinate cases (out-of-bounds reads in the reference) inside the allocation
  const uint8_t* addr(int x, int y, int s) const {
    x = x < 0 ? 0 : (x >= w ? w - 1 : x);
    y = y < 0 ? 0 : (y >= h ? h - 1 : y);
    return data + ((static_cast<size_t>(y) * w + x) * samples + s) * pf_size(pf);
  }
  vec4 texel(int x, int y, int s) const { return load_texel(pf, addr(x, y, s)); }
  vec4 texel_lerp(int x0, int y0, int x1, int y1, float tx, float ty, int s) const {
    return lerp_2d(pf, addr(x0, y0, s), addr(x1, y0, s), addr(x0, y1, s), addr(x1, y1, s), tx, ty);
  }
};

struct sampler_params {
  int min_filter, mag_filter, mip_filter, mip_qual, addr_u, addr_v;
  float lod_bias;
  uint32_t max_aniso;
  vec4 border;
};

struct texture_view {
  std::vector<surf_view> levels;  // index = mip level
  int max_lod, min_lod;           // texture.h:21-22 (max_lod = finest level)
  uint32_t size[3];               // level-0 width, height, depth (=1)
};

// ---- floor used by every SSE addresser: trunc via cvtt, minus one where it rounded up
inline float sse_floor(float f) {
  float ip = static_cast<float>(cvtt(f));
  if (ip > f) ip = ip - 1.0f;
  return ip;
}

// ---- 2-D addressers, same mode on u and v (sampler.cpp:19-392, SSE branches) -------------
// Returns texel coordinates; for point sampling lo is the texel.
struct coord2 { int x0, y0, x1, y1; float tx, ty; };

inline void wrap_point_2d(float u, float v, int W, int H, int& ix, int& iy) {
  // sampler.cpp:27-44
  float c[2] = {u, v};
  float sz[2] = {static_cast<float>(W), static_cast<float>(H)};
  int out[2];
  for (int k = 0; k < 2; ++k) {
    float f = c[k] - static_cast<float>(cvtt(c[k]));
    f = sz[k] * f;
    float ip = sse_floor(f);
    float m = ip + sz[k] * 8192.0f;
    float dv = static_cast<float>(cvtt(m / sz[k]));
    out[k] = cvtt(m - dv * sz[k]);
  }
  ix = out[0];
  iy = out[1];
}

inline coord2 wrap_linear_2d(float u, float v, int W, int H) {
  // sampler.cpp:64-96
  float c[2] = {u, v};
  float sz[2] = {static_cast<float>(W), static_cast<float>(H)};
  float ipart[2], frac[2];
  for (int k = 0; k < 2; ++k) {
    float f = c[k] - static_cast<float>(cvtt(c[k]));
    f = sz[k] * f;
    f = f - 0.5f;
    ipart[k] = sse_floor(f);
    frac[k] = f - ipart[k];
  }
  // lanes (ipx, ipy, ipx+1, ipy+1) + size*8192, then a float modulo
  int r[4];
  for (int l = 0; l < 4; ++l) {
    float s = sz[l & 1];
    float m = ipart[l & 1] + ((l >= 2) ? 1.0f : 0.0f);
    m = m + s * 8192.0f;
    float dv = static_cast<float>(cvtt(m / s));
    r[l] = cvtt(m - dv * s);
  }
  return coord2{r[0], r[1], r[2], r[3], frac[0], frac[1]};
}

inline float mirror_fold(float c, int size) {
  // sampler.cpp:133-139 (without the -0.5 of the linear form)
  int sel = fast_floori(c);
  return ((sel & 1) ? static_cast<float>(1 + sel) - c : c - static_cast<float>(sel)) * static_cast<float>(size);
}

inline void mirror_point_2d(float u, float v, int W, int H, int& ix, int& iy) {
  // sampler.cpp:131-155
  float o[2] = {mirror_fold(u, W), mirror_fold(v, H)};
  float szm1[2] = {static_cast<float>(W) - 1.0f, static_cast<float>(H) - 1.0f};
  int out[2];
  for (int k = 0; k < 2; ++k) {
    float ip = sse_floor(o[k]);
    out[k] = cvtt(sse_min(sse_max(ip, 0.0f), szm1[k]));
  }
  ix = out[0];
  iy = out[1];
}

inline coord2 mirror_linear_2d(float u, float v, int W, int H) {
  // sampler.cpp:168-202
  float o[2] = {mirror_fold(u, W) - 0.5f, mirror_fold(v, H) - 0.5f};
  float szm1[2] = {static_cast<float>(W) - 1.0f, static_cast<float>(H) - 1.0f};
  float ipart[2], frac[2];
  for (int k = 0; k < 2; ++k) {
    ipart[k] = sse_floor(o[k]);
    frac[k] = o[k] - ipart[k];
  }
  int r[4];
  for (int l = 0; l < 4; ++l) {
    float m = ipart[l & 1] + ((l >= 2) ? 1.0f : 0.0f);
    r[l] = cvtt(sse_min(sse_max(m, 0.0f), szm1[l & 1]));
  }
  return coord2{r[0], r[1], r[2], r[3], frac[0], frac[1]};
}

inline void clamp_point_2d(float u, float v, int W, int H, int& ix, int& iy) {
  // sampler.cpp:234-255
  float c[2] = {u, v};
  float sz[2] = {static_cast<float>(W), static_cast<float>(H)};
  int out[2];
  for (int k = 0; k < 2; ++k) {
    float f = c[k] * sz[k];
    f = sse_max(f, 0.5f);
    f = sse_min(f, sz[k] - 0.5f);
    float ip = sse_floor(f);
    out[k] = cvtt(sse_min(sse_max(ip, 0.0f), sz[k] - 1.0f));
  }
  ix = out[0];
  iy = out[1];
}

inline coord2 clamp_linear_2d(float u, float v, int W, int H) {
  // sampler.cpp:269-296
  float c[2] = {u, v};
  float sz[2] = {static_cast<float>(W), static_cast<float>(H)};
  float ipart[2], frac[2];
  for (int k = 0; k < 2; ++k) {
    float f = c[k] * sz[k];
    f = sse_max(f, 0.5f);
    f = sse_min(f, sz[k] - 0.5f);
    f = f - 0.5f;
    ipart[k] = sse_floor(f);
    frac[k] = f - ipart[k];
  }
  int r[4];
  for (int l = 0; l < 4; ++l) {
    float m = ipart[l & 1] + ((l >= 2) ? 1.0f : 0.0f);
    r[l] = cvtt(sse_min(sse_max(m, 0.0f), sz[l & 1] - 1.0f));
  }
  return coord2{r[0], r[1], r[2], r[3], frac[0], frac[1]};
}

inline void border_point_2d(float u, float v, int W, int H, int& ix, int& iy) {
  // sampler.cpp:322-345: texel index, or -1 at or beyond the size
  float c[2] = {u, v};
  int szi[2] = {W, H};
  int out[2];
  for (int k = 0; k < 2; ++k) {
    float sz = static_cast<float>(szi[k]);
    float f = c[k] * sz;
    f = sse_max(f, -0.5f);
    f = sse_min(f, sz - (-0.5f));
    int ip = cvtt(sse_floor(f));
    out[k] = ip >= szi[k] ? -1 : ip;
  }
  ix = out[0];
  iy = out[1];
}

// sampler.cpp:347-390: texel indices of the bilinear footprint, -1 at or beyond the size (and
// -1 below 0, where the coordinate is clamped to -0.5 texel)
inline coord2 border_linear_2d(float u, float v, int W, int H) {
  coord2 c;
  float cc[2] = {u, v};
  int szi[2] = {W, H};
  int lo[2], hi[2];
  float fr[2];
  for (int k = 0; k < 2; ++k) {
    float sz = static_cast<float>(szi[k]);
    float f = cc[k] * sz;
    f = sse_max(f, -0.5f);
    f = sse_min(f, sz - (-0.5f));
    f = f + (-0.5f);
    float ip = sse_floor(f);
    fr[k] = f - ip;
    int i = cvtt(ip);
    lo[k] = i >= szi[k] ? -1 : i;
    hi[k] = i + 1 >= szi[k] ? -1 : i + 1;
  }
  c.x0 = lo[0];
  c.x1 = hi[0];
  c.tx = fr[0];
  c.y0 = lo[1];
  c.y1 = hi[1];
  c.ty = fr[1];
  return c;
}

// ---- 1-D addressers for mixed modes (sampler.cpp:20-25,127-130,205-210,300-304,394-412) ----
inline float do_coordf(int mode, float coord, int size) {
  switch (mode) {
  case ADDR_WRAP: return (coord - fast_floor(coord)) * static_cast<float>(size) - 0.5f;
  case ADDR_MIRROR: {
    int sel = fast_floori(coord);
    return ((sel & 1) ? static_cast<float>(1 + sel) - coord : coord - static_cast<float>(sel)) * static_cast<float>(size) - 0.5f;
  }
  case ADDR_CLAMP: return std_clamp(coord * static_cast<float>(size), 0.5f, static_cast<float>(size) - 0.5f) - 0.5f;
  case ADDR_BORDER: return std_clamp(coord * static_cast<float>(size), -0.5f, static_cast<float>(size) + 0.5f) - 0.5f;
  }
  return 0.0f;
}
inline int do_coordi_point_1d(int mode, int coord, int size) {
  switch (mode) {
  case ADDR_WRAP: return (size * 8192 + coord) % size;
  case ADDR_MIRROR:
  case ADDR_CLAMP: return std_clamp(coord, 0, size - 1);
  case ADDR_BORDER: return coord >= size ? -1 : coord;
  }
  return 0;
}

// ---- surface filters (sampler.cpp:414-481) ----
inline vec4 filter_point(const surf_view& s, const sampler_params& p, float x, float y, int sample) {
  if (p.addr_u == p.addr_v) {
    int ix = 0, iy = 0;
    switch (p.addr_u) {
    case ADDR_WRAP: wrap_point_2d(x, y, s.w, s.h, ix, iy); break;
    case ADDR_MIRROR: mirror_point_2d(x, y, s.w, s.h, ix, iy); break;
    case ADDR_CLAMP: clamp_point_2d(x, y, s.w, s.h, ix, iy); break;
    case ADDR_BORDER: border_point_2d(x, y, s.w, s.h, ix, iy); break;
    }
    if (0 <= ix && ix < s.w && 0 <= iy && iy < s.h) return s.texel(ix, iy, sample);
    return p.border;
  }
  float ox = do_coordf(p.addr_u, x, s.w);
  int ix = do_coordi_point_1d(p.addr_u, fast_floori(ox + 0.5f), s.w);
  float oy = do_coordf(p.addr_v, y, s.h);
  int iy = do_coordi_point_1d(p.addr_v, fast_floori(oy + 0.5f), s.h);
  if (ix < 0 || iy < 0) return p.border;
  return s.texel(ix, iy, sample);
}

// Border addressing hands index -1 to the bilinear fetch (sampler.cpp:347-390, 436-447), which
// reads texel offset ((y*w + x)*samples + s)*elem in size_t arithmetic (surface.cpp:277-295): x = -1
// is the last texel of row y - 1, while y = -1, or x = -1 on row 0, lies before the surface.  The
// first is followed; the second (undefined in the reference) reads the nearest texel instead,
// which gives the reference's result whenever that corner's weight is zero.
inline vec4 filter_linear(const surf_view& s, const sampler_params& p, float x, float y, int sample) {
  coord2 c;
  if (p.addr_u == p.addr_v) {
    switch (p.addr_u) {
    case ADDR_WRAP: c = wrap_linear_2d(x, y, s.w, s.h); break;
    case ADDR_MIRROR: c = mirror_linear_2d(x, y, s.w, s.h); break;
    case ADDR_CLAMP: c = clamp_linear_2d(x, y, s.w, s.h); break;
    default: c = border_linear_2d(x, y, s.w, s.h); break;
    }
  } else {
    float ox = do_coordf(p.addr_u, x, s.w);
    int ipx = fast_floori(ox);
    c.x0 = do_coordi_point_1d(p.addr_u, ipx, s.w);
    c.x1 = do_coordi_point_1d(p.addr_u, ipx + 1, s.w);
    c.tx = ox - static_cast<float>(ipx);
    float oy = do_coordf(p.addr_v, y, s.h);
    int ipy = fast_floori(oy);
    c.y0 = do_coordi_point_1d(p.addr_v, ipy, s.h);
    c.y1 = do_coordi_point_1d(p.addr_v, ipy + 1, s.h);
    c.ty = oy - static_cast<float>(ipy);
  }
  if (p.addr_u == ADDR_BORDER || p.addr_v == ADDR_BORDER) {
    // the four corners, each moved to the texel its linear offset names
    int xs[4] = {c.x0, c.x1, c.x0, c.x1}, ys[4] = {c.y0, c.y0, c.y1, c.y1};
    for (int k = 0; k < 4; ++k) {
      if (xs[k] < 0) {
        xs[k] = s.w - 1;
        ys[k] -= 1;
      }
    }
    return lerp_2d(s.pf, s.addr(xs[0], ys[0], sample), s.addr(xs[1], ys[1], sample), s.addr(xs[2], ys[2], sample),
                   s.addr(xs[3], ys[3], sample), c.tx, c.ty);
  }
  return s.texel_lerp(c.x0, c.y0, c.x1, c.y1, c.tx, c.ty, sample);
}

inline vec4 sample_surface(const surf_view& s, const sampler_params& p, int filter, float x, float y) {
  return filter == FILTER_LINEAR ? filter_linear(s, p, x, y, 0) : filter_point(s, p, x, y, 0);
}

// ---- LOD (sampler.cpp:521-603) ----
inline float calc_lod(const texture_view& t, const sampler_params& p, vec4 ddx, vec4 ddy, float bias) {
  vec4 size = v4(static_cast<float>(t.size[0]), static_cast<float>(t.size[1]), static_cast<float>(t.size[2]), 0.0f);
  if (p.mip_qual == MIP_LO) {
    vec4 maxd = v4(std_max(std::fabs(ddx[0]), std::fabs(ddy[0])), std_max(std::fabs(ddx[1]), std::fabs(ddy[1])),
                   std_max(std::fabs(ddx[2]), std::fabs(ddy[2])), 0.0f);
    maxd = mul(maxd, size);
    float rho = std_max(std_max(maxd[0], maxd[1]), maxd[2]);
    return fast_log2(rho) + bias;
  }
  vec4 dxt = mul(ddx, size), dyt = mul(ddy, size);
  float rho;
  if (p.mip_qual == MIP_HI) {
    float A = dxt[0] * dxt[0] + dyt[0] * dyt[0];
    float B = -2.0f * (dxt[0] * dxt[1] + dyt[0] * dyt[1]);
    float C = dxt[1] * dxt[1] + dyt[1] * dyt[1];
    float F = A * C - B * B * 0.25f;
    float invF = 1.0f / F;
    A *= invF;
    B *= invF;
    C *= invF;
    float AsubC = A - C;
    float R = std::sqrt(AsubC * AsubC + B * B);
    rho = std::sqrt(2.0f / (A + C - R));
  } else {
    float lx = std::sqrt(length_sqr3(dxt[0], dxt[1], dxt[2]));
    float ly = std::sqrt(length_sqr3(dyt[0], dyt[1], dyt[2]));
    rho = std_max(lx, ly);
  }
  if (rho == 0.0f) rho = 0.000001f;
  return fast_log2(rho) + bias;
}

struct aniso_info {
  float lod, probe_count, weight_D;
  vec4 delta_uv;
};

// sampler.cpp:875-958
inline aniso_info calc_aniso(const texture_view& t, const sampler_params& p, vec4 ddx, vec4 ddy, float bias) {
  vec4 size = v4(static_cast<float>(t.size[0]), static_cast<float>(t.size[1]), static_cast<float>(t.size[2]), 0.0f);
  vec4 dxt = mul(ddx, size), dyt = mul(ddy, size);
  float ddx_len = std::sqrt(length_sqr2(dxt[0], dxt[1]));
  float ddy_len = std::sqrt(length_sqr2(dyt[0], dyt[1]));
  vec4 d0 = dxt - dyt, d1 = dxt + dyt;
  float diag0 = std::sqrt(length_sqr2(d0[0], d0[1]));
  float diag1 = std::sqrt(length_sqr2(d1[0], d1[1]));
  float minor = std_min(std_min(diag0, diag1), std_min(ddx_len, ddy_len));
  if (minor == 0.0f) minor = 0.000001f;
  vec4 long_axis;
  float long_len;
  if (ddx_len > ddy_len) {
    long_len = ddx_len;
    long_axis = dxt;
  } else {
    long_len = ddy_len;
    long_axis = dyt;
  }
  float probe = (2.0f * long_len / minor) - 1.0f;
  float rounded = fast_round(probe);
  rounded = std_min(static_cast<float>(p.max_aniso), rounded);
  if (rounded < probe) minor = 2.0f * long_len / (rounded + 1.0f);
  aniso_info a;
  a.lod = fast_log2(minor) + bias;
  a.probe_count = rounded;
  if (rounded <= 1.0f) {
    a.delta_uv = v4(0, 0, 0, 0);
    a.weight_D = 0.0f;
  } else {
    float r = minor / long_len;
    a.delta_uv = ((long_axis * (1.0f - r)) * 2.0f) * (1.0f / (rounded - 1.0f));
    a.weight_D = static_cast<float>(256) * length_sqr4(a.delta_uv) * 0.25f / (long_len * long_len);
    a.delta_uv[0] /= size[0];
    a.delta_uv[1] /= size[1];
  }
  return a;
}

// EWA weight i of the 256-entry table (sampler.cpp:618-677); defined next to the pipeline.
float ewa(int i);

// sampler.cpp:680-764 (2-D textures)
inline vec4 sample_impl(const texture_view& t, const sampler_params& p, float u, float v, float miplevel, const aniso_info* af) {
  bool is_mag = (p.mip_filter == FILTER_POINT) ? (miplevel < 0.5f) : (miplevel < 0.0f);
  if (is_mag) return sample_surface(t.levels[t.max_lod], p, p.mag_filter, u, v);
  if (p.mip_filter == FILTER_POINT) {
    int ml = fast_floori(0.5 + miplevel);
    ml = std_clamp(ml, t.max_lod, t.min_lod);
    return sample_surface(t.levels[ml], p, p.min_filter, u, v);
  }
  // The reference dereferences a null anisotropic_info when the cpp tex2d path meets an
  // anisotropic sampler (sampler.cpp:730-737); both implementations take the linear-mip path there.
  if (p.mip_filter == FILTER_LINEAR || af == nullptr) {
    int lo = fast_floori(miplevel);
    int hi = lo + 1;
    float frac = miplevel - static_cast<float>(lo);
    int lo_c = std_clamp(lo, t.max_lod, t.min_lod);
    int hi_c = std_clamp(hi, t.max_lod, t.min_lod);
    vec4 c0 = sample_surface(t.levels[lo_c], p, p.min_filter, u, v);
    vec4 c1 = sample_surface(t.levels[hi_c], p, p.min_filter, u, v);
    return lerp4(c0, c1, frac);
  }
  // anisotropic: EWA-weighted probes on one rounded level
  float start = -0.5f * (af->probe_count - 1.0f);
  float x = u + af->delta_uv[0] * start;
  float y = v + af->delta_uv[1] * start;
  int lo = fast_roundi(miplevel);
  size_t lo_s = static_cast<size_t>(lo);  // size_t clamp as in the reference
  size_t lo_c = lo_s < static_cast<size_t>(t.max_lod) ? static_cast<size_t>(t.max_lod)
                : (static_cast<size_t>(t.min_lod) < lo_s ? static_cast<size_t>(t.min_lod) : lo_s);
  vec4 color = v4(0, 0, 0, 0);
  float w_sum = 0.0f;
  int n = static_cast<int>(af->probe_count);
  for (int i = -n + 1; i < n; i += 2) {
    vec4 c0 = sample_surface(t.levels[lo_c], p, p.min_filter, x, y);
    float wf = static_cast<float>(i * i) * af->weight_D;
    int wi = (wf >= 0.0f) ? static_cast<int>(sse_min(wf, 255.0f)) : 0;  // the reference reads past its table there
    float w = ewa(wi);
    color = color + c0 * w;
    w_sum += w;
    x += af->delta_uv[0];
    y += af->delta_uv[1];
  }
  float inv = 1 / w_sum;  // vector_generic.h:75-79: /= multiplies by the reciprocal
  return color * inv;
}

// sampler.cpp:854-873
inline vec4 sample_2d_grad(const texture_view& t, const sampler_params& p, float u, float v, float dxu, float dxv,
                           float dyu, float dyv, float lod_bias) {
  vec4 ddx = v4(dxu, dxv, 0, 0), ddy = v4(dyu, dyv, 0, 0);
  aniso_info af{};
  float lod;
  if (p.mip_filter == FILTER_ANISO && p.max_aniso > 1) {
    af = calc_aniso(t, p, ddx, ddy, lod_bias);
    lod = af.lod;
  } else {
    lod = calc_lod(t, p, ddx, ddy, lod_bias);
  }
  return sample_impl(t, p, u, v, lod, &af);
}

// sampler.cpp:834-852 (calc_lod_2d) for the cpp pixel shader's tex2d; aniso is undefined there
inline float calc_lod_2d(const texture_view& t, const sampler_params& p, float dxu, float dxv, float dyu, float dyv) {
  vec4 ddx = v4(dxu, dxv, 0, 0), ddy = v4(dyu, dyv, 0, 0);
  if (p.mip_filter == FILTER_ANISO && p.max_aniso > 1) return calc_aniso(t, p, ddx, ddy, 0.0f).lod;
  return calc_lod(t, p, ddx, ddy, 0.0f);
}

}  // namespace srxo
