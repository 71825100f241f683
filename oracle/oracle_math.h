// ORACLE -- test infrastructure only. Nothing in the product links, includes or calls this.
//
// Scalar restatement of the arithmetic helpers, colour conversions and texture sampler of
// the Salvia software renderer.  Each function follows the reference's x86 SSE path
// operation by operation (every SSE lane op is one IEEE single-precision op), because that
// is the code the reference runs on x86-64.  Compile with -ffp-contract=off and without
// -ffast-math so that no operation is fused or reassociated.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace srxo {

// ---------------------------------------------------------------- scalar helpers

// eflib/include/eflib/math/math.h:130-144 -- fast_roundi(d) = (int)floor(d + 0.5)
inline int fast_roundi(double d) { return static_cast<int>(std::floor(d + 0.5)); }
// math.h:146-158; DOUBLE_MAGIC_ROUND_EPS = (0.5f - 1.5e-8) evaluated in double.
inline int fast_ceili(double d) {
  const double eps = (0.5f - 1.5e-8);
  return fast_roundi(d + eps);
}
inline int fast_floori(double d) {
  const double eps = (0.5f - 1.5e-8);
  return fast_roundi(d - eps);
}

inline float bits_f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline uint32_t f_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

// math.h:89-105 -- bit-trick log2 with a quadratic mantissa polynomial.
inline float fast_log2(float val) {
  int32_t x = static_cast<int32_t>(f_bits(val));
  int log_2 = ((x >> 23) & 255) - 128;
  x &= ~(255 << 23);
  x += 127 << 23;
  float f = bits_f(static_cast<uint32_t>(x));
  f = ((-1.0f / 3) * f + 2) * f - 2.0f / 3;
  return f + static_cast<float>(log_2);
}

// math.h:111-124 -- round by adding and subtracting 2^23 carrying val's sign.
inline float fast_round(float val) {
  uint32_t n = f_bits(val);
  float bias = bits_f(((23u + 127u) << 23) + (n & 0x80000000u));
  float f = val + bias;
  f = f - bias;
  return f;
}
// math.h:126-134
inline float fast_floor(float val) {
  float f = fast_round(val);
  return (f > val) ? f - 1 : f;
}

// eflib::equal<float> (math.h:21-26): |a - b| <= FLT_EPSILON
inline bool equal_eps(float a, float b) { return std::fabs(a - b) <= 1.1920928955078125e-07f; }

// x86 cvttps2dq / cvttss2si: truncation; NaN and out-of-range give INT_MIN.
inline int32_t cvtt(float f) {
  if (!(f >= -2147483648.0f && f < 2147483648.0f)) return INT32_MIN;
  return static_cast<int32_t>(f);
}
// cvtps2dq under the default MXCSR: round half to even.
inline int32_t cvt_rne(float f) {
  if (!(f >= -2147483648.0f && f < 2147483648.0f)) return INT32_MIN;
  return static_cast<int32_t>(std::nearbyint(f));
}
// maxps / minps: the second operand is returned when either is NaN.
inline float sse_max(float a, float b) { return a > b ? a : b; }
inline float sse_min(float a, float b) { return a < b ? a : b; }
// std::min / std::max / std::clamp as the reference calls them (eflib uses std::clamp).
inline float std_min(float a, float b) { return (b < a) ? b : a; }
inline float std_max(float a, float b) { return (a < b) ? b : a; }
template <class T> inline T std_clamp(T v, T lo, T hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }

// ---------------------------------------------------------------- vec4

struct vec4 {
  float v[4];
  float& operator[](int i) { return v[i]; }
  float operator[](int i) const { return v[i]; }
};
inline vec4 v4(float x, float y, float z, float w) { return vec4{{x, y, z, w}}; }
inline vec4 operator+(vec4 a, vec4 b) { return v4(a[0] + b[0], a[1] + b[1], a[2] + b[2], a[3] + b[3]); }
inline vec4 operator-(vec4 a, vec4 b) { return v4(a[0] - b[0], a[1] - b[1], a[2] - b[2], a[3] - b[3]); }
inline vec4 operator*(vec4 a, float s) { return v4(a[0] * s, a[1] * s, a[2] * s, a[3] * s); }
inline vec4 mul(vec4 a, vec4 b) { return v4(a[0] * b[0], a[1] * b[1], a[2] * b[2], a[3] * b[3]); }
// vector_generic.h:105-113: total = 0; total += d_i*d_i in order
inline float length_sqr2(float x, float y) { float t = 0.0f; t += x * x; t += y * y; return t; }
inline float length_sqr3(float x, float y, float z) { float t = 0.0f; t += x * x; t += y * y; t += z * z; return t; }
inline float length_sqr4(vec4 a) { float t = 0.0f; for (int i = 0; i < 4; ++i) t += a[i] * a[i]; return t; }

// ---------------------------------------------------------------- colours
// salvia/include/salvia/common/colors.h; pixel format codes colors_convertors.h:40-47
enum { PF_RGBA32F = 0, PF_RGB32F = 1, PF_BGRA8 = 2, PF_RGBA8 = 3, PF_R32F = 4, PF_RG32F = 5, PF_R32I = 6 };
inline int pf_size(int pf) {
  switch (pf) {
  case PF_RGBA32F: return 16;
  case PF_RGB32F: return 12;
  case PF_BGRA8: return 4;
  case PF_RGBA8: return 4;
  case PF_R32F: return 4;
  case PF_RG32F: return 8;
  case PF_R32I: return 4;
  }
  return 0;
}

// colors.h:182-193: clamp(x*255, 0, 255) with maxps/minps, then cvtps2dq (half-even).
inline uint8_t unorm8_store(float x) {
  float m = x * 255.0f;
  m = sse_max(m, 0.0f);
  m = sse_min(m, 255.0f);
  return static_cast<uint8_t>(cvt_rne(m));
}

// to_rgba32f of one stored texel (colors.h to_rgba32f members)
inline vec4 load_texel(int pf, const uint8_t* p) {
  const float inv_255 = 1.0f / 255;
  float f[4];
  switch (pf) {
  case PF_RGBA32F: std::memcpy(f, p, 16); return v4(f[0], f[1], f[2], f[3]);
  case PF_RGB32F: std::memcpy(f, p, 12); return v4(f[0], f[1], f[2], 1.0f);
  case PF_RGBA8: return v4(p[0] * inv_255, p[1] * inv_255, p[2] * inv_255, p[3] * inv_255);
  case PF_BGRA8: return v4(p[2] * inv_255, p[1] * inv_255, p[0] * inv_255, p[3] * inv_255);
  case PF_R32F: std::memcpy(f, p, 4); return v4(f[0], 0.0f, 0.0f, 0.0f);
  case PF_RG32F: std::memcpy(f, p, 8); return v4(f[0], f[1], 0.0f, 0.0f);
  case PF_R32I: { int32_t i; std::memcpy(&i, p, 4); return v4(static_cast<float>(i), 0.0f, 0.0f, 0.0f); }
  }
  return v4(0, 0, 0, 0);
}

// assign from color_rgba32f (colors.h assign members)
inline void store_texel(int pf, uint8_t* p, vec4 c) {
  switch (pf) {
  case PF_RGBA32F: std::memcpy(p, c.v, 16); return;
  case PF_RGB32F: std::memcpy(p, c.v, 12); return;
  case PF_RGBA8: p[0] = unorm8_store(c[0]); p[1] = unorm8_store(c[1]); p[2] = unorm8_store(c[2]); p[3] = unorm8_store(c[3]); return;
  case PF_BGRA8: p[0] = unorm8_store(c[2]); p[1] = unorm8_store(c[1]); p[2] = unorm8_store(c[0]); p[3] = unorm8_store(c[3]); return;
  case PF_R32F: std::memcpy(p, c.v, 4); return;
  case PF_RG32F: std::memcpy(p, c.v, 8); return;
  case PF_R32I: { int32_t i = static_cast<int32_t>(c[0] + 0.5f); std::memcpy(p, &i, 4); return; }
  }
}

// colors.h:303-314 -- lerp(c0, c1, t) = c0 + (c1 - c0) * t per channel (SSE order)
inline vec4 lerp4(vec4 a, vec4 b, float t) {
  return v4(a[0] + (b[0] - a[0]) * t, a[1] + (b[1] - a[1]) * t, a[2] + (b[2] - a[2]) * t, a[3] + (b[3] - a[3]) * t);
}

// Bilinear lerp of four stored texels (colors.h:372-488, per format).
inline vec4 lerp_2d(int pf, const uint8_t* a0, const uint8_t* a1, const uint8_t* a2, const uint8_t* a3, float tx, float ty) {
  switch (pf) {
  case PF_RGBA8:
  case PF_BGRA8: {
    // bytes widened to int then cvtepi32_ps; lerps in 0..255; * (1/255) at the end
    auto ld = [pf](const uint8_t* p) {
      return pf == PF_RGBA8 ? v4(p[0], p[1], p[2], p[3]) : v4(p[2], p[1], p[0], p[3]);
    };
    vec4 c0 = ld(a0), c1 = ld(a1), c2 = ld(a2), c3 = ld(a3);
    vec4 c01 = lerp4(c0, c1, tx);
    vec4 c23 = lerp4(c2, c3, tx);
    vec4 r = lerp4(c01, c23, ty);
    return r * (1.0f / 255);
  }
  case PF_RGBA32F: {
    vec4 c0 = load_texel(pf, a0), c1 = load_texel(pf, a1), c2 = load_texel(pf, a2), c3 = load_texel(pf, a3);
    vec4 c01 = lerp4(c0, c1, tx);
    vec4 c23 = lerp4(c2, c3, tx);
    return lerp4(c01, c23, ty);
  }
  case PF_R32F: {
    float c0, c1, c2, c3;
    std::memcpy(&c0, a0, 4); std::memcpy(&c1, a1, 4); std::memcpy(&c2, a2, 4); std::memcpy(&c3, a3, 4);
    float c01 = c0 + (c1 - c0) * tx;
    float c23 = c2 + (c3 - c2) * tx;
    return v4(c01 + (c23 - c01) * ty, 0.0f, 0.0f, 0.0f);
  }
  case PF_RG32F: {
    // colors.h:469-478: the .g channel uses the .r base (reference quirk, kept)
    float c0[2], c1[2], c2[2], c3[2];
    std::memcpy(c0, a0, 8); std::memcpy(c1, a1, 8); std::memcpy(c2, a2, 8); std::memcpy(c3, a3, 8);
    float c01r = c0[0] + (c1[0] - c0[0]) * tx, c01g = c0[0] + (c1[1] - c0[1]) * tx;
    float c23r = c2[0] + (c3[0] - c2[0]) * tx, c23g = c2[0] + (c3[1] - c2[1]) * tx;
    return v4(c01r + (c23r - c01r) * ty, c01r + (c23g - c01g) * ty, 0.0f, 0.0f);
  }
  case PF_RGB32F: {
    float c0[3], c1[3], c2[3], c3[3];
    std::memcpy(c0, a0, 12); std::memcpy(c1, a1, 12); std::memcpy(c2, a2, 12); std::memcpy(c3, a3, 12);
    float c01[3] = {c0[0] + (c1[0] - c0[0]) * tx, c0[0] + (c1[1] - c0[1]) * tx, c0[0] + (c1[2] - c0[2]) * tx};
    float c23[3] = {c2[0] + (c3[0] - c2[0]) * tx, c2[0] + (c3[1] - c2[1]) * tx, c2[0] + (c3[2] - c2[2]) * tx};
    return v4(c01[0] + (c23[0] - c01[0]) * ty, c01[0] + (c23[1] - c01[1]) * ty, c01[0] + (c23[2] - c01[2]) * ty, 1.0f);
  }
  case PF_R32I: {
    int32_t c0, c1, c2, c3;
    std::memcpy(&c0, a0, 4); std::memcpy(&c1, a1, 4); std::memcpy(&c2, a2, 4); std::memcpy(&c3, a3, 4);
    float c01 = c0 + (c1 - c0) * tx;
    float c23 = c2 + (c3 - c2) * tx;
    return v4(c01 + (c23 - c01) * ty, 0.0f, 0.0f, 0.0f);
  }
  }
  return v4(0, 0, 0, 0);
}

}  // namespace srxo
