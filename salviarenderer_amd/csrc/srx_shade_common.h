// Device helpers of the pixel-shading stage shared by the precompiled kernels (srx_kernels.hip)
// and the SASL shading kernels compiled at run time (srx_sasl_jit.cpp): the per-triangle
// interpolation data, the quad interpolation chain and the sample addressing of a surface.
#pragma once

#include "srx_device.h"
#include "srx_sampler.h"

namespace srx {

constexpr uint32_t NO_WINNER = 0xFFFFFFFFu;

// Setup records of clipped triangle id = 3 * primitive + fan index live at slot
// fan index * prim_count + primitive: the fan-0 records of consecutive primitives (all of them when
// nothing is clipped) are contiguous, so setup writes them densely.
__device__ __forceinline__ size_t tri_slot(const draw_params& P, uint32_t id) {
  const uint32_t p = id / 3u;
  return static_cast<size_t>(id - 3u * p) * P.prim_count + p;
}
// Per-triangle interpolation data: v0 (the reordered first vertex) and the screen-space
// derivatives of every register (compute_derivative_n, shader.cpp:414-450).
template <int NA>
struct attr_setup {
  float4 v0[NA + 1], dx[NA + 1], dy[NA + 1];
};

template <int NA>
__device__ __forceinline__ void load_attr(const draw_params& P, const float4* __restrict__ tri_attr, uint32_t id,
                                          attr_setup<NA>& a) {
  const float4* src = tri_attr + tri_slot(P, id) * 3 * (NA + 1);
#pragma unroll
  for (int r = 0; r <= NA; ++r) {
    a.v0[r] = src[r];
    a.dx[r] = src[NA + 1 + r];
    a.dy[r] = src[2 * (NA + 1) + r];
  }
}

// Interpolated registers of single pixels of a 2x2 quad, computed only where read: pixel i's value
// follows the quad's chain (p00 = v0 + (ddx*dx + ddy*dy), +ddx for the right column, then +ddy for
// the bottom row), perspective-divided by its own 1/w, so every pixel gets the value a whole-quad
// interpolation would give it.
template <int NA>
struct quad_lazy {
  const attr_setup<NA>& t;
  const draw_params& P;
  float dx, dy;
  __device__ __forceinline__ quad_lazy(const attr_setup<NA>& at, const draw_params& p, int px0, int py0)
      : t(at), P(p), dx((0.5f + static_cast<float>(px0)) - at.v0[0].x), dy((0.5f + static_cast<float>(py0)) - at.v0[0].y) {}
  __device__ __forceinline__ static float chain(float v, float a, float b, float dx, float dy, int i) {
    float c = v + ((a * dx) + (b * dy));
    c = (i & 1) ? c + a : c;
    return (i & 2) ? c + b : c;
  }
  __device__ __forceinline__ float inv_w(int i) const {
    return 1.0f / chain(t.v0[0].w, t.dx[0].w, t.dy[0].w, dx, dy, i);
  }
  // register r (1..NA) of pixel i, components [0, nc)
  __device__ __forceinline__ float4 reg(int r, int i, float iw, int nc = 4) const {
    const float4 v = t.v0[r], a = t.dx[r], b = t.dy[r];
    const uint32_t mod = P.attr_mod[r - 1];
    float4 o;
    if (mod & AM_NOINTERPOLATION) {
      o = v;
    } else {
      o.x = chain(v.x, a.x, b.x, dx, dy, i);
      o.y = nc > 1 ? chain(v.y, a.y, b.y, dx, dy, i) : 0.0f;
      o.z = nc > 2 ? chain(v.z, a.z, b.z, dx, dy, i) : 0.0f;
      o.w = nc > 3 ? chain(v.w, a.w, b.w, dx, dy, i) : 0.0f;
    }
    if (!(mod & AM_NOPERSPECTIVE)) o = make_float4(o.x * iw, o.y * iw, o.z * iw, o.w * iw);
    return o;
  }
};

__device__ __forceinline__ uint8_t* sample_addr(const surface_desc& s, int x, int y, int smp) {
  return s.data + ((static_cast<size_t>(y) * s.w + x) * s.samples + smp) * pf_size(s.pf);
}


}  // namespace srx
