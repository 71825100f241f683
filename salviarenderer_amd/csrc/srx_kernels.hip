// HIP kernels of the Salvia rasterization pipeline for MI355X (gfx950).
//
// Draw pipeline (one HIP stream per renderer):
//   vs_kernel       vertex fetch + vertex shader, one thread per vertex    (stream_assembler.cpp,
//                   default_vertex_cache.cpp: the shader runs once per vertex instead of per cache miss)
//   setup_kernel    index fetch, near/far trivial accept, cull, winding, viewport transform and
//                   triangle setup, one thread per primitive                (clipper.cpp, rasterizer.cpp:864-945)
//   clip_kernel     Sutherland-Hodgman near/far clipping for the few primitives that need it; the
//                   draw's record total from setup's per-block totals
//   emit_kernel     per-triangle GPU-tile records written in clipped-primitive order (offsets from
//                   the per-block totals and a workgroup scan)
//   sort            by GPU tile: a counting sort (order-independent draws) or a stable 8-bit LSD
//                   radix sort -> per-tile lists in primitive order (the reference's per-tile
//                   std::sort, rasterizer.cpp:947-994)
//   raster_kernel   one workgroup per GPU tile, one thread per 2x2 quad; walks its tile's list in
//                   order and evaluates the reference's 64->16->4->sample decision chain per quad
//                   (rasterizer.cpp:298-773), early/late depth-stencil, pixel shader and blending
//                   (framebuffer.cpp).  Opaque early-Z draws defer the pixel shader to one pass over
//                   the final per-sample winners.
#include "srx_kernels.h"
#include "srx_sampler.h"
#include "srx_shade_common.h"
#include "srx_sasl_exec.h"

#include <type_traits>

namespace srx {

__constant__ float c_ewa[256];

enum { VS_PASSTHROUGH_COLOR = 1, VS_WVP_NORMAL_UV = 2, VS_WVP_COLOR = 3, VS_WVP_COLOR_LIGHT = 4, VS_COLORIZED = 5,
       VS_MESH = 6, VS_SSM_DRAW = 7, VS_CUP = 8, VS_BOX = 9 };
enum { PS_GOURAUD = 1, PS_LAMBERT_TEX = 2, PS_TEX_CPP = 3, PS_CONSTANT = 4, PS_SHADOW_PCF = 5, PS_COLORIZED = 6,
       PS_MESH = 7, PS_SSM = 8, PS_CUP = 9, PS_BOX = 10 };
enum { BLEND_OFF = 0, BLEND_ON = 1, BLEND_DISABLED = 2 };
enum { TOPO_TRIANGLE_LIST = 2, TOPO_TRIANGLE_STRIP = 4 };
enum { FMT_RGBA32F = 2, FMT_RGBA32I = 4, FMT_RGB32F = 6, FMT_RG32F = 16, FMT_R32F = 41 };
enum { CULL_NONE = 0, CULL_FRONT = 1, CULL_BACK = 2 };
enum { SOP_KEEP = 1, SOP_ZERO, SOP_REPLACE, SOP_INCR_SAT, SOP_DECR_SAT, SOP_INVERT, SOP_INCR_WRAP, SOP_DECR_WRAP };

template <int NA>
struct vtx {
  float4 r[NA + 1];
};

// ============================================================== vertex stage
// eflib::transform(out, v, m): out[i] = dot_prod4(v, column i) (eflib/src/math.cpp:142-154)
__device__ __forceinline__ float4 transform_rowvec(float4 v, const float* m) {
  float o[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) o[i] = v.x * m[0 * 4 + i] + v.y * m[1 * 4 + i] + v.z * m[2 * 4 + i] + v.w * m[3 * 4 + i];
  return make_float4(o[0], o[1], o[2], o[3]);
}

// eflib::transform33(out, v, m): out[i] = dot_prod3(v.xyz, column i.xyz), out.w = 0 (math.cpp:184-197)
__device__ __forceinline__ float4 transform33_rowvec(float4 v, const float* m) {
  float o[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) o[i] = v.x * m[0 * 4 + i] + v.y * m[1 * 4 + i] + v.z * m[2 * 4 + i];
  return make_float4(o[0], o[1], o[2], 0.0f);
}
// a float4 constant minus a register, component by component
__device__ __forceinline__ float4 sub4(const float* c, float4 v) {
  return make_float4(c[0] - v.x, c[1] - v.y, c[2] - v.z, c[3] - v.w);
}

__device__ __forceinline__ float4 fetch_elem(const input_elem_dev& e, uint32_t v) {
  uint32_t need;
  switch (e.fmt) {
  case FMT_R32F: need = 4; break;
  case FMT_RG32F: need = 8; break;
  case FMT_RGB32F: need = 12; break;
  default: need = 16; break;
  }
  uint64_t off = static_cast<uint64_t>(e.stride) * v;
  if (off + need > e.avail) return make_float4(0, 0, 0, 0);  // the reference reads out of bounds
  const float* f = reinterpret_cast<const float*>(e.base + off);
  // get_vec4 (stream_assembler.cpp:26-44)
  switch (e.fmt) {
  case FMT_R32F: return make_float4(f[0], 0.0f, 0.0f, e.default_w);
  case FMT_RG32F: return make_float4(f[0], f[1], 0.0f, e.default_w);
  case FMT_RGB32F: return make_float4(f[0], f[1], f[2], e.default_w);
  default: return make_float4(f[0], f[1], f[2], f[3]);
  }
}

// a[i] for i < n, else 0, loaded without a branch (the index is clamped): guarded loads in an
// unrolled loop would each sit in a branch of its own and wait for its data before the next issues.
// n must be at least 1.
__device__ __forceinline__ uint32_t load_or_zero(const uint32_t* __restrict__ a, uint32_t i, uint32_t n) {
  const uint32_t x = a[min(i, n - 1u)];
  return i < n ? x : 0u;
}

// Coalesced stores of a wave's per-lane blocks of W float4 (lane i's block at dst + i*W): the lanes
// in `mask` put their blocks in the wave's LDS staging area, then the wave stores the span with
// consecutive lanes on consecutive float4, skipping the blocks of lanes not in `mask` (their memory
// is left as it was).  Blocks wider than the staging area are staged a group of lanes at a time.
constexpr int WAVE_STAGE = 320;  // float4 of LDS staging per wave
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
template <int W>
__device__ __forceinline__ void wave_store_blocks(float4* __restrict__ dst, const float4 (&blk)[W], uint64_t mask,
                                                  float4* __restrict__ stage) {
  constexpr int LG = WAVE_STAGE / W < 64 ? WAVE_STAGE / W : 64;  // lanes per pass
  const int lane = __lane_id();
  for (int g = 0; g < 64; g += LG) {
    const uint64_t pm = (mask >> g) & (LG >= 64 ? ~0ull : ((1ull << LG) - 1ull));
    if (!pm) continue;  // wave-uniform
    if (lane >= g && lane < g + LG && ((mask >> lane) & 1ull)) {
#pragma unroll
      for (int k = 0; k < W; ++k) stage[(lane - g) * W + k] = blk[k];
    }
    wave_lds_sync();
    const int n = min(LG, 64 - g) * W;
    for (int j = lane; j < n; j += 64)
      if ((pm >> (j / W)) & 1ull) dst[static_cast<size_t>(g) * W + j] = stage[j];
    wave_lds_sync();  // the stores have read the staging area before the next group writes it
  }
}
// The same with a run-time block width w <= MAX_REGS (lane i's block: blk[0..w)).
__device__ __forceinline__ void wave_store_blocks_n(float4* __restrict__ dst, const float4 (&blk)[MAX_REGS], int w,
                                                    uint64_t mask, float4* __restrict__ stage) {
  const int lg = min(WAVE_STAGE / w, 64);
  const int lane = __lane_id();
  for (int g = 0; g < 64; g += lg) {
    const uint64_t pm = (mask >> g) & (lg >= 64 ? ~0ull : ((1ull << lg) - 1ull));
    if (!pm) continue;  // wave-uniform
    if (lane >= g && lane < g + lg && ((mask >> lane) & 1ull)) {
#pragma unroll
      for (int k = 0; k < MAX_REGS; ++k)
        if (k < w) stage[(lane - g) * w + k] = blk[k];
    }
    wave_lds_sync();
    const int n = min(lg, 64 - g) * w;
    for (int j = lane; j < n; j += 64)
      if ((pm >> (j / w)) & 1ull) dst[static_cast<size_t>(g) * w + j] = stage[j];
    wave_lds_sync();
  }
}

// Kernel bodies take their workgroup index `bx` (and grid size where they stride) as arguments:
// the single-draw kernels pass blockIdx.x, the batched kernels a job's local workgroup index.
__device__ __forceinline__ void vs_body(const draw_params& P, float4* __restrict__ vbuf, uint32_t bx) {
  uint32_t v = bx * blockDim.x + threadIdx.x;
  const bool live = v <= P.nverts;  // slot nverts holds the output for an all-out-of-bounds fetch
  float4 in[MAX_INPUTS];
#pragma unroll
  for (int i = 0; i < MAX_INPUTS; ++i) in[i] = make_float4(0, 0, 0, 0);
  if (v < P.nverts)
    for (int i = 0; i < P.nin; ++i) {
      const float4 x = fetch_elem(P.in[i], v);
#pragma unroll
      for (int r = 0; r < MAX_INPUTS; ++r)  // register-resident: no dynamically indexed array
        if (P.in[i].reg == r) in[r] = x;
    }
  float4 out[MAX_REGS];
#pragma unroll
  for (int r = 0; r < MAX_REGS; ++r) out[r] = make_float4(0, 0, 0, 0);
  const float* k = P.vs_consts;
  switch (P.vs_id) {
  case VS_PASSTHROUGH_COLOR:
    out[0] = in[0];
    out[1] = in[1];
    break;
  case VS_WVP_COLOR:
    out[0] = transform_rowvec(in[0], k);
    out[1] = in[1];
    break;
  case VS_WVP_COLOR_LIGHT:  // light-space position for the shadow lookup (StandardShadowMap.cpp)
    out[0] = transform_rowvec(in[0], k);
    out[1] = in[1];
    out[2] = transform_rowvec(in[0], k + 16);
    break;
  case VS_COLORIZED:  // ColorizedTriangle.cpp:44-52
    out[0] = transform_rowvec(in[0], k);
    out[1] = in[1];
    out[2] = sub4(k + 16, in[0]);
    out[3] = sub4(k + 20, in[0]);
    out[4] = sub4(k + 24, in[0]);
    break;
  case VS_MESH: {  // ComplexMesh.cpp:202-211
    const float4 es = transform_rowvec(in[0], k);
    out[0] = transform_rowvec(es, k + 16);
    out[1] = make_float4(k[32] - es.x, k[33] - es.y, k[34] - es.z, 1.0f);
    out[2] = make_float4(k[36] - es.x, k[37] - es.y, k[38] - es.z, 1.0f);
    out[3] = transform33_rowvec(in[1], k);
    break;
  }
  case VS_SSM_DRAW:  // resources/ssm/Draw.savs: consts cameraPos, cameraWvp, lightPos, lightWvp
    out[0] = transform_rowvec(in[0], k + 4);
    out[1] = in[2];
    out[2] = in[1];
    out[3] = sub4(k + 20, in[0]);
    out[4] = sub4(k, in[0]);
    out[5] = transform_rowvec(in[0], k + 24);
    break;
  case VS_CUP:  // StencilMirror.cpp:47-55: consts wvp, eyePos, lightPos
    out[0] = transform_rowvec(in[0], k);
    out[1] = in[2];
    out[2] = in[1];
    out[3] = sub4(k + 20, in[0]);
    out[4] = sub4(k + 16, in[0]);
    break;
  case VS_BOX:  // TextureAndBlending.cpp:76-81
    out[0] = transform_rowvec(in[0], k);
    out[1] = in[0];
    out[2] = in[1];
    break;
  default:  // VS_WVP_NORMAL_UV
    out[0] = transform_rowvec(in[0], k);
    out[1] = in[1];
    out[2] = in[2];
    break;
  }
  // the wave's 64 consecutive vertices' outputs are stored coalesced (blockDim.x is 256)
  __shared__ float4 stage[4 * WAVE_STAGE];
  const uint64_t mask = __ballot(live);
  if (mask) {
    const int w = P.nattr + 1;
    const size_t v0 = static_cast<size_t>(bx) * blockDim.x + (threadIdx.x & ~63u);
    wave_store_blocks_n(vbuf + v0 * w, out, w, mask, stage + (threadIdx.x >> 6) * WAVE_STAGE);
  }
}
// A single draw's first launch: it also zeroes the draw's counters and ntiles per-tile record counts
// and cursors, which the launches after it add to.
__global__ void __launch_bounds__(256) vs_kernel(draw_params P, float4* __restrict__ vbuf, uint32_t* __restrict__ counters,
                                                 uint32_t* __restrict__ tiles, uint32_t ntiles) {
  const uint32_t nz = max(ntiles, static_cast<uint32_t>(CT_COUNT));
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nz; i += gridDim.x * blockDim.x) {
    if (i < static_cast<uint32_t>(CT_COUNT)) counters[i] = 0u;
    if (i < ntiles) tiles[i] = 0u;
  }
  vs_body(P, vbuf, blockIdx.x);
}

// ============================================================== primitive assembly helpers
__device__ __forceinline__ uint32_t fetch_index(const draw_params& P, uint32_t prim, int k) {
  uint32_t id;
  if (P.topo == TOPO_TRIANGLE_LIST) {
    id = prim * 3 + k;
  } else {  // strip: odd primitives swap ids 0 and 2 (index_fetcher.cpp:69-77)
    id = prim + k;
    if (prim & 1) id = (k == 0) ? prim + 2 : (k == 2 ? prim : prim + 1);
  }
  if (!P.indexed) return id + static_cast<uint32_t>(P.basevert);
  uint64_t off = static_cast<uint64_t>(P.istart_bytes) + static_cast<uint64_t>(id) * P.istride;
  uint32_t v = 0;
  if (off + P.istride <= P.ib_size)
    v = (P.istride == 2) ? *reinterpret_cast<const uint16_t*>(P.ib + off) : *reinterpret_cast<const uint32_t*>(P.ib + off);
  return v + static_cast<uint32_t>(P.basevert);
}

template <int NA>
__device__ __forceinline__ void load_vertex(const draw_params& P, const float4* vbuf, uint32_t index, vtx<NA>& out) {
  uint32_t slot = index < P.nverts ? index : P.nverts;
  const float4* src = vbuf + static_cast<size_t>(slot) * (NA + 1);
#pragma unroll
  for (int r = 0; r <= NA; ++r) out.r[r] = src[r];
}

__device__ __forceinline__ float dot4(float4 p, float4 v) { return p.x * v.x + p.y * v.y + p.z * v.z + p.w * v.w; }

// raster_state cull functions (raster_state.cpp:10-32)
__device__ __forceinline__ bool cull_fn(const draw_params& P, float area) {
  switch (P.cull_mode) {
  case CULL_FRONT: return P.front_ccw ? (area <= 0) : (area >= 0);
  case CULL_BACK: return P.front_ccw ? (area >= 0) : (area <= 0);
  default: return false;
  }
}

// viewport_transform + project_n (shader.cpp:499-511, 117-135)
template <int NA>
__device__ __forceinline__ void viewport_project(vtx<NA>& v, const draw_params& P) {
  float4 p = v.r[0];
  float invw = equal_eps(p.w, 0.0f) ? 1.0f : 1.0f / p.w;
  float px = p.x * invw, py = p.y * invw, pz = p.z * invw;
  float ox = (P.vp[0] + P.vp[2]) * 0.5f, oy = (P.vp[1] + P.vp[3]) * 0.5f;
  v.r[0].x = (P.vp[2] * 0.5f) * px + ox;
  v.r[0].y = (P.vp[3] * 0.5f) * -py + oy;
  v.r[0].z = (P.vp[5] - P.vp[4]) * pz + P.vp[4];
  v.r[0].w = invw;
#pragma unroll
  for (int a = 0; a < NA; ++a)
    if (!(P.attr_mod[a] & AM_NOPERSPECTIVE)) {
      v.r[a + 1].x *= invw;
      v.r[a + 1].y *= invw;
      v.r[a + 1].z *= invw;
      v.r[a + 1].w *= invw;
    }
}

template <int NA>
struct tri_setup {
  vtx<NA> v0, ddx, ddy;
  float e[3][3];
  float bb[4];
  bool front, valid;
};

// compute_triangle_info (rasterizer.cpp:864-945) + compute_derivative_n (shader.cpp:414-450)
// Register selects of whole vertices, one float at a time: a conditional over vertex lvalues
// would select addresses and push the vertices into scratch memory.
__device__ __forceinline__ float4 sel4(bool c, float4 a, float4 b) {
  return make_float4(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z, c ? a.w : b.w);
}
template <int NA>
__device__ __forceinline__ vtx<NA> vsel(bool c, const vtx<NA>& a, const vtx<NA>& b) {
  vtx<NA> o;
#pragma unroll
  for (int r = 0; r <= NA; ++r) o.r[r] = sel4(c, a.r[r], b.r[r]);
  return o;
}
template <int NA>
__device__ __forceinline__ vtx<NA> pick3(int i, const vtx<NA>& a, const vtx<NA>& b, const vtx<NA>& c) {
  vtx<NA> o;
#pragma unroll
  for (int r = 0; r <= NA; ++r) o.r[r] = sel4(i == 0, a.r[r], sel4(i == 1, b.r[r], c.r[r]));
  return o;
}

template <int NA>
__device__ __forceinline__ void triangle_info(const vtx<NA>& a, const vtx<NA>& b, const vtx<NA>& c, tri_setup<NA>& t) {
  const vtx<NA>* V[3] = {&a, &b, &c};  // constant indices only (after unrolling)
  double dist[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) dist[i] = static_cast<double>(fabsf(V[i]->r[0].x) + fabsf(V[i]->r[0].y));
  int r0;
  if (dist[0] < dist[1]) r0 = dist[0] < dist[2] ? 0 : 2;
  else r0 = dist[1] < dist[2] ? 1 : 2;
  int r1 = (r0 + 1) % 3, r2 = (r1 + 1) % 3;
  // cyclic reorder by selects, so the vertices stay in registers
  const vtx<NA> R0 = pick3<NA>(r0, a, b, c);
  const vtx<NA> R1 = pick3<NA>(r1, a, b, c);
  const vtx<NA> R2 = pick3<NA>(r2, a, b, c);
  float e01[NA + 1][4], e02[NA + 1][4];
#pragma unroll
  for (int r = 0; r <= NA; ++r) {
    e01[r][0] = R1.r[r].x - R0.r[r].x; e01[r][1] = R1.r[r].y - R0.r[r].y;
    e01[r][2] = R1.r[r].z - R0.r[r].z; e01[r][3] = R1.r[r].w - R0.r[r].w;
    e02[r][0] = R2.r[r].x - R0.r[r].x; e02[r][1] = R2.r[r].y - R0.r[r].y;
    e02[r][2] = R2.r[r].z - R0.r[r].z; e02[r][3] = R2.r[r].w - R0.r[r].w;
  }
  float area = e02[0][0] * e01[0][1] - e02[0][1] * e01[0][0];
  t.valid = !equal_eps(area, 0.0f);
  if (!t.valid) return;
  t.front = area > 0.0f;
  float inv_area = 1.0f / area;
  t.bb[0] = std_min(std_min(a.r[0].x, b.r[0].x), c.r[0].x);
  t.bb[1] = std_max(std_max(a.r[0].x, b.r[0].x), c.r[0].x);
  t.bb[2] = std_min(std_min(a.r[0].y, b.r[0].y), c.r[0].y);
  t.bb[3] = std_max(std_max(a.r[0].y, b.r[0].y), c.r[0].y);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float4& s = V[i]->r[0];
    const float4& e = V[(i + 1) % 3]->r[0];
    t.e[i][0] = s.y - e.y;
    t.e[i][1] = e.x - s.x;
    t.e[i][2] = e.x * s.y - e.y * s.x;
  }
  float e01x = e01[0][0], e01y = e01[0][1], e02x = e02[0][0], e02y = e02[0][1];
#pragma unroll
  for (int r = 0; r <= NA; ++r) {
    float dx[4], dy[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      float xd = e02[r][k] * e01y - e01[r][k] * e02y;
      float yd = e01[r][k] * e02x - e02[r][k] * e01x;
      dx[k] = xd * inv_area;
      dy[k] = yd * inv_area;
    }
    t.ddx.r[r] = make_float4(dx[0], dx[1], dx[2], dx[3]);
    t.ddy.r[r] = make_float4(dy[0], dy[1], dy[2], dy[3]);
  }
  t.v0 = R0;
}

// compute_front (clipper.cpp:92-101)
__device__ __forceinline__ bool compute_front(float4 p0, float4 p1, float4 p2) {
  float i0 = 1.0f / p0.w, i1 = 1.0f / p1.w, i2 = 1.0f / p2.w;
  float x0 = p0.x * i0, y0 = p0.y * i0, x1 = p1.x * i1, y1 = p1.y * i1, x2 = p2.x * i2, y2 = p2.y * i2;
  float ax = x2 - x0, ay = y2 - y0, bx = x1 - x0, by = y1 - y0;
  return (ax * by - ay * bx) > 0.0f;
}

// Screen-space vertices of clipped triangle `id` (= 3 * primitive + fan index).
template <int NA>
__device__ __forceinline__ void load_clipped_triangle(const draw_params& P, const float4* vbuf, const uint8_t* clip_flag,
                                                      const int32_t* clip_slot, const float4* clip_verts,
                                                      const uint32_t* clip_hdr, uint32_t id, vtx<NA>* out) {
  uint32_t prim = id / 3, k = id - prim * 3;
  if (!clip_flag[prim]) {
    vtx<NA> v[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) load_vertex<NA>(P, vbuf, fetch_index(P, prim, i), v[i]);
    bool front = compute_front(v[0].r[0], v[1].r[0], v[2].r[0]);
    // winding normalisation (clipper.cpp:59-65) by selects
    out[0] = v[0];
    out[1] = vsel<NA>(front, v[1], v[2]);
    out[2] = vsel<NA>(front, v[2], v[1]);
#pragma unroll
    for (int i = 0; i < 3; ++i) viewport_project<NA>(out[i], P);
    return;
  }
  int32_t slot = clip_slot[prim];
  uint32_t hdr = clip_hdr[slot];
  bool front = (hdr >> 8) & 1;
  const float4* base = clip_verts + static_cast<size_t>(slot) * 5 * (NA + 1);
  int ia = 0, ib = static_cast<int>(k) + 1, ic = static_cast<int>(k) + 2;
  if (!front) {
    int tmp = ib;
    ib = ic;
    ic = tmp;
  }
  int idx[3] = {ia, ib, ic};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int r = 0; r <= NA; ++r) out[i].r[r] = base[idx[i] * (NA + 1) + r];
}

// ============================================================== binning geometry
// Visits, in order, the GPU tiles that may hold pixels of a triangle, with the reference's
// 64x64 tile-level decision (threaded_dispatch_primitive, rasterizer.cpp:775-862).
// Binning of a triangle to GPU tiles.  The reference bins by 64x64 tile (rasterizer.cpp:947-994,
// the tile's corner test rejects or fully accepts it); a GPU tile of a partially covered 64x64 tile
// gets a record when its rows and columns meet the triangle's bounding box and one of its 16x16
// stage-16 regions passes the rasteriser's first subdivision test (subdivide_tile on the 64x64
// tile, rasterizer.cpp:662-675, the same float expressions as the visibility and raster kernels):
// a region that test rejects produces no fragment, so its record could change nothing.
// Larger triangles skip the region test: their partially covered 64x64 tiles are few next to the
// inner ones, and the test would lengthen the per-triangle counting loop.
constexpr int REGION_TEST_TILES = 16;
struct tri_bin {
  int sx, sy, ex, ey;  // 64x64 tiles [sx, ex) x [sy, ey)
  bool small;          // one 64x64 tile: the corner test is skipped (rasterizer.cpp:957)
  bool regions;        // at most REGION_TEST_TILES 64x64 tiles: GPU tiles are region-tested
  bool mx[3], my[3];
  float r2a[3], s1x[3], s1y[3];
};
__device__ __forceinline__ void bin_setup(const draw_params& P, const float* e, const float* bb, tri_bin& t) {
  const int tx = P.tiles_x, ty = P.tiles_y;
  t.sx = min(fast_floori(std_max(0.0f, bb[0]) / TILE), tx);
  t.sy = min(fast_floori(std_max(0.0f, bb[2]) / TILE), ty);
  t.ex = min(fast_ceili(std_max(0.0f, bb[1]) / TILE) + 1, tx);
  t.ey = min(fast_ceili(std_max(0.0f, bb[3]) / TILE) + 1, ty);
  t.small = (t.sx + 1 == t.ex) && (t.sy + 1 == t.ey);
  t.regions = (t.ex - t.sx) * (t.ey - t.sy) <= REGION_TEST_TILES;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float ex = e[k * 3 + 0], ey = e[k * 3 + 1];
    t.mx[k] = ex > 0;
    t.my[k] = ey > 0;
    float stx = TILE * ex, sty = TILE * ey;
    t.r2a[k] = -fabsf(stx) - fabsf(sty);
    stx *= 0.25f;
    sty *= 0.25f;
    t.s1x[k] = stx;
    t.s1y[k] = sty;
  }
}
// The GPU tiles of 64x64 tile (x, y) that get a record, bit j * sub + i for GPU tile (i, j) of it
// (row-major, the order records are written in); full: the tile is fully accepted.
__device__ __forceinline__ uint64_t tile_gmask(const draw_params& P, const float* e, const float* bb, const tri_bin& t,
                                               int x, int y, bool& full) {
  full = false;
  if (!t.small) {
    bool rej = false, acc = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float ev = e[k * 3 + 2] - ((static_cast<float>(x + t.mx[k]) * TILE) * e[k * 3 + 0] +
                                 (static_cast<float>(y + t.my[k]) * TILE) * e[k * 3 + 1]);
      rej |= (0 < ev);
      acc &= (t.r2a[k] >= ev);
    }
    if (rej) return 0;
    full = acc;
  }
  const int sub = TILE / P.gt;
  const float xmin = bb[0] - static_cast<float>(x * TILE), xmax = bb[1] - static_cast<float>(x * TILE);
  const float ymin = bb[2] - static_cast<float>(y * TILE), ymax = bb[3] - static_cast<float>(y * TILE);
  uint32_t rows = 0, cols = 0;
  for (int j = 0; j < sub; ++j) {
    if (y * sub + j >= P.gtiles_y) break;
    if (full || !(ymin >= static_cast<float>(j * P.gt + P.gt) || ymax < static_cast<float>(j * P.gt))) rows |= 1u << j;
  }
  for (int i = 0; i < sub; ++i) {
    if (x * sub + i >= P.gtiles_x) break;
    if (full || !(xmin >= static_cast<float>(i * P.gt + P.gt) || xmax < static_cast<float>(i * P.gt))) cols |= 1u << i;
  }
  if (!rows || !cols) return 0;
  uint64_t all = 0;  // every GPU tile in the bounding-box rows and columns
  for (uint32_t rm = rows; rm; rm &= rm - 1) all |= static_cast<uint64_t>(cols) << (__builtin_ctz(rm) * sub);
  if (full || !t.regions) return all;
  // stage-16 regions the first subdivision rejects (bit ry * 4 + rx)
  uint32_t rrej = 0;
  {
    const float tvx = static_cast<float>(x * TILE), tvy = static_cast<float>(y * TILE);
    float ev1[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float ex = e[k * 3], ey = e[k * 3 + 1], ez = e[k * 3 + 2];
      float part = static_cast<float>(t.mx[k] * TILE) * ex + static_cast<float>(t.my[k] * TILE) * ey;
      part *= 0.25f;
      ev1[k] = (ez - part) - ((tvx * ex) + (tvy * ey));
    }
    // regions inside the bounding-box rows and columns only
    const int r0 = (__builtin_ctz(rows) * P.gt) >> 4, r1 = ((31 - __builtin_clz(rows)) * P.gt + P.gt - 1) >> 4;
    const int c0 = (__builtin_ctz(cols) * P.gt) >> 4, c1 = ((31 - __builtin_clz(cols)) * P.gt + P.gt - 1) >> 4;
    for (int ry = r0; ry <= r1; ++ry)
      for (int rx = c0; rx <= c1; ++rx) {
        bool rej = false;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const float r = (t.s1x[k] * static_cast<float>(rx)) + (t.s1y[k] * static_cast<float>(ry));
          rej |= r < ev1[k];
        }
        if (rej) rrej |= 1u << (ry * 4 + rx);
      }
  }
  if (!rrej) return all;
  if (P.gt == 16) return all & ~static_cast<uint64_t>(rrej);  // GPU tile (i, j) is region (i, j)
  uint64_t m = 0;
  for (uint32_t rm = rows; rm; rm &= rm - 1) {
    const int j = __builtin_ctz(rm);
    const int ry0 = (j * P.gt) >> 4, ry1 = (j * P.gt + P.gt - 1) >> 4;
    for (uint32_t cm = cols; cm; cm &= cm - 1) {
      const int i = __builtin_ctz(cm);
      const int rx0 = (i * P.gt) >> 4, rx1 = (i * P.gt + P.gt - 1) >> 4;
      bool any = false;  // a region of the GPU tile the first subdivision keeps
      for (int ry = ry0; ry <= ry1; ++ry)
        for (int rx = rx0; rx <= rx1; ++rx) any |= !((rrej >> (ry * 4 + rx)) & 1u);
      if (any) m |= 1ull << (j * sub + i);
    }
  }
  return m;
}

template <class F>
__device__ __forceinline__ void for_each_gtile(const draw_params& P, const float* e, const float* bb, F&& f) {
  tri_bin t;
  bin_setup(P, e, bb, t);
  const int sub = TILE / P.gt;
  for (int y = t.sy; y < t.ey; ++y)
    for (int x = t.sx; x < t.ex; ++x) {
      bool full;
      for (uint64_t m = tile_gmask(P, e, bb, t, x, y, full); m; m &= m - 1) {
        const int b = __builtin_ctzll(m);
        f(static_cast<uint32_t>((y * sub + b / sub) * P.gtiles_x + x * sub + b % sub), full);
      }
    }
}

// 64x64 tiles in a triangle's bounding box (as bin_setup clamps it)
__device__ __forceinline__ int bin_span(const draw_params& P, const float* bb) {
  const int sx = min(fast_floori(std_max(0.0f, bb[0]) / TILE), P.tiles_x);
  const int sy = min(fast_floori(std_max(0.0f, bb[2]) / TILE), P.tiles_y);
  const int ex = min(fast_ceili(std_max(0.0f, bb[1]) / TILE) + 1, P.tiles_x);
  const int ey = min(fast_ceili(std_max(0.0f, bb[3]) / TILE) + 1, P.tiles_y);
  return (ex > sx && ey > sy) ? (ex - sx) * (ey - sy) : 0;
}
// Triangles over more than this many 64x64 tiles are counted by a whole wave (clip launch), not by
// setup's one lane.
constexpr int COUNT_WIDE = 32;

// count_gtiles by a whole wave (every lane calls it with the same triangle): lanes take the 64x64
// tiles, the wave sums their GPU tiles.
__device__ __forceinline__ uint32_t count_gtiles_wave(const draw_params& P, const float* e, const float* bb) {
  tri_bin t;
  bin_setup(P, e, bb, t);
  const int w = t.ex - t.sx;
  const int R = (t.ex > t.sx && t.ey > t.sy) ? w * (t.ey - t.sy) : 0;
  uint32_t n = 0;
  for (int tt = __lane_id(); tt < R; tt += 64) {
    bool full;
    n += static_cast<uint32_t>(__popcll(tile_gmask(P, e, bb, t, t.sx + tt % w, t.sy + tt / w, full)));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) n += static_cast<uint32_t>(__shfl_xor(static_cast<int>(n), off, 64));
  return n;
}

// The number of GPU tiles for_each_gtile visits.
__device__ __forceinline__ uint32_t count_gtiles(const draw_params& P, const float* e, const float* bb) {
  tri_bin t;
  bin_setup(P, e, bb, t);
  uint32_t n = 0;
  for (int y = t.sy; y < t.ey; ++y)
    for (int x = t.sx; x < t.ex; ++x) {
      bool full;
      n += static_cast<uint32_t>(__popcll(tile_gmask(P, e, bb, t, x, y, full)));
    }
  return n;
}

// count_gtiles that also keeps the first four GPU tiles for_each_gtile visits, packed as
// tile << 1 | full (setup's pairs, which spare the emission the walk of triangles with at most four
// records).
constexpr uint32_t NO_PAIRS = 0xFFFFFFFFu;  // no tile is that large
__device__ __forceinline__ uint32_t count_gtiles_pairs(const draw_params& P, const float* e, const float* bb, uint4& k) {
  tri_bin t;
  bin_setup(P, e, bb, t);
  const int sub = TILE / P.gt;
  uint32_t n = 0;
  k = make_uint4(NO_PAIRS, NO_PAIRS, NO_PAIRS, NO_PAIRS);
  for (int y = t.sy; y < t.ey; ++y)
    for (int x = t.sx; x < t.ex; ++x) {
      bool full;
      uint64_t m = tile_gmask(P, e, bb, t, x, y, full);
      const uint32_t c = static_cast<uint32_t>(__popcll(m));
      for (uint32_t i = n; m && i < 4; m &= m - 1, ++i) {
        const int b = __builtin_ctzll(m);
        const uint32_t g = static_cast<uint32_t>((y * sub + b / sub) * P.gtiles_x + x * sub + b % sub);
        const uint32_t v = (g << 1) | (full ? 1u : 0u);
        if (i == 0) k.x = v;
        else if (i == 1) k.y = v;
        else if (i == 2) k.z = v;
        else k.w = v;
      }
      n += c;
    }
  return n;
}

template <int NA>
__device__ __forceinline__ void tri_rec_words(const tri_setup<NA>& t, float4 (&rec)[5]) {
  rec[0] = make_float4(t.e[0][0], t.e[0][1], t.e[0][2], t.e[1][0]);
  rec[1] = make_float4(t.e[1][1], t.e[1][2], t.e[2][0], t.e[2][1]);
  rec[2] = make_float4(t.e[2][2], t.bb[0], t.bb[1], t.bb[2]);
  rec[3] = make_float4(t.bb[3], t.v0.r[0].x, t.v0.r[0].y, t.v0.r[0].z);
  rec[4] = make_float4(t.ddx.r[0].z, t.ddy.r[0].z, t.front ? 1.0f : 0.0f, 0.0f);
}
template <int NA>
__device__ __forceinline__ void tri_attr_words(const tri_setup<NA>& t, float4 (&a)[3 * (NA + 1)]) {
#pragma unroll
  for (int r = 0; r <= NA; ++r) {
    a[r] = t.v0.r[r];
    a[NA + 1 + r] = t.ddx.r[r];
    a[2 * (NA + 1) + r] = t.ddy.r[r];
  }
}

template <int NA>
__device__ __forceinline__ void write_tri_rec(float4* rec, const tri_setup<NA>& t) {
  float4 w[5];
  tri_rec_words<NA>(t, w);
#pragma unroll
  for (int k = 0; k < 5; ++k) rec[k] = w[k];
}

template <int NA>
__device__ __forceinline__ void write_tri_attr(float4* a, const tri_setup<NA>& t) {
  float4 w[3 * (NA + 1)];
  tri_attr_words<NA>(t, w);
#pragma unroll
  for (int k = 0; k < 3 * (NA + 1); ++k) a[k] = w[k];
}

__device__ __forceinline__ void read_tri_rec(const float4* rec, float* e, float* bb) {
  float4 a = rec[0], b = rec[1], c = rec[2], d = rec[3];
  e[0] = a.x; e[1] = a.y; e[2] = a.z; e[3] = a.w; e[4] = b.x; e[5] = b.y; e[6] = b.z; e[7] = b.w; e[8] = c.x;
  bb[0] = c.y; bb[1] = c.z; bb[2] = c.w; bb[3] = d.x;
}

// Sums v over the workgroup and stores the total in partial[blockIdx.x] (no global atomics: many
// same-address device-scope atomics from every XCD serialise badly).  All threads must call it.
__device__ __forceinline__ void block_partial(unsigned long long* __restrict__ partial, unsigned long long v) {
  __shared__ unsigned long long wsum[16];
  unsigned long long s = v;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  const int wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) wsum[wave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < nw; ++w) t += wsum[w];
    partial[blockIdx.x] = t;
  }
}

// Sum of v over the workgroup, returned to thread 0.  All threads must call it.
__device__ __forceinline__ uint32_t block_sum_u32(uint32_t v) {
  __shared__ uint32_t wsum[16];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), off, 64));
  const int wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) wsum[wave] = v;
  __syncthreads();
  uint32_t t = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < nw; ++w) t += wsum[w];
  return t;
}

// stats[slot] += sum(partial[0..n))
__global__ void __launch_bounds__(256) reduce_partials_kernel(const unsigned long long* __restrict__ partial, uint32_t n,
                                                               unsigned long long* __restrict__ stats, int slot) {
  __shared__ unsigned long long wsum[4];
  unsigned long long t = 0;
  for (uint32_t i = threadIdx.x; i < n; i += 256) t += partial[i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) stats[slot] += wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// ============================================================== setup (fast path)
template <int NA>
__device__ __forceinline__ void setup_body(const draw_params& P, const float4* __restrict__ vbuf, float4* __restrict__ tri_rec,
                                           float4* __restrict__ tri_attr, uint32_t* __restrict__ cnt,
                                           uint8_t* __restrict__ clip_flag, uint32_t* __restrict__ counters,
                                           uint32_t* __restrict__ wide_list, unsigned long long* __restrict__ partial,
                                           uint32_t* __restrict__ bsum, uint4* __restrict__ pairs, uint32_t bx) {
  uint32_t p = bx * blockDim.x + threadIdx.x;
  unsigned long long cprims = 0;
  float4 rec[5], attr[3 * (NA + 1)];  // this lane's setup words, stored by the whole wave below
  bool have = false;
  uint32_t c0 = 0;  // records of fan triangle 0 counted here
  uint4 kp = make_uint4(NO_PAIRS, NO_PAIRS, NO_PAIRS, NO_PAIRS);  // its first four (tile, full) pairs
  if (p < P.prim_count) {
    cnt[3 * p + 1] = 0;
    cnt[3 * p + 2] = 0;
    vtx<NA> v[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) load_vertex<NA>(P, vbuf, fetch_index(P, p, i), v[i]);
    // near (0,0,1,0) and far (0,0,-1,1) planes (clipper.cpp:21-27,104-120)
    const float4 pn = make_float4(0.0f, 0.0f, 1.0f, 0.0f), pf = make_float4(0.0f, 0.0f, -1.0f, 1.0f);
    bool in = true;
#pragma unroll
    for (int i = 0; i < 3; ++i) in &= !(dot4(pn, v[i].r[0]) < 0);
    if (in)
#pragma unroll
      for (int i = 0; i < 3; ++i) in &= !(dot4(pf, v[i].r[0]) < 0);
    clip_flag[p] = in ? 0 : 1;
    if (!in) {
      atomicAdd(&counters[CT_CLIP_PRIMS], 1u);
    } else {
      bool front = compute_front(v[0].r[0], v[1].r[0], v[2].r[0]);
      if (!cull_fn(P, front ? 1.0f : -1.0f)) {
        cprims = 1;
        // winding normalisation (clipper.cpp:59-65) by selects
        vtx<NA> o[3] = {v[0], vsel<NA>(front, v[1], v[2]), vsel<NA>(front, v[2], v[1])};
#pragma unroll
        for (int i = 0; i < 3; ++i) viewport_project<NA>(o[i], P);
        tri_setup<NA> t;
        triangle_info<NA>(o[0], o[1], o[2], t);
        if (t.valid) {
          have = true;
          tri_rec_words<NA>(t, rec);
          tri_attr_words<NA>(t, attr);
          float e[9], bb[4];
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) e[i * 3 + j] = t.e[i][j];
          for (int i = 0; i < 4; ++i) bb[i] = t.bb[i];
          if (bin_span(P, bb) > COUNT_WIDE)
            wide_list[atomicAdd(&counters[CT_WIDE], 1u)] = p;  // counted in the clip launch
          else if (pairs)
            c0 = count_gtiles_pairs(P, e, bb, kp);
          else
            c0 = count_gtiles(P, e, bb);
        }
      }
    }
    cnt[3 * p] = c0;
    if (pairs) {
      if (c0 > 4u) kp.x = NO_PAIRS;
      pairs[p] = kp;
    }
  }
  // The fan-0 slots of the wave's 64 consecutive primitives are consecutive: store the set-up
  // triangles' words coalesced (blockDim.x is 256).
  {
    __shared__ float4 stage[4 * WAVE_STAGE];
    const uint64_t mask = __ballot(have);
    if (mask) {
      const size_t p0 = static_cast<size_t>(bx) * blockDim.x + (threadIdx.x & ~63u);
      float4* ws = stage + (threadIdx.x >> 6) * WAVE_STAGE;
      wave_store_blocks<5>(tri_rec + p0 * 5, rec, mask, ws);
      wave_store_blocks<3 * (NA + 1)>(tri_attr + p0 * 3 * (NA + 1), attr, mask, ws);
    }
  }
  if (bsum) {  // uniform over the launch: the block's total; the clip launch adds the counts it makes
    const uint32_t t = block_sum_u32(c0);
    if (threadIdx.x == 0) bsum[bx] = t;
  }
  if (P.count_stats) block_partial(partial, cprims);  // uniform over the launch
}
template <int NA>
__global__ void __launch_bounds__(256) setup_kernel(draw_params P, const float4* __restrict__ vbuf, float4* __restrict__ tri_rec,
                                                     float4* __restrict__ tri_attr,
                                                     uint32_t* __restrict__ cnt, uint8_t* __restrict__ clip_flag,
                                                     uint32_t* __restrict__ counters, uint32_t* __restrict__ wide_list,
                                                     unsigned long long* __restrict__ partial, uint32_t* __restrict__ bsum,
                                                     uint4* __restrict__ pairs) {
  setup_body<NA>(P, vbuf, tri_rec, tri_attr, cnt, clip_flag, counters, wide_list, partial, bsum, pairs, blockIdx.x);
}

// ============================================================== clipping (slow path)
template <int NA>
__device__ __forceinline__ void lerp_vtx(vtx<NA>& o, const vtx<NA>& a, const vtx<NA>& b, float t, const draw_params& P) {
  o.r[0] = make_float4(a.r[0].x + (b.r[0].x - a.r[0].x) * t, a.r[0].y + (b.r[0].y - a.r[0].y) * t,
                       a.r[0].z + (b.r[0].z - a.r[0].z) * t, a.r[0].w + (b.r[0].w - a.r[0].w) * t);
#pragma unroll
  for (int i = 0; i < NA; ++i) {
    o.r[i + 1] = a.r[i + 1];
    if (!(P.attr_mod[i] & AM_NOINTERPOLATION))
      o.r[i + 1] = make_float4(o.r[i + 1].x + (b.r[i + 1].x - a.r[i + 1].x) * t, o.r[i + 1].y + (b.r[i + 1].y - a.r[i + 1].y) * t,
                               o.r[i + 1].z + (b.r[i + 1].z - a.r[i + 1].z) * t, o.r[i + 1].w + (b.r[i + 1].w - a.r[i + 1].w) * t);
  }
}

// clip_triangle_to_poly_general + clip_solid_triangle (clipper.cpp:43-228) for primitive p
template <int NA>
__device__ uint32_t clip_primitive(const draw_params& P, uint32_t p, const float4* __restrict__ vbuf, float4* __restrict__ tri_rec,
                                   float4* __restrict__ tri_attr, int32_t* __restrict__ clip_slot,
                               float4* __restrict__ clip_verts, uint32_t* __restrict__ clip_hdr, uint32_t capacity,
                               uint32_t* __restrict__ counters, unsigned long long* __restrict__ stats) {
  vtx<NA> cv[2][5];
  uint32_t nc[2] = {3, 0};
  for (int i = 0; i < 3; ++i) load_vertex<NA>(P, vbuf, fetch_index(P, p, i), cv[0][i]);
  const float4 planes[2] = {make_float4(0.0f, 0.0f, 1.0f, 0.0f), make_float4(0.0f, 0.0f, -1.0f, 1.0f)};
  int src = 0, dst = 1;
  bool front = false;
  for (int pl = 0; pl < 2; ++pl) {
    nc[dst] = 0;
    float d0 = 0.0f;
    if (nc[src] != 0) d0 = dot4(planes[pl], cv[src][0].r[0]);
    for (uint32_t i = 0, j = 1; i < nc[src]; ++i, ++j) {
      j %= nc[src];
      float d1 = dot4(planes[pl], cv[src][j].r[0]);
      if (d0 >= 0.0f) {
        cv[dst][nc[dst]++] = cv[src][i];
        if (d1 < 0.0f) lerp_vtx<NA>(cv[dst][nc[dst]++], cv[src][i], cv[src][j], d0 / (d0 - d1), P);
      } else if (d1 >= 0.0f) {
        lerp_vtx<NA>(cv[dst][nc[dst]++], cv[src][j], cv[src][i], d1 / (d1 - d0), P);
      }
      d0 = d1;
    }
    if (pl == 0 && nc[dst] >= 3) {
      float px[3], py[3];
      for (int v = 0; v < 3; ++v) {
        float iaw = 1 / fabsf(cv[dst][v].r[0].w);
        px[v] = cv[dst][v].r[0].x * iaw;
        py[v] = cv[dst][v].r[0].y * iaw;
      }
      float ax = px[2] - px[0], ay = py[2] - py[0], bx = px[1] - px[0], by = py[1] - py[0];
      float area = ax * by - ay * bx;
      front = area > 0.0f;
      if (cull_fn(P, area)) return 0;
    }
    src ^= 1;
    dst ^= 1;
  }
  uint32_t n = nc[src];
  if (n < 3) return 0;
  atomicAdd(&stats[ST_CPRIMITIVES], static_cast<unsigned long long>(n - 2));
  uint32_t slot = atomicAdd(&counters[CT_CLIP_SLOTS], 1u);
  if (slot >= capacity) {
    atomicAdd(&counters[CT_CLIP_OVERFLOW], 1u);
    return 0;
  }
  clip_slot[p] = static_cast<int32_t>(slot);
  clip_hdr[slot] = n | (front ? 0x100u : 0u);
  float4* out = clip_verts + static_cast<size_t>(slot) * 5 * (NA + 1);
  for (uint32_t i = 0; i < n; ++i) {
    viewport_project<NA>(cv[src][i], P);  // once per polygon vertex, as for the unique vertices
    for (int r = 0; r <= NA; ++r) out[i * (NA + 1) + r] = cv[src][i].r[r];
  }
  uint32_t fans = 0;
  for (uint32_t t = 1; t + 1 < n; ++t) {
    const vtx<NA>& a = cv[src][0];
    const vtx<NA>& b = front ? cv[src][t] : cv[src][t + 1];
    const vtx<NA>& c = front ? cv[src][t + 1] : cv[src][t];
    tri_setup<NA> ts;
    triangle_info<NA>(a, b, c, ts);
    uint32_t id = 3 * p + (t - 1);
    if (!ts.valid) continue;
    write_tri_rec<NA>(tri_rec + tri_slot(P, id) * 5, ts);
    write_tri_attr<NA>(tri_attr + tri_slot(P, id) * 3 * (NA + 1), ts);
    fans |= 1u << (t - 1);
  }
  return fans;
}

// Clipping launch: a bounded grid strides over the primitives setup flagged (none, usually).  With
// `bsum` (single draws) the counts it makes are added to their 256-primitive block's total.
template <int NA>
__device__ __forceinline__ void clip_work(const draw_params& P, const float4* __restrict__ vbuf, float4* __restrict__ tri_rec,
                                          float4* __restrict__ tri_attr, uint32_t* __restrict__ cnt,
                                          const uint8_t* __restrict__ clip_flag, int32_t* __restrict__ clip_slot,
                                          float4* __restrict__ clip_verts, uint32_t* __restrict__ clip_hdr, uint32_t capacity,
                                          uint32_t* __restrict__ counters, unsigned long long* __restrict__ stats,
                                          const uint32_t* __restrict__ wide_list, uint32_t* __restrict__ bsum,
                                          uint32_t nclip, uint32_t nwide, uint32_t bx, uint32_t gx) {
  // triangles over many 64x64 tiles: one wave counts each one's records
  const uint32_t wave = (bx * blockDim.x + threadIdx.x) >> 6, waves = (gx * blockDim.x) >> 6;
  for (uint32_t w = wave; w < nwide; w += waves) {
    const uint32_t p = wide_list[w];
    float e[9], bb[4];
    read_tri_rec(tri_rec + static_cast<size_t>(p) * 5, e, bb);  // fan 0 of primitive p
    const uint32_t n = count_gtiles_wave(P, e, bb);
    if (__lane_id() == 0) {
      cnt[3 * p] = n;
      if (bsum) atomicAdd(&bsum[p / 256], n);
    }
  }
  if (nclip == 0) return;
  // lanes clip the flagged primitives among the wave's 64; the wave then counts the records of the
  // resulting triangles together (clipped triangles are few and often reach across the target)
  const int lane = __lane_id();
  for (uint32_t b = wave * 64; b < P.prim_count; b += waves * 64) {
    const uint32_t p = b + static_cast<uint32_t>(lane);
    uint32_t fans = 0;  // bit k: fan triangle 3p + k was set up
    if (p < P.prim_count && clip_flag[p])
      fans = clip_primitive<NA>(P, p, vbuf, tri_rec, tri_attr, clip_slot, clip_verts, clip_hdr, capacity, counters, stats);
    __threadfence();  // the fans' records are read back by every lane of the wave
    for (unsigned long long m = __ballot(fans != 0); m; m &= m - 1) {
      const int src = __builtin_ctzll(m);
      const uint32_t fs = static_cast<uint32_t>(__shfl(static_cast<int>(fans), src, 64));
      const uint32_t ps = b + static_cast<uint32_t>(src);
      for (int k = 0; k < 3; ++k) {
        if (!(fs & (1u << k))) continue;
        const uint32_t id = 3 * ps + k;
        float e[9], bb[4];
        read_tri_rec(tri_rec + tri_slot(P, id) * 5, e, bb);
        const uint32_t n = count_gtiles_wave(P, e, bb);
        if (lane == 0) {
          cnt[id] = n;
          if (bsum) atomicAdd(&bsum[ps / 256], n);
        }
      }
    }
  }
}

// counters[CT_TOTAL] = the sum of the block totals (every thread of the workgroup calls it).
__device__ __forceinline__ void record_total(const draw_params& P, const uint32_t* __restrict__ bsum,
                                             uint32_t* __restrict__ counters) {
  const uint32_t nb = (P.prim_count + 255) / 256;
  uint32_t t = 0;
  for (uint32_t i0 = 0; i0 < nb; i0 += 8 * blockDim.x) {
    uint32_t v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t i = i0 + k * blockDim.x + threadIdx.x;
      const uint32_t x = __hip_atomic_load(bsum + min(i, nb - 1u), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      v[k] = i < nb ? x : 0u;  // no branch per load (see load_or_zero)
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) t += v[k];
  }
  t = block_sum_u32(t);
  if (threadIdx.x == 0) counters[CT_TOTAL] = t;
}

template <int NA>
__device__ __forceinline__ void clip_body(const draw_params& P, const float4* __restrict__ vbuf, float4* __restrict__ tri_rec,
                                          float4* __restrict__ tri_attr, uint32_t* __restrict__ cnt,
                                          const uint8_t* __restrict__ clip_flag, int32_t* __restrict__ clip_slot,
                                          float4* __restrict__ clip_verts, uint32_t* __restrict__ clip_hdr, uint32_t capacity,
                                          uint32_t* __restrict__ counters, unsigned long long* __restrict__ stats,
                                          const uint32_t* __restrict__ wide_list, uint32_t* __restrict__ bsum, uint32_t bx,
                                          uint32_t gx) {
  const uint32_t nclip = counters[CT_CLIP_PRIMS], nwide = counters[CT_WIDE];
  const bool flagged = nclip != 0 || nwide != 0;  // uniform over the launch
  if (flagged)
    clip_work<NA>(P, vbuf, tri_rec, tri_attr, cnt, clip_flag, clip_slot, clip_verts, clip_hdr, capacity, counters, stats,
                  wide_list, bsum, nclip, nwide, bx, gx);
  if (!bsum) return;
  // single draws: the record total, once every workgroup's counts are in the block totals (the last
  // workgroup to finish sums them; with nothing flagged, setup's totals are final and workgroup 0 does)
  if (flagged) {
    __shared__ uint32_t s_last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) s_last = atomicAdd(&counters[CT_CLIP_DONE], 1u) == gx - 1 ? 1u : 0u;
    __syncthreads();
    if (!s_last) return;
    __threadfence();
  } else if (bx != 0) {
    return;
  }
  record_total(P, bsum, counters);
}
template <int NA>
__global__ void clip_kernel(draw_params P, const float4* __restrict__ vbuf, float4* __restrict__ tri_rec,
                            float4* __restrict__ tri_attr, uint32_t* __restrict__ cnt, const uint8_t* __restrict__ clip_flag, int32_t* __restrict__ clip_slot,
                            float4* __restrict__ clip_verts, uint32_t* __restrict__ clip_hdr, uint32_t capacity,
                            uint32_t* __restrict__ counters, unsigned long long* __restrict__ stats,
                            const uint32_t* __restrict__ wide_list, uint32_t* __restrict__ bsum) {
  clip_body<NA>(P, vbuf, tri_rec, tri_attr, cnt, clip_flag, clip_slot, clip_verts, clip_hdr, capacity, counters, stats,
                wide_list, bsum, blockIdx.x, gridDim.x);
}

// ============================================================== record emission
// off: exclusive scan of the per-triangle record counts; a triangle owns [off[id], off[id+1]).
// One thread per primitive, over its (at most 3) clipped triangles in fan order.  Records beyond
// `capacity` are dropped; the host compares the total with it and emits again.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  int lane = __lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// Records of one clipped triangle emitted by a whole wave: lanes take the triangle's 64x64 tiles in
// for_each_gtile's order, count each tile's GPU tiles with the same tests, and write at their
// wave-scanned offsets -- for triangles that reach many tiles, which one lane would walk serially.
__device__ __forceinline__ void emit_wide(const draw_params& P, const float4* __restrict__ tri_rec, uint32_t id,
                                          uint32_t first, uint32_t next, uint32_t capacity,
                                          uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int lane = __lane_id();
  float e[9], bb[4];
  read_tri_rec(tri_rec + tri_slot(P, id) * 5, e, bb);
  tri_bin t;
  bin_setup(P, e, bb, t);
  const int sub = TILE / P.gt, w = t.ex - t.sx;
  const int R = (t.ex > t.sx && t.ey > t.sy) ? w * (t.ey - t.sy) : 0;
  uint32_t base = first;
  for (int t0 = 0; t0 < R; t0 += 64) {
    const int tt = t0 + lane;
    uint64_t m = 0;
    bool full = false;
    int x = 0, y = 0;
    if (tt < R) {
      x = t.sx + tt % w;
      y = t.sy + tt / w;
      m = tile_gmask(P, e, bb, t, x, y, full);
    }
    const uint32_t cnt = static_cast<uint32_t>(__popcll(m));
    const uint32_t incl = wave_incl_scan(cnt);
    uint32_t o = base + incl - cnt;
    for (; m; m &= m - 1) {
      const int b = __builtin_ctzll(m);
      if (o < next && o < capacity) {
        keys[o] = static_cast<uint32_t>((y * sub + b / sub) * P.gtiles_x + x * sub + b % sub);
        vals[o] = (id << 1) | (full ? 1u : 0u);
      }
      ++o;
    }
    base += static_cast<uint32_t>(__shfl(static_cast<int>(incl), 63, 64));
  }
}

// Records beyond which a triangle is emitted by its whole wave (so is any triangle over more than
// COUNT_WIDE 64x64 tiles).
constexpr uint32_t EMIT_WIDE = 64;

// Records of primitive p's fan triangles: triangle 3p+k owns [o[k], o[k+1]).  Every lane of the wave
// calls it (p >= prim_count: nothing of its own).
// pr: setup's first four (tile << 1 | full) pairs of fan triangle 0, pr.x == NO_PAIRS when it kept
// none (then, and for the other fan triangles, the tiles are walked again).
__device__ __forceinline__ void emit_records(const draw_params& P, const float4* __restrict__ tri_rec, uint32_t p,
                                             const uint32_t (&o)[4], uint4 pr, uint32_t capacity,
                                             uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  uint32_t wide = 0;  // fan triangles left to the wave
  if (p < P.prim_count) {
    const uint32_t id3 = 3 * p;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const uint32_t id = id3 + k;
      const uint32_t first = o[k], next = o[k + 1];
      if (next == first) continue;
      if (k == 0 && pr.x != NO_PAIRS) {  // at most four records, as setup's walk found them
        const uint32_t v[4] = {pr.x, pr.y, pr.z, pr.w};
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i)
          if (first + i < next && first + i < capacity) {
            keys[first + i] = v[i] >> 1;
            vals[first + i] = (id << 1) | (v[i] & 1u);
          }
        continue;
      }
      if (next - first > EMIT_WIDE) {
        wide |= 1u << k;
        continue;
      }
      float e[9], bb[4];
      read_tri_rec(tri_rec + tri_slot(P, id) * 5, e, bb);
      if (bin_span(P, bb) > COUNT_WIDE) {  // few records over a large box: the walk is over the box
        wide |= 1u << k;
        continue;
      }
      uint32_t q = first;
      for_each_gtile(P, e, bb, [&](uint32_t g, bool full) {
        if (q < next && q < capacity) {  // the enumeration repeats the counting pass exactly
          keys[q] = g;
          vals[q] = (id << 1) | (full ? 1u : 0u);
        }
        ++q;
      });
    }
  }
  // every lane of the wave takes part in its lanes' wide triangles, one after another
  unsigned long long pending = __ballot(wide != 0);
  while (pending) {
    const int src = __ffsll(static_cast<long long>(pending)) - 1;
    pending &= pending - 1;
    const uint32_t ws = static_cast<uint32_t>(__shfl(static_cast<int>(wide), src, 64));
    const uint32_t ps = static_cast<uint32_t>(__shfl(static_cast<int>(p), src, 64));
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (!(ws & (1u << k))) continue;
      const uint32_t first = static_cast<uint32_t>(__shfl(static_cast<int>(o[k]), src, 64));
      const uint32_t next = static_cast<uint32_t>(__shfl(static_cast<int>(o[k + 1]), src, 64));
      emit_wide(P, tri_rec, 3 * ps + k, first, next, capacity, keys, vals);
    }
  }
}
// off: exclusive scan of the counts, followed by the total (batches)
__device__ __forceinline__ void emit_body(const draw_params& P, const float4* __restrict__ tri_rec,
                                          const uint32_t* __restrict__ off, const uint32_t* __restrict__ total,
                                          uint32_t capacity, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                          uint32_t bx) {
  const uint32_t p = bx * blockDim.x + threadIdx.x;
  const uint32_t n = 3 * P.prim_count;
  uint32_t o[4] = {0, 0, 0, 0};
  if (p < P.prim_count) {
    const uint32_t id3 = 3 * p;
    o[0] = off[id3];
    o[1] = off[id3 + 1];
    o[2] = off[id3 + 2];
    o[3] = (id3 + 3 < n) ? off[id3 + 3] : *total;
  }
  emit_records(P, tri_rec, p, o, make_uint4(NO_PAIRS, NO_PAIRS, NO_PAIRS, NO_PAIRS), capacity, keys, vals);
}
// Single draws: the offsets come from the counts of the workgroup's 256 primitives (a workgroup
// scan) and the totals of the blocks before it, which setup and the clip launch left in `bsum`, so
// no scan of the counts runs: triangle id's first record is the sum of the counts of the triangles
// before it, as an exclusive scan of the counts gives it.
__global__ void __launch_bounds__(256) emit_kernel(draw_params P, const float4* __restrict__ tri_rec,
                                                   const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ bsum,
                                                   const uint4* __restrict__ pairs, uint32_t capacity,
                                                   uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  __shared__ uint32_t wtot[4], wbefore[4];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t p = blockIdx.x * 256 + tid;
  uint32_t c[3] = {0, 0, 0};
  uint4 pr = make_uint4(NO_PAIRS, NO_PAIRS, NO_PAIRS, NO_PAIRS);
  if (p < P.prim_count) {
    c[0] = cnt[3 * p];
    c[1] = cnt[3 * p + 1];
    c[2] = cnt[3 * p + 2];
    pr = pairs[p];
  }
  uint32_t before = 0;  // records of the blocks before this one
  for (uint32_t i0 = 0; i0 < blockIdx.x; i0 += 4 * 256) {
    uint32_t v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t i = i0 + k * 256 + tid;
      v[k] = load_or_zero(bsum, i, blockIdx.x);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) before += v[k];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) before += static_cast<uint32_t>(__shfl_xor(static_cast<int>(before), off, 64));
  const uint32_t t = c[0] + c[1] + c[2];
  const uint32_t incl = wave_incl_scan(t);
  if (lane == 63) {
    wtot[wave] = incl;
    wbefore[wave] = before;
  }
  __syncthreads();
  uint32_t o[4];
  o[0] = wbefore[0] + wbefore[1] + wbefore[2] + wbefore[3] + incl - t;
  for (uint32_t w = 0; w < wave; ++w) o[0] += wtot[w];
  o[1] = o[0] + c[0];
  o[2] = o[1] + c[1];
  o[3] = o[2] + c[2];
  emit_records(P, tri_rec, p, o, pr, capacity, keys, vals);
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t n, uint32_t v) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ============================================================== scan (exclusive, u32)
constexpr int SCAN_PER_THREAD = 16;
static_assert(SCAN_PER_THREAD % 4 == 0, "scan blocks move as 16-byte words");
constexpr int SCAN_BLOCK = 256 * SCAN_PER_THREAD;  // elements per workgroup

// Exclusive scan of one SCAN_BLOCK-element block in place.  The block total goes to sums[block]
// (when sums is given); a single-block launch also stores it to *total.
__global__ void __launch_bounds__(256) scan_block_kernel(uint32_t* data, uint32_t n, uint32_t* sums, uint32_t* total) {
  __shared__ uint32_t wsum[4];
  // 64-bit positions: n may come within one block of 2^32
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * SCAN_BLOCK + threadIdx.x * SCAN_PER_THREAD;
  // a full block of 16-byte-aligned data moves as four 16-byte words per lane
  const bool vec = (static_cast<uint64_t>(blockIdx.x) + 1) * SCAN_BLOCK <= n && (reinterpret_cast<uintptr_t>(data) & 15) == 0;
  uint32_t v[SCAN_PER_THREAD];
  uint32_t t = 0;
  if (vec) {
    const uint4* d4 = reinterpret_cast<const uint4*>(data + base);
#pragma unroll
    for (int i = 0; i < SCAN_PER_THREAD / 4; ++i) {
      const uint4 q = d4[i];
      v[4 * i] = q.x;
      v[4 * i + 1] = q.y;
      v[4 * i + 2] = q.z;
      v[4 * i + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < SCAN_PER_THREAD; ++i) v[i] = (base + i < n) ? data[base + i] : 0u;
  }
#pragma unroll
  for (int i = 0; i < SCAN_PER_THREAD; ++i) t += v[i];
  const uint32_t incl = wave_incl_scan(t);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  uint32_t wprefix = 0;
  for (int w = 0; w < wave; ++w) wprefix += wsum[w];
  uint32_t run = wprefix + incl - t;
#pragma unroll
  for (int i = 0; i < SCAN_PER_THREAD; ++i) {
    const uint32_t x = v[i];
    v[i] = run;
    run += x;
  }
  if (vec) {
    uint4* d4 = reinterpret_cast<uint4*>(data + base);
#pragma unroll
    for (int i = 0; i < SCAN_PER_THREAD / 4; ++i) d4[i] = make_uint4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
  } else {
#pragma unroll
    for (int i = 0; i < SCAN_PER_THREAD; ++i)
      if (base + i < n) data[base + i] = v[i];
  }
  if (threadIdx.x == 255) {
    const uint32_t block_total = wprefix + incl;
    if (sums) sums[blockIdx.x] = block_total;
    if (total && gridDim.x == 1) *total = block_total;
  }
}

__global__ void __launch_bounds__(256) scan_add_kernel(uint32_t* data, uint32_t n, const uint32_t* sums) {
  const uint32_t add = sums[blockIdx.x];
  const uint64_t b0 = static_cast<uint64_t>(blockIdx.x) * SCAN_BLOCK;
#pragma unroll
  for (int k = 0; k < SCAN_PER_THREAD; ++k) {
    const uint64_t i = b0 + k * 256 + threadIdx.x;
    if (i < n) data[i] += add;
  }
}

size_t scan_tmp_size(uint32_t n) {
  size_t total = 0;
  uint64_t m = n;
  while (true) {
    const uint64_t nb = (m + SCAN_BLOCK - 1) / SCAN_BLOCK;
    if (nb <= 1) break;
    total += nb;
    m = nb;
  }
  return total + 4;
}

// Reduce-then-scan: blocks scan in place, the block totals are scanned the same way (one block up
// to SCAN_BLOCK^2 elements), then added back.  The top-level single block yields the total.
static void scan_rec(uint32_t* data, uint32_t n, uint32_t* tmp, uint32_t* total_dev, hipStream_t s) {
  const uint32_t nb = static_cast<uint32_t>((static_cast<uint64_t>(n) + SCAN_BLOCK - 1) / SCAN_BLOCK);
  if (nb <= 1) {
    hipLaunchKernelGGL(scan_block_kernel, dim3(1), dim3(256), 0, s, data, n, nullptr, total_dev);
    return;
  }
  uint32_t* sums = tmp;
  hipLaunchKernelGGL(scan_block_kernel, dim3(nb), dim3(256), 0, s, data, n, sums, nullptr);
  scan_rec(sums, nb, tmp + nb, total_dev, s);
  hipLaunchKernelGGL(scan_add_kernel, dim3(nb), dim3(256), 0, s, data, n, sums);
}

void launch_exclusive_scan(uint32_t* data, uint32_t n, uint32_t* tmp, uint32_t* total_dev, hipStream_t s) {
  scan_rec(data, n, tmp, total_dev, s);
}

// ============================================================== radix sort (stable LSD, 8-bit digits)
constexpr int SORT_BLOCK = 4096;  // 256 threads x 16 items

__global__ void __launch_bounds__(256) radix_hist_kernel(const uint32_t* __restrict__ keys, uint32_t n, int shift,
                                                          uint32_t nblocks, uint32_t* __restrict__ hist) {
  __shared__ uint32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  uint32_t b0 = blockIdx.x * SORT_BLOCK;
  uint32_t end = min(n, b0 + SORT_BLOCK);
  for (uint32_t i = b0 + threadIdx.x; i < end; i += 256) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
  __syncthreads();
  hist[threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// Each wave ranks its 1024 consecutive items in 16 rounds of 64; a lane finds the lanes sharing its
// digit with eight ballots (one per digit bit), so ranks follow input order (stability) at a cost
// independent of how many digits a round holds.
__global__ void __launch_bounds__(256) radix_scatter_kernel(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                             uint32_t n, int shift, uint32_t nblocks,
                                                             const uint32_t* __restrict__ hist_scanned,
                                                             uint32_t* __restrict__ kout, uint32_t* __restrict__ vout) {
  __shared__ uint32_t wcnt[4][256];
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = threadIdx.x; i < 4 * 256; i += 256) (&wcnt[0][0])[i] = 0;
  __syncthreads();
  uint32_t base = blockIdx.x * SORT_BLOCK + wave * 1024;
  uint32_t rank[16], key[16], val[16];
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    uint32_t i = base + r * 64 + lane;
    bool valid = i < n;
    key[r] = valid ? kin[i] : 0u;
    val[r] = valid ? vin[i] : 0u;
    uint32_t d = (key[r] >> shift) & 255u;
    // m: the valid lanes holding this lane's digit, from one ballot per digit bit
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = ((d >> b) & 1u) != 0;
      const uint64_t bal = __ballot(bit);
      m &= bit ? bal : ~bal;
    }
    const uint32_t before = valid ? wcnt[wave][d] : 0u;
    rank[r] = before + static_cast<uint32_t>(__popcll(m & lt));
    __builtin_amdgcn_wave_barrier();
    if (valid && (m & lt) == 0) wcnt[wave][d] = before + static_cast<uint32_t>(__popcll(m));  // lowest lane of the digit
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  {
    uint32_t d = threadIdx.x, s = 0;
    for (int w = 0; w < 4; ++w) {
      uint32_t c = wcnt[w][d];
      wcnt[w][d] = s;
      s += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    uint32_t i = base + r * 64 + lane;
    if (i < n) {
      uint32_t d = (key[r] >> shift) & 255u;
      uint32_t pos = hist_scanned[d * nblocks + blockIdx.x] + wcnt[wave][d] + rank[r];
      if (pos < n) {
        kout[pos] = key[r];
        vout[pos] = val[r];
      }
    }
  }
}

int launch_sort(const draw_buffers& B, uint32_t n, int bits, hipStream_t s) {
  int cur = 0;
  if (n == 0) return cur;
  uint32_t nblocks = (n + SORT_BLOCK - 1) / SORT_BLOCK;
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(radix_hist_kernel, dim3(nblocks), dim3(256), 0, s, B.keys[cur], n, shift, nblocks, B.hist);
    launch_exclusive_scan(B.hist, 256 * nblocks, B.scan_tmp, nullptr, s);
    hipLaunchKernelGGL(radix_scatter_kernel, dim3(nblocks), dim3(256), 0, s, B.keys[cur], B.vals[cur], n, shift, nblocks,
                       B.hist, B.keys[cur ^ 1], B.vals[cur ^ 1]);
    cur ^= 1;
  }
  return cur;
}

// [ts, te) of every GPU tile in the sorted records.  With T > 0 the thread at the last record of a
// tile (or before the first record) also writes the empty ranges of the tiles up to the next
// record's tile, so every tile is written exactly once without a separate clear (dense records).
// ---------------------------------------------------------------- one-pass bucket sort by GPU tile
// For the order-independent path the order of a tile's records is immaterial, so one counting pass
// replaces the stable radix passes: per-workgroup tile histograms in LDS, one scan of the
// tile-major histogram, and a scatter whose in-workgroup ranks come from LDS atomics.
constexpr int BUCKET_PER_THREAD = 16;
constexpr int BUCKET_ITEMS = 256 * BUCKET_PER_THREAD;  // records per workgroup
constexpr uint32_t BUCKET_MAX_TILES = 8192;

// PT records per thread: 256 * PT records per workgroup (the batched path tunes PT)
// Column of workgroup b in a tile's row of the bucket histogram: the workgroups of one XCD
// (b mod 8, as workgroups are dealt to the XCDs) take adjacent columns, so each histogram line is
// written from one XCD's L2.  Any order of the workgroups within a tile is valid here (the tile's
// records carry no order on this path); the map is a bijection on [0, nblocks).
__device__ __forceinline__ uint32_t bucket_column(uint32_t b, uint32_t nblocks) {
  const uint32_t q = nblocks / 8, r = nblocks % 8, x = b % 8;
  return x * q + min(x, r) + b / 8;
}
// The per-workgroup tile counts live in LDS for BUCKET_MAX_TILES tiles at a time: a target with more
// GPU tiles is counted range by range, each pass over the workgroup's keys taking the tiles of one
// range (the keys stay in registers).
template <int PT>
__device__ __forceinline__ void bucket_hist_body(const uint32_t* __restrict__ keys, uint32_t n, uint32_t T,
                                                 uint32_t nblocks, uint32_t* __restrict__ hist, uint32_t bx) {
  __shared__ uint32_t cnt[BUCKET_MAX_TILES];
  const uint32_t b0 = bx * (256 * PT);
  uint32_t k[PT];
#pragma unroll
  for (int j = 0; j < PT; ++j) {
    const uint32_t i = b0 + j * 256 + threadIdx.x;
    k[j] = i < n ? keys[i] : 0xFFFFFFFFu;
  }
  const uint32_t col = bucket_column(bx, nblocks);
  for (uint32_t r0 = 0; r0 < T; r0 += BUCKET_MAX_TILES) {
    const uint32_t tn = min(T - r0, BUCKET_MAX_TILES);
    if (r0) __syncthreads();  // the previous range's counts are written out
    for (uint32_t d = threadIdx.x; d < tn; d += 256) cnt[d] = 0;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PT; ++j)
      if (k[j] - r0 < tn) atomicAdd(&cnt[k[j] - r0], 1u);  // also skips the 0xFFFFFFFF padding
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < tn; d += 256) hist[(r0 + d) * nblocks + col] = cnt[d];
  }
}

// Only the tiles present in the workgroup's records fetch their (strided) offsets.  Positions at
// or beyond `limit` (the output's size) are dropped.
template <int PT>
__device__ __forceinline__ void bucket_scatter_body(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                    uint32_t n, uint32_t T, uint32_t nblocks,
                                                    const uint32_t* __restrict__ hist_scanned, uint32_t* __restrict__ vout,
                                                    uint32_t limit, uint32_t bx) {
  __shared__ uint32_t off[BUCKET_MAX_TILES];
  const uint32_t b0 = bx * (256 * PT);
  uint32_t k[PT], v[PT];
#pragma unroll
  for (int j = 0; j < PT; ++j) {
    const uint32_t i = b0 + j * 256 + threadIdx.x;
    k[j] = i < n ? keys[i] : 0xFFFFFFFFu;
    v[j] = i < n ? vals[i] : 0u;
  }
  for (uint32_t r0 = 0; r0 < T; r0 += BUCKET_MAX_TILES) {
    const uint32_t tn = min(T - r0, BUCKET_MAX_TILES);
    if (r0) __syncthreads();  // the previous range's offsets are consumed
    for (uint32_t d = threadIdx.x; d < tn; d += 256) off[d] = 0xFFFFFFFFu;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PT; ++j)
      if (k[j] - r0 < tn) off[k[j] - r0] = 0u;  // mark the tiles present
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < tn; d += 256)
      if (off[d] == 0u) off[d] = hist_scanned[(r0 + d) * nblocks + bucket_column(bx, nblocks)];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PT; ++j)
      if (k[j] - r0 < tn) {
        const uint32_t pos = atomicAdd(&off[k[j] - r0], 1u);
        if (pos < limit) vout[pos] = v[j];
      }
  }
}

// atomicAdd(&counters[key], 1) for every active lane, one atomic per distinct key of the wave;
// returns the lane's old value plus its rank among the lanes with the same key.
__device__ __forceinline__ uint32_t wave_counted_add(uint32_t* counters, uint32_t key) {
  const int lane = __lane_id();
  uint32_t result = 0;
  while (true) {
    const uint32_t k0 = __builtin_amdgcn_readfirstlane(key);
    const unsigned long long m = __ballot(key == k0);
    if (key == k0) {
      const uint32_t rank = static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)));
      uint32_t old = 0;
      if (rank == 0) old = atomicAdd(&counters[k0], static_cast<uint32_t>(__popcll(m)));
      result = static_cast<uint32_t>(__shfl(static_cast<int>(old), __builtin_ctzll(m), 64)) + rank;
      break;
    }
  }
  return result;
}
// Record-count class of a GPU tile for the scheduling order: class 11 - floor(log2 n) (heaviest
// first, 2048 records and more in class 0), class 15 for an empty tile.
__device__ __forceinline__ uint32_t tile_class(uint32_t n) {
  if (n == 0) return 15u;
  const int lg = 31 - __clz(n);
  return static_cast<uint32_t>(min(max(11 - lg, 0), 14));
}

// tile t's records are [scanned[t * nblocks], scanned[(t + 1) * nblocks]) (`end` for the last tile)
__device__ __forceinline__ void bucket_range_body(const uint32_t* __restrict__ hist_scanned, uint32_t end, uint32_t T,
                                                  uint32_t nblocks, uint32_t* __restrict__ ts, uint32_t* __restrict__ te,
                                                  uint32_t t) {
  if (t >= T) return;
  ts[t] = hist_scanned[t * nblocks];
  te[t] = (t + 1 < T) ? hist_scanned[(t + 1) * nblocks] : end;
}

uint32_t bucket_hist_size(uint32_t T, uint32_t n) {
  return T * std::max<uint32_t>((n + BUCKET_ITEMS - 1) / BUCKET_ITEMS, 1);
}

// Counting sort of one draw's records by GPU tile (order-independent draws): the order within a tile
// is immaterial, so each workgroup of BUCKET_ITEMS records counts its tiles in LDS, noting the
// tiles it touches, and adds the counts of those tiles to the per-tile totals; one workgroup turns
// the totals into tile ranges; the scatter counts again, reserves its range in each touched tile
// with one atomic and places its records at the LDS ranks.  tile_count and tile_cursor are zero
// at the start of the draw.
// Counts the workgroup's records of tiles [r0, r0 + tn) in LDS (rank[j]: the record's rank among the
// workgroup's records of its tile) and lists the tiles it touches (relative to r0).
__device__ __forceinline__ void tile_count_lds(const uint32_t (&k)[BUCKET_PER_THREAD], uint32_t (&rank)[BUCKET_PER_THREAD],
                                               uint32_t r0, uint32_t tn, uint32_t* cnt, uint16_t* touched,
                                               uint32_t* ntouched) {
  for (uint32_t d = threadIdx.x; d < tn; d += 256) cnt[d] = 0;
  if (threadIdx.x == 0) *ntouched = 0;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < BUCKET_PER_THREAD; ++j)
    if (k[j] - r0 < tn) {  // also skips the 0xFFFFFFFF padding
      rank[j] = atomicAdd(&cnt[k[j] - r0], 1u);
      if (rank[j] == 0) touched[atomicAdd(ntouched, 1u)] = static_cast<uint16_t>(k[j] - r0);
    }
  __syncthreads();
}
__device__ __forceinline__ void tile_load_keys(const uint32_t* __restrict__ keys, uint32_t n, uint32_t b0,
                                               uint32_t (&k)[BUCKET_PER_THREAD]) {
#pragma unroll
  for (int j = 0; j < BUCKET_PER_THREAD; ++j) {
    const uint32_t i = b0 + j * 256 + threadIdx.x;
    k[j] = i < n ? keys[i] : 0xFFFFFFFFu;
  }
}
__global__ void __launch_bounds__(256) tile_count_kernel(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ count,
                                                         uint32_t cap, uint32_t T, uint32_t* __restrict__ tile_count) {
  __shared__ uint32_t cnt[BUCKET_MAX_TILES];
  __shared__ uint16_t touched[BUCKET_ITEMS];
  __shared__ uint32_t ntouched;
  const uint32_t n = min(*count, cap);
  const uint32_t b0 = blockIdx.x * BUCKET_ITEMS;
  if (b0 >= n) return;
  uint32_t k[BUCKET_PER_THREAD], rank[BUCKET_PER_THREAD];
  tile_load_keys(keys, n, b0, k);
  for (uint32_t r0 = 0; r0 < T; r0 += BUCKET_MAX_TILES) {
    const uint32_t tn = min(T - r0, BUCKET_MAX_TILES);
    if (r0) __syncthreads();  // the previous range's touched tiles are consumed
    tile_count_lds(k, rank, r0, tn, cnt, touched, &ntouched);
    for (uint32_t i = threadIdx.x; i < ntouched; i += 256) {
      const uint32_t t = touched[i];
      atomicAdd(&tile_count[r0 + t], cnt[t]);
    }
  }
}
// counters[key] += the number of valid lanes with this lane's key (< 16), one LDS atomic per distinct
// key; returns the counter's old value plus the lane's rank among them.  Lanes with equal keys are
// matched by four ballots, one per key bit.  All lanes of the wave call it.
__device__ __forceinline__ uint32_t wave_counted_add16(uint32_t* counters, uint32_t key, bool valid) {
  const int lane = __lane_id();
  uint64_t m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const bool bit = ((key >> b) & 1u) != 0;
    const uint64_t bal = __ballot(bit);
    m &= bit ? bal : ~bal;
  }
  const uint64_t lt = (1ull << lane) - 1ull;
  uint32_t old = 0;
  if (valid && (m & lt) == 0) old = atomicAdd(&counters[key], static_cast<uint32_t>(__popcll(m)));
  const int leader = valid ? __builtin_ctzll(m) : lane;
  old = static_cast<uint32_t>(__shfl(static_cast<int>(old), leader, 64));
  return old + static_cast<uint32_t>(__popcll(m & lt));
}

// One workgroup: tile ranges from the totals (exclusive prefix, BUCKET_MAX_TILES tiles per step), the
// tiles per record-count class, and, when `order` is given, the tiles' scheduling order: tiles of
// class c after every tile of the classes before c, within a class in the atomics' order (tiles are
// independent, so the order never changes a pixel).
__global__ void __launch_bounds__(1024) tile_prefix_kernel(const uint32_t* __restrict__ tile_count, uint32_t T,
                                                          uint32_t* __restrict__ ts, uint32_t* __restrict__ te,
                                                          uint32_t* __restrict__ counters, uint32_t* __restrict__ order) {
  constexpr int PER = BUCKET_MAX_TILES / 1024;
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t classes[16];
  const uint32_t tid = threadIdx.x;
  if (tid < 16) classes[tid] = 0;
  uint32_t carry = 0;  // records of the tiles before this step
  for (uint32_t base = 0; base < T; base += BUCKET_MAX_TILES) {
    const uint32_t t0 = base + tid * PER;
    uint32_t c[PER], sum = 0;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      c[i] = load_or_zero(tile_count, t0 + i, T);
      sum += c[i];
    }
    const uint32_t incl = wave_incl_scan(sum);
    if (base) __syncthreads();  // the previous step's wave sums are read
    if ((tid & 63) == 63) wsum[tid >> 6] = incl;
    __syncthreads();
    uint32_t run = carry + incl - sum, step = 0;
    for (uint32_t w = 0; w < 16; ++w) {
      if (w < (tid >> 6)) run += wsum[w];
      step += wsum[w];
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      if (t0 + i < T) {
        ts[t0 + i] = run;
        te[t0 + i] = run + c[i];
      }
      run += c[i];
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) (void)wave_counted_add16(classes, tile_class(c[i]), t0 + i < T);
    carry += step;
  }
  __syncthreads();
  if (tid < 16) counters[CT_CLASS + tid] = classes[tid];
  if (!order) return;
  // classes[] becomes each class's next position
  if (tid == 0) {
    uint32_t b = 0;
    for (int k = 0; k < 16; ++k) {
      const uint32_t n = classes[k];
      classes[k] = b;
      b += n;
    }
  }
  __syncthreads();
  for (uint32_t base = 0; base < T; base += BUCKET_MAX_TILES) {
    const uint32_t t0 = base + tid * PER;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const bool valid = t0 + i < T;
      const uint32_t pos = wave_counted_add16(classes, tile_class(load_or_zero(tile_count, t0 + i, T)), valid);
      if (valid) order[pos] = t0 + i;
    }
  }
}
// The same for targets of at most PREFIX_MAX_TILES tiles, one workgroup per PREFIX_TILES tiles and
// no communication between them: every workgroup reads all T tile totals (8 per thread at 1080p),
// from which it takes the records of the tiles before its own, the tiles of each record-count class
// in the whole target and before its own tiles, and writes its tiles' ranges and scheduling
// positions.  The order is class by class (heaviest first), by tile within a class.
constexpr uint32_t PREFIX_TILES = 1024;
constexpr uint32_t PREFIX_MAX_TILES = 65536;
__global__ void __launch_bounds__(1024) tile_prefix_par_kernel(const uint32_t* __restrict__ tile_count, uint32_t T,
                                                              uint32_t* __restrict__ ts, uint32_t* __restrict__ te,
                                                              uint32_t* __restrict__ counters, uint32_t* __restrict__ order) {
  __shared__ uint32_t w_all[16][16], w_before[16][16], w_mine[16][16];  // [wave][class]
  __shared__ uint32_t w_rec[16], w_scan[16], c_base[16], s_start[16], s_pos[16][16];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, bx = blockIdx.x;
  const uint32_t nj = (T + PREFIX_TILES - 1) / PREFIX_TILES;  // at most 64
  // this lane's tiles per class, one 8-bit field per class (classes 0-7 in [0], 8-15 in [1]): all of
  // them, and those before the workgroup's
  unsigned long long all[2] = {0, 0}, before[2] = {0, 0};
  uint32_t rec_before = 0, mine = 0;
  for (uint32_t j0 = 0; j0 < nj; j0 += 8) {
    uint32_t v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const uint32_t i = (j0 + q) * PREFIX_TILES + tid;
      v[q] = load_or_zero(tile_count, i, T);
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const uint32_t j = j0 + q;  // uniform
      if (j >= nj) break;
      const bool valid = j * PREFIX_TILES + tid < T;
      const uint32_t c = tile_class(v[q]);
      if (j < bx) rec_before += v[q];
      if (j == bx) mine = v[q];
      const unsigned long long one = valid ? 1ull << (8 * (c & 7)) : 0ull;
      all[c >> 3] += one;
      if (j < bx) before[c >> 3] += one;
    }
  }
  const uint32_t t = bx * PREFIX_TILES + tid;
  const bool mine_valid = t < T;
  const uint32_t cm = tile_class(mine);
  // rank among the wave's own tiles of the same class (four ballots, one per class bit)
  uint64_t m = __ballot(mine_valid);
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const bool bit = ((cm >> b) & 1u) != 0;
    const uint64_t bal = __ballot(bit);
    m &= bit ? bal : ~bal;
  }
  const uint32_t rank = static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)));
  unsigned long long mine_f[2] = {0, 0};
  mine_f[cm >> 3] = mine_valid ? 1ull << (8 * (cm & 7)) : 0ull;
  // the wave's per-class sums: the 8-bit fields widened to 16 bits (4 classes per word, at most
  // 64 lanes x 64 chunks per class), added over the wave
  unsigned long long wa[4], wb[4], wm[4];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const unsigned long long lo = 0x00FF00FF00FF00FFull;
    wa[2 * h] = all[h] & lo;
    wa[2 * h + 1] = (all[h] >> 8) & lo;
    wb[2 * h] = before[h] & lo;
    wb[2 * h + 1] = (before[h] >> 8) & lo;
    wm[2 * h] = mine_f[h] & lo;
    wm[2 * h + 1] = (mine_f[h] >> 8) & lo;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      wa[w] += __shfl_xor(wa[w], off, 64);
      wb[w] += __shfl_xor(wb[w], off, 64);
      wm[w] += __shfl_xor(wm[w], off, 64);
    }
  // class k = 8h + 2f + e lives in word 2h + e, 16-bit field f (fields of a word: classes of one parity)
  auto field = [](const unsigned long long (&w)[4], uint32_t k) -> uint32_t {
    const bool h = (k >> 3) != 0, e = (k & 1) != 0;
    const unsigned long long x = h ? (e ? w[3] : w[2]) : (e ? w[1] : w[0]);  // selects, no indexing
    return static_cast<uint32_t>((x >> (16 * ((k & 7) >> 1))) & 0xFFFFull);
  };
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) rec_before += static_cast<uint32_t>(__shfl_xor(static_cast<int>(rec_before), off, 64));
  const uint32_t incl = wave_incl_scan(mine);
  if (lane < 16) {
    w_all[wave][lane] = field(wa, lane);
    w_before[wave][lane] = field(wb, lane);
    w_mine[wave][lane] = field(wm, lane);
  }
  if (lane == 0) w_rec[wave] = rec_before;
  if (lane == 63) w_scan[wave] = incl;
  __syncthreads();
  // Per-wave and per-class bases, each a fixed-length sum over the 16 waves (unrolled, the loads
  // issued together):
  //   c_base[k]     tiles of classes < k in the target + tiles of class k before the workgroup's
  //   s_start[w]    records of the tiles before the workgroup's + of waves < w's tiles
  //   s_pos[w][k]   c_base[k] + this workgroup's class-k tiles in waves < w
  if (tid < 64) {
    const uint32_t k = tid & 15;
    uint32_t tot = 0, bef = 0;
#pragma unroll
    for (int w = 0; w < 16; ++w) {
      tot += w_all[w][k];
      bef += w_before[w][k];
    }
    uint32_t below = tot;  // inclusive scan over the 16 classes, then exclusive
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
      const uint32_t y = static_cast<uint32_t>(__shfl_up(static_cast<int>(below), off, 16));
      if (k >= static_cast<uint32_t>(off)) below += y;
    }
    below -= tot;
    if (tid < 16) {
      c_base[k] = below + bef;
      if (bx == 0) counters[CT_CLASS + k] = tot;
    }
  } else if (tid < 80) {
    const uint32_t w = tid - 64;
    uint32_t rec = 0, pre = 0;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      rec += w_rec[v];
      pre += static_cast<uint32_t>(v) < w ? w_scan[v] : 0u;
    }
    s_start[w] = rec + pre;
  }
  __syncthreads();
  if (tid < 256) {
    const uint32_t w = tid >> 4, k = tid & 15;
    uint32_t pre = 0;
#pragma unroll
    for (int v = 0; v < 16; ++v) pre += static_cast<uint32_t>(v) < w ? w_mine[v][k] : 0u;
    s_pos[w][k] = c_base[k] + pre;
  }
  __syncthreads();
  if (mine_valid) {
    const uint32_t start = s_start[wave] + incl - mine;
    ts[t] = start;
    te[t] = start + mine;
    order[s_pos[wave][cm] + rank] = t;
  }
}
__global__ void __launch_bounds__(256) tile_scatter_kernel(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                           const uint32_t* __restrict__ count, uint32_t cap, uint32_t T,
                                                           const uint32_t* __restrict__ ts, uint32_t* __restrict__ tile_cursor,
                                                           uint32_t* __restrict__ vout) {
  __shared__ uint32_t cnt[BUCKET_MAX_TILES];
  __shared__ uint16_t touched[BUCKET_ITEMS];
  __shared__ uint32_t ntouched;
  const uint32_t n = min(*count, cap);
  const uint32_t b0 = blockIdx.x * BUCKET_ITEMS;
  if (b0 >= n) return;
  uint32_t k[BUCKET_PER_THREAD], rank[BUCKET_PER_THREAD];
  tile_load_keys(keys, n, b0, k);
  for (uint32_t r0 = 0; r0 < T; r0 += BUCKET_MAX_TILES) {
    const uint32_t tn = min(T - r0, BUCKET_MAX_TILES);
    if (r0) __syncthreads();  // the previous range's positions are consumed
    tile_count_lds(k, rank, r0, tn, cnt, touched, &ntouched);
    // this workgroup's range in each touched tile; cnt[t] becomes its first position
    for (uint32_t i = threadIdx.x; i < ntouched; i += 256) {
      const uint32_t t = touched[i];
      cnt[t] = ts[r0 + t] + atomicAdd(&tile_cursor[r0 + t], cnt[t]);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < BUCKET_PER_THREAD; ++j)
      if (k[j] - r0 < tn) vout[cnt[k[j] - r0] + rank[j]] = vals[b0 + j * 256 + threadIdx.x];
  }
}

bool bucket_sort_applicable(const draw_params& P) { return P.gtiles_x > 0 && P.gtiles_y > 0; }
bool launch_bucket_sort(const draw_params& P, const draw_buffers& B, uint32_t cap, uint32_t* order, hipStream_t s) {
  const uint32_t T = static_cast<uint32_t>(P.gtiles_x * P.gtiles_y);
  if (!bucket_sort_applicable(P)) return false;
  const uint32_t* count = B.counters + CT_TOTAL;
  const uint32_t nblocks = std::max<uint32_t>((cap + BUCKET_ITEMS - 1) / BUCKET_ITEMS, 1);
  hipLaunchKernelGGL(tile_count_kernel, dim3(nblocks), dim3(256), 0, s, B.keys[0], count, cap, T, B.tile_count);
  if (order && T <= PREFIX_MAX_TILES)
    hipLaunchKernelGGL(tile_prefix_par_kernel, dim3((T + PREFIX_TILES - 1) / PREFIX_TILES), dim3(1024), 0, s, B.tile_count, T,
                       B.tstart, B.tend, B.counters, order);
  else
    hipLaunchKernelGGL(tile_prefix_kernel, dim3(1), dim3(1024), 0, s, B.tile_count, T, B.tstart, B.tend, B.counters, order);
  hipLaunchKernelGGL(tile_scatter_kernel, dim3(nblocks), dim3(256), 0, s, B.keys[0], B.vals[0], count, cap, T, B.tstart,
                     B.tile_count + T, B.vals[1]);
  return true;
}

__global__ void tile_range_kernel(const uint32_t* __restrict__ keys, uint32_t n, uint32_t T, uint32_t* __restrict__ ts,
                                  uint32_t* __restrict__ te) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // T == 0: the ranges were cleared beforehand, only the non-empty tiles are written
  uint32_t k = keys[i];
  if (i == 0 && T) {
    for (uint32_t t = 0; t < k; ++t) ts[t] = te[t] = 0;
  }
  if (i == 0 || keys[i - 1] != k) ts[k] = i;
  if (i == n - 1 || keys[i + 1] != k) {
    te[k] = i + 1;
    const uint32_t next = !T ? k + 1 : ((i == n - 1) ? T : keys[i + 1]);
    for (uint32_t t = k + 1; t < next; ++t) ts[t] = te[t] = i + 1;
  }
}

// Scheduling order of the GPU tiles, heaviest first so that the longest tiles start early.  Tiles
// are independent, so the order never changes a pixel.  One workgroup buckets the tiles by record
// count / 8 (heaviest bucket first) with LDS counters.
__device__ __forceinline__ uint32_t tile_bucket(const uint32_t* ts, const uint32_t* te, uint32_t t) {
  return 255u - min((te[t] - ts[t]) >> 3, 255u);
}
__device__ __forceinline__ void tile_order_body(const uint32_t* __restrict__ ts, const uint32_t* __restrict__ te,
                                                uint32_t T, uint32_t* __restrict__ order) {
  __shared__ uint32_t cnt[256];
  __shared__ uint32_t wtot[4];
  const uint32_t tid = threadIdx.x;
  if (tid < 256) cnt[tid] = 0;
  __syncthreads();
  for (uint32_t t = tid; t < T; t += blockDim.x) (void)wave_counted_add(cnt, tile_bucket(ts, te, t));
  __syncthreads();
  // exclusive scan of the 256 bucket counts by the first four waves
  uint32_t c = 0, incl = 0;
  if (tid < 256) {
    c = cnt[tid];
    incl = wave_incl_scan(c);
    if ((tid & 63) == 63) wtot[tid >> 6] = incl;
  }
  __syncthreads();
  if (tid < 256) {
    uint32_t before = 0;
    for (uint32_t w = 0; w < (tid >> 6); ++w) before += wtot[w];
    cnt[tid] = before + incl - c;
  }
  __syncthreads();
  for (uint32_t t = tid; t < T; t += blockDim.x) order[wave_counted_add(cnt, tile_bucket(ts, te, t))] = t;
}
__global__ void __launch_bounds__(1024) tile_order_kernel(const uint32_t* __restrict__ ts, const uint32_t* __restrict__ te,
                                                          uint32_t T, uint32_t* __restrict__ order) {
  tile_order_body(ts, te, T, order);
}

__global__ void tile_class_count_kernel(const uint32_t* __restrict__ ts, const uint32_t* __restrict__ te, uint32_t T,
                                        uint32_t* __restrict__ counters) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < T) (void)wave_counted_add(counters + CT_CLASS, tile_class(te[t] - ts[t]));
}
// Tiles of class c take the positions after every tile of the classes before c; within a class the
// order is the atomics' (tiles are independent, so the order never changes a pixel).
__global__ void tile_class_order_kernel(const uint32_t* __restrict__ ts, const uint32_t* __restrict__ te, uint32_t T,
                                        uint32_t* __restrict__ counters, uint32_t* __restrict__ order) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const uint32_t c = tile_class(te[t] - ts[t]);
  uint32_t base = 0;
  for (uint32_t k = 0; k < c; ++k) base += counters[CT_CLASS + k];
  order[base + wave_counted_add(counters + CT_CURSOR, c)] = t;
}
void launch_tile_order(const draw_params& P, const draw_buffers& B, uint32_t* order, bool counted, hipStream_t s) {
  const uint32_t T = static_cast<uint32_t>(P.gtiles_x * P.gtiles_y);
  if (!T) return;
  if (!counted) hipLaunchKernelGGL(tile_class_count_kernel, dim3((T + 255) / 256), dim3(256), 0, s, B.tstart, B.tend, T, B.counters);
  hipLaunchKernelGGL(tile_class_order_kernel, dim3((T + 255) / 256), dim3(256), 0, s, B.tstart, B.tend, T, B.counters, order);
}

// records per workgroup beyond which a tile of a depth-only draw is split
constexpr uint32_t VIS_PART = 256;

// Work table of the order-independent visibility kernel for draws without shading (depth-only
// passes), whose time is the longest tile's: a GPU tile with more than VIS_PART records is split
// into parts of at most VIS_PART records, each a workgroup of the first launch that leaves its
// partial per-sample keys in `partial`; the second launch merges a split tile's parts (a minimum,
// so the order of the parts is immaterial) and finishes it.  (Draws that shade keep whole tiles:
// their shading fills the GPU while the long tiles run.)  Entries follow `order`, heaviest first.
//   entries[i] = (tile, part, parts, first partial slot)   i < counts[0]
//   merges[i]  = (tile, parts, first partial slot, 0)      i < counts[1]
__global__ void __launch_bounds__(1024) vis_table_kernel(const uint32_t* __restrict__ ts, const uint32_t* __restrict__ te,
                                                         const uint32_t* __restrict__ order, uint32_t T,
                                                         uint4* __restrict__ entries, uint4* __restrict__ merges,
                                                         uint32_t* __restrict__ counts) {
  __shared__ uint32_t wtot[3][16];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t per = (T + 1023) / 1024, p0 = tid * per, p1 = min(T, p0 + per);
  uint32_t c[3] = {0, 0, 0};  // blocks, partial slots, merges in this thread's segment of `order`
  for (uint32_t p = p0; p < p1; ++p) {
    const uint32_t t = order[p], n = te[t] - ts[t];
    const uint32_t parts = n > VIS_PART ? (n + VIS_PART - 1) / VIS_PART : 1u;
    c[0] += parts;
    if (parts > 1) { c[1] += parts; c[2] += 1; }
  }
  uint32_t incl[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    incl[k] = wave_incl_scan(c[k]);
    if (lane == 63) wtot[k][wave] = incl[k];
  }
  __syncthreads();
  uint32_t base[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    uint32_t before = 0, total = 0;
    for (uint32_t w = 0; w < 16; ++w) {
      if (w < wave) before += wtot[k][w];
      total += wtot[k][w];
    }
    base[k] = before + incl[k] - c[k];
    if (tid == 0) counts[k == 0 ? 0 : (k == 1 ? 2 : 1)] = total;
  }
  uint32_t b = base[0], sl = base[1], m = base[2];
  for (uint32_t p = p0; p < p1; ++p) {
    const uint32_t t = order[p], n = te[t] - ts[t];
    const uint32_t parts = n > VIS_PART ? (n + VIS_PART - 1) / VIS_PART : 1u;
    for (uint32_t k = 0; k < parts; ++k) entries[b + k] = make_uint4(t, k, parts, sl);
    b += parts;
    if (parts > 1) {
      merges[m++] = make_uint4(t, parts, sl, 0u);
      sl += parts;
    }
  }
}

void launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s) {
  if (n) hipLaunchKernelGGL(fill_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
}

void launch_tile_ranges(const draw_params& P, const draw_buffers& B, int which, uint32_t n, hipStream_t s) {
  uint32_t T = static_cast<uint32_t>(P.gtiles_x * P.gtiles_y);
  if (!T) return;
  if (n < 4 * T) {  // sparse records: clear every range first, so no thread walks a long empty run
    hipLaunchKernelGGL(fill_u32_kernel, dim3((T + 255) / 256), dim3(256), 0, s, B.tstart, T, 0u);
    hipLaunchKernelGGL(fill_u32_kernel, dim3((T + 255) / 256), dim3(256), 0, s, B.tend, T, 0u);
    if (n) hipLaunchKernelGGL(tile_range_kernel, dim3((n + 255) / 256), dim3(256), 0, s, B.keys[which], n, 0u, B.tstart, B.tend);
    return;
  }
  hipLaunchKernelGGL(tile_range_kernel, dim3((n + 255) / 256), dim3(256), 0, s, B.keys[which], n, T, B.tstart, B.tend);
}

// ============================================================== pixel shaders
__device__ __forceinline__ void normalize3(float x, float y, float z, float& ox, float& oy, float& oz) {
  float t = 0.0f;
  t += x * x;
  t += y * y;
  t += z * z;
  float len = sqrtf(t);
  if (equal_eps(len, 0.0f)) len = 1.0f;
  float inv = 1.0f / len;
  ox = x * inv;
  oy = y * inv;
  oz = z * inv;
}

// 3x3 percentage-closer shadow lookup of one pixel (attr0 colour, attr1 light clip position): the
// tap layout of StandardShadowMap.cpp:86-106, each tap a tex2dlod at lod 0, lit when the light-space
// depth minus the bias does not exceed it.
__device__ __forceinline__ float4 shadow_pcf(const draw_params& P, const ps_samplers* __restrict__ SM, float4 a, float4 l) {
  const float lx = l.x / l.w, ly = l.y / l.w, lz = l.z / l.w;
  const float cu = (lx + 1.0f) * 0.5f, cv = 1.0f - (ly + 1.0f) * 0.5f;
  const float o = P.ps_consts[2];
  int count = 0;
  if (P.has_sampler[0]) {
    for (int ty = -1; ty <= 1; ++ty)
      for (int tx = -1; tx <= 1; ++tx) {
        const float u = cu + static_cast<float>(tx) * o, v = cv + static_cast<float>(ty) * o;
        const float4 t = sample_lod0(SM->s[0], u, v);
        if (lz - P.ps_consts[1] <= t.x) ++count;
      }
  }
  const float lit = static_cast<float>(count) / 9.0f;
  const float k = P.ps_consts[0] + (1.0f - P.ps_consts[0]) * lit;
  return make_float4(a.x * k, a.y * k, a.z * k, a.w);
}

// cpp_pixel_shader::tex2d(s, reg) (cpp_pixel_shader.cpp:13-31): one LOD per quad from the register's
// p1 - p0 and p2 - p0 (calc_lod_2d, sampler.cpp:834-852), then sample_2d_lod at pixel pi
template <int NA>
__device__ __forceinline__ float4 tex2d_cpp(const sampler_desc_dev& sp, const quad_lazy<NA>& q, int reg, int pi) {
  const float4 u0 = q.reg(reg, 0, q.inv_w(0), 2), u1 = q.reg(reg, 1, q.inv_w(1), 2), u2 = q.reg(reg, 2, q.inv_w(2), 2);
  const float dxu = u1.x - u0.x, dxv = u1.y - u0.y;
  const float dyu = u2.x - u0.x, dyv = u2.y - u0.y;
  const float lod = (sp.mip_filter == FILTER_ANISO && sp.max_aniso > 1) ? calc_aniso(sp, dxu, dxv, dyu, dyv, 0.0f).lod
                                                                        : calc_lod(sp, dxu, dxv, dyu, dyv, 0.0f);
  const float4 uv = q.reg(reg, pi, q.inv_w(pi), 2);
  return sample_impl(sp, uv.x, uv.y, lod, nullptr);
}
// eflib vector helpers on .xyz (vector_generic.h:110-118, math.cpp:39-41,85-87)
__device__ __forceinline__ float dot3(float4 a, float4 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float length3(float4 a) {
  float t = 0.0f;
  t += a.x * a.x;
  t += a.y * a.y;
  t += a.z * a.z;
  return sqrtf(t);
}
__device__ __forceinline__ float4 normalize3v(float4 v) {
  float x, y, z;
  normalize3(v.x, v.y, v.z, x, y, z);
  return make_float4(x, y, z, 0.0f);
}
__device__ __forceinline__ float4 reflect3(float4 i, float4 n) {
  const float k = 2.0f * dot3(i, n);
  return make_float4(i.x - n.x * k, i.y - n.y * k, i.z - n.z * k, 0.0f);
}

// ColorizedTriangle.cpp:55-92: three lights, attr0 normal, attr1..3 light vectors
template <int NA>
__device__ __forceinline__ float4 ps_colorized(const quad_lazy<NA>& q, int pi) {
  const float iw = q.inv_w(pi);
  const float4 n = q.reg(1, pi, iw, 3);
  float4 ld[3];
  float inv[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    ld[i] = q.reg(2 + i, pi, iw, 3);
    inv[i] = 1.0f / length3(ld[i]);
  }
  const float4 nn = normalize3v(n);
  float refl[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) refl[i] = dot3(nn, make_float4(ld[i].x * inv[i], ld[i].y * inv[i], ld[i].z * inv[i], 0.0f));
  const float c[3][3] = {{0.7f, 0.1f, 0.3f}, {0.1f, 0.3f, 0.7f}, {0.3f, 0.7f, 0.1f}};
  float o[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float a = ((c[0][k] * refl[0]) * inv[0]) * inv[0];
    float b = ((c[1][k] * refl[1]) * inv[1]) * inv[1];
    float d = ((c[2][k] * refl[2]) * inv[2]) * inv[2];
    o[k] = std_clamp((a + b) + d, 0.0f, 1.0f);
  }
  return make_float4(o[0], o[1], o[2], 1.0f);
}

// ComplexMesh.cpp:227-246
template <int NA>
__device__ __forceinline__ float4 ps_mesh(const quad_lazy<NA>& q, int pi) {
  const float iw = q.inv_w(pi);
  const float4 l = normalize3v(q.reg(1, pi, iw, 3)), e = normalize3v(q.reg(2, pi, iw, 3)), n = normalize3v(q.reg(3, pi, iw, 3));
  const float n_dot_l = dot3(n, l);
  const float4 h = normalize3v(make_float4(l.x + e.x, l.y + e.y, l.z + e.z, 0.0f));
  const float spec = 3.5f * powf(std_max(dot3(h, n), 0.0f), 5.0f);  // (roughness + 2) / 2 with roughness 5
  const float clr = n_dot_l * (0.8f + spec * 0.4f);
  return make_float4(clr, clr, clr, 1.0f);
}

// StandardShadowMap.cpp:62-142 (exponential shadow map, esm_constant 25000, 3x3 Gaussian weights)
template <int NA>
__device__ __forceinline__ float4 ps_ssm(const draw_params& P, const ps_samplers* __restrict__ SM, const quad_lazy<NA>& q,
                                         int pi) {
  const float iw = q.inv_w(pi);
  float occlusion = 0.0f;
  if (P.has_sampler[0]) {
    const float4 l4 = q.reg(5, pi, iw);
    const float lx = l4.x / l4.w, ly = l4.y / l4.w, lz = l4.z / l4.w;
    const float cx = (lx + 1.0f) * 0.5f, cy = 1.0f - (ly + 1.0f) * 0.5f;
    const float o = 1 / 512.0f;
    const float ox[9] = {-o, 0.0f, +o, -o, 0.0f, +o, -o, 0.0f, +o};
    const float oy[9] = {-o, -o, -o, 0.0f, 0.0f, 0.0f, +o, +o, +o};
    const float gw[9] = {0.027681f, 0.111014f, 0.027681f, 0.111014f, 0.445213f, 0.111014f, 0.027681f, 0.111014f, 0.027681f};
    float sd[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) sd[i] = sample_lod0(SM->s[0], cx + ox[i], cy + oy[i]).x;
    const float esm = 25000.0f;
    float occ = 0.0f;
#pragma unroll
    for (int i = 1; i < 9; ++i) occ += gw[i] * expf(esm * (sd[i] - sd[0]));
    occ += gw[0];
    occ = logf(occ);
    occ += esm * sd[0];
    occlusion = std_clamp(expf(occ - esm * lz), 0.0f, 1.0f);
  }
  float4 tex = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
  if (P.has_sampler[1]) tex = tex2d_cpp(SM->s[1], q, 1, pi);
  const float4 n = normalize3v(q.reg(2, pi, iw, 3)), ld = normalize3v(q.reg(3, pi, iw, 3)), ed = normalize3v(q.reg(4, pi, iw, 3));
  const float idiff = std_clamp(dot3(ld, n), 0.0f, 1.0f);
  const float4 r = reflect3(ld, n);
  const float ispec = std_clamp(dot3(make_float4(-r.x, -r.y, -r.z, 0.0f), ed), 0.0f, 1.0f);
  // pow(float, int) promotes to double (C++11 <cmath>), then the vec4 product rounds it to float
  const float p = static_cast<float>(pow(static_cast<double>(ispec), static_cast<double>(__float_as_int(P.ps_consts[12]))));
  const float* amb = P.ps_consts;
  const float* dif = P.ps_consts + 4;
  const float* spc = P.ps_consts + 8;
  float4 out;
  out.x = tex.x * (amb[0] + (dif[0] * idiff + spc[0] * p) * occlusion);
  out.y = tex.y * (amb[1] + (dif[1] * idiff + spc[1] * p) * occlusion);
  out.z = tex.z * (amb[2] + (dif[2] * idiff + spc[2] * p) * occlusion);
  out.w = 1.0f;
  return out;
}

// StencilMirror.cpp:66-110
template <int NA>
__device__ __forceinline__ float4 ps_cup(const draw_params& P, const ps_samplers* __restrict__ SM, const quad_lazy<NA>& q,
                                         int pi) {
  const float iw = q.inv_w(pi);
  float4 tex = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
  if (P.has_sampler[0]) tex = tex2d_cpp(SM->s[0], q, 1, pi);
  const float4 n = normalize3v(q.reg(2, pi, iw, 3)), ld = normalize3v(q.reg(3, pi, iw, 3)), ed = normalize3v(q.reg(4, pi, iw, 3));
  const float idiff = std_clamp(dot3(ld, n), 0.0f, 1.0f);
  const float ispec = std_clamp(dot3(reflect3(ld, n), ed), 0.0f, 1.0f);
  const float* amb = P.ps_consts;
  const float* dif = P.ps_consts + 4;
  const float* spc = P.ps_consts + 8;
  float4 out;
  out.x = tex.x * ((amb[0] + dif[0] * idiff) + spc[0] * ispec);
  out.y = tex.y * ((amb[1] + dif[1] * idiff) + spc[1] * ispec);
  out.z = tex.z * ((amb[2] + dif[2] * idiff) + spc[2] * ispec);
  out.w = 1.0f;
  return out;
}

// Pixel shader (cpp_pixel_shader::execute, cpp_pixel_shader.cpp:62-74) colour of quad pixel pi
// only; the other three pixels supply derivatives.  The sample shaders never discard.
template <int NA>
__device__ __forceinline__ float4 run_ps_pixel(const draw_params& P, const ps_samplers* __restrict__ SM,
                                               const quad_lazy<NA>& q, int pi) {
  switch (P.ps_id) {
  case PS_GOURAUD:
    return q.reg(1, pi, q.inv_w(pi));
  case PS_CONSTANT:
    return make_float4(P.ps_consts[0], P.ps_consts[1], P.ps_consts[2], P.ps_consts[3]);
  case PS_LAMBERT_TEX:
    if constexpr (NA >= 2) {
      // SASL tex2D: ddx per row (p1-p0, p3-p2), ddy per column (p2-p0, p3-p1)
      // (sasl/src/codegen/cgs_simd.cpp:275-312), then sample_2d_grad (sampler_api.cpp:13-18);
      // uv at pi and its row and column neighbours
      const int pr = pi ^ 1, pc = pi ^ 2;
      const float iw = q.inv_w(pi), iwr = q.inv_w(pr), iwc = q.inv_w(pc);
      const float4 uv = q.reg(2, pi, iw, 2), uvr = q.reg(2, pr, iwr, 2), uvc = q.reg(2, pc, iwc, 2);
      const bool col1 = (pi & 1) != 0, row1 = (pi & 2) != 0;
      // row: (left, right) = col1 ? (neighbour, pi) : (pi, neighbour); column likewise
      const float4 r0 = sel4(col1, uvr, uv), r1 = sel4(col1, uv, uvr);
      const float4 c0 = sel4(row1, uvc, uv), c1 = sel4(row1, uv, uvc);
      const float dxu = r1.x - r0.x, dxv = r1.y - r0.y;
      const float dyu = c1.x - c0.x, dyv = c1.y - c0.y;
      const float4 nrm = q.reg(1, pi, iw, 3);
      float4 t = make_float4(0, 0, 0, 0);
      if (P.has_sampler[0]) {
        const sampler_desc_dev& sp = SM->s[0];
        t = sp.fast_aniso ? sample_2d_grad_fast(sp, uv.x, uv.y, dxu, dxv, dyu, dyv, 0.0f)
                          : sample_2d_grad(sp, uv.x, uv.y, dxu, dxv, dyu, dyv, 0.0f);
      }
      float nx, ny, nz;
      normalize3(nrm.x, nrm.y, nrm.z, nx, ny, nz);
      float d = nx * P.ps_consts[0] + ny * P.ps_consts[1] + nz * P.ps_consts[2];
      float k = std_clamp(d, 0.0f, 1.0f);
      return make_float4(t.x * k, t.y * k, t.z * k, 1.0f);
    }
    return make_float4(0, 0, 0, 0);
  case PS_TEX_CPP: {
    // AnisotropicFilter / TextureAndBlending ps_plane: tex2d(s0, attr0) with alpha 1
    if (!P.has_sampler[0]) return make_float4(0, 0, 0, 1.0f);
    float4 t = tex2d_cpp(SM->s[0], q, 1, pi);
    t.w = 1.0f;
    return t;
  }
  case PS_COLORIZED:
    if constexpr (NA >= 4) return ps_colorized(q, pi);
    return make_float4(0, 0, 0, 0);
  case PS_MESH:
    if constexpr (NA >= 3) return ps_mesh(q, pi);
    return make_float4(0, 0, 0, 0);
  case PS_SSM:
    if constexpr (NA >= 5) return ps_ssm(P, SM, q, pi);
    return make_float4(0, 0, 0, 0);
  case PS_CUP:
    if constexpr (NA >= 4) return ps_cup(P, SM, q, pi);
    return make_float4(0, 0, 0, 0);
  case PS_BOX: {  // TextureAndBlending.cpp:96-114
    if constexpr (NA >= 2) {
      if (!P.has_sampler[0]) return make_float4(1.0f, 1.0f, 1.0f, 0.5f);
      float4 t = tex2d_cpp(SM->s[0], q, 2, pi);
      t.w = 0.5f;
      return t;
    }
    return make_float4(0, 0, 0, 0);
  }
  case PS_SHADOW_PCF:
    if constexpr (NA >= 2) {
      const float iw = q.inv_w(pi);
      return shadow_pcf(P, SM, q.reg(1, pi, iw), q.reg(2, pi, iw));
    }
    return make_float4(0, 0, 0, 0);
  default:
    return make_float4(0, 0, 0, 0);
  }
}

// ============================================================== raster
struct prim_lds {
  float e[9];
  float bb[4];     // relative to the 64x64 tile origin
  float ev1[3];    // stage-16 rejection value, already offset by the 64x64 tile origin
  float s1x[3], s1y[3], ra1[3];
  float ev2[3], s2x[3], s2y[3], ra2[3];
  float zx, zy, z0, dzx, dzy;
  float aa[4];
  uint32_t idf;    // triangle id << 1 | tile-level accept
  uint32_t r16;    // bit k: 16x16 region k of the GPU tile rejected; bit 16+k: accepted
  uint32_t front;
};

__device__ __forceinline__ bool cmpf(int f, float a, float b) {
  switch (f) {
  case 0: return false;
  case 1: return a < b;
  case 2: return a == b;
  case 3: return a <= b;
  case 4: return a > b;
  case 5: return a != b;
  case 6: return a >= b;
  default: return true;
  }
}
__device__ __forceinline__ bool cmpu(int f, uint32_t a, uint32_t b) {
  switch (f) {
  case 0: return false;
  case 1: return a < b;
  case 2: return a == b;
  case 3: return a <= b;
  case 4: return a > b;
  case 5: return a != b;
  case 6: return a >= b;
  default: return true;
  }
}
// framebuffer.cpp:136-166
__device__ __forceinline__ uint32_t stencil_op(int op, uint32_t ref, uint32_t cur) {
  switch (op) {
  case SOP_ZERO: return 0;
  case SOP_REPLACE: return ref;
  case SOP_INCR_SAT: return min(0xFFu, cur + 1);
  case SOP_DECR_SAT: return cur - 1;  // std::max<uint32_t>(0, cur - 1) never clamps
  case SOP_INVERT: return ~cur;
  case SOP_INCR_WRAP: return (cur + 1) & 0xFF;
  case SOP_DECR_WRAP: return (cur - 1 + 256) & 0xFF;
  default: return cur;
  }
}


__device__ __forceinline__ void blend_store(const draw_params& P, int x, int y, int smp, float4 src) {
  if (!P.color.data) return;
  uint8_t* a = sample_addr(P.color, x, y, smp);
  if (P.blend_id == BLEND_OFF) {
    store_texel(P.color.pf, a, src);
  } else if (P.blend_id == BLEND_ON) {
    store_texel(P.color.pf, a, lerp4(load_texel(P.color.pf, a), src, src.w));
  }
}


// Pixel shader and backend for one triangle on a 2x2 quad: `pm` holds the samples to store, bits
// 4i..4i+S-1 for pixel i.  Interpolants are formed per pixel where read (quad_lazy), so only one
// pixel's registers are live at a time; the colour of each pixel is the quad shader's.
template <int NA, int S, bool SASL = false>
__device__ __forceinline__ void shade_store(const draw_params& P, const ps_samplers* __restrict__ SM,
                                         const float4* __restrict__ tri_attr, uint32_t id, int px0, int py0, uint32_t pm) {
  attr_setup<NA> at;
  load_attr<NA>(P, tri_attr, id, at);
  const quad_lazy<NA> q(at, P, px0, py0);
  if constexpr (SASL) {
    // pixel_shader_unit::update + execute (shader_unit.cpp:106-176) on the whole quad: input i is
    // attribute inputs[i].attr of each pixel, the program runs the four pixels together
    const sasl_program* __restrict__ prog = P.sasl;
    sasl_regs R;
    const uint32_t nregs = prog->nregs;
    for (uint32_t r = 0; r < nregs; ++r)
      for (int px = 0; px < 4; ++px) R.r[r][px] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const uint32_t nin = prog->ninputs;
    for (int px = 0; px < 4; ++px) {
      const float iw = q.inv_w(px);
      for (uint32_t k = 0; k < nin; ++k) {
        const sasl_input inp = prog->inputs[k];
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int a = 0; a < NA; ++a)
          if (a == inp.attr) v = q.reg(a + 1, px, iw);
        R.r[inp.reg][px] = make_float4(v.x, inp.ncomp > 1 ? v.y : 0.0f, inp.ncomp > 2 ? v.z : 0.0f, inp.ncomp > 3 ? v.w : 0.0f);
      }
    }
    sasl_exec(prog, R, [&](int slot, int, float u, float v, float dxu, float dxv, float dyu, float dyv) -> float4 {
      if (!P.has_sampler[slot]) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      const sampler_desc_dev& sp = SM->s[slot];
      return sp.fast_aniso ? sample_2d_grad_fast(sp, u, v, dxu, dxv, dyu, dyv, 0.0f)
                           : sample_2d_grad(sp, u, v, dxu, dxv, dyu, dyv, 0.0f);
    });
    for (int i = 0; i < 4; ++i) {
      const uint32_t m = (pm >> (4 * i)) & 15u;
      if (!m) continue;
      const float4 c = sasl_color0(prog, R, i);
      for (int s = 0; s < S; ++s)
        if (m & (1u << s)) blend_store(P, px0 + (i & 1), py0 + (i >> 1), s, c);
    }
    return;
  }
#pragma unroll 1
  for (int i = 0; i < 4; ++i) {
    const uint32_t m = (pm >> (4 * i)) & 15u;
    if (!m) continue;
    const float4 c = run_ps_pixel<NA>(P, SM, q, i);
    for (int s = 0; s < S; ++s)
      if (m & (1u << s)) blend_store(P, px0 + (i & 1), py0 + (i >> 1), s, c);
  }
}

// One workgroup per GPU tile, one thread per 2x2 quad.  VIS: opaque early-Z draws -- record for every
// sample the last triangle that passed (the colour it ends with) into `vis`; the pixel shader runs
// later in shade_kernel.  Otherwise every quad runs the pixel shader and the backend in primitive order.
template <int NA, int S, bool VIS, bool SASL = false>
__global__ void __launch_bounds__(256) raster_kernel(draw_params P, const ps_samplers* __restrict__ SM,
                                                      const uint32_t* __restrict__ order, const uint32_t* __restrict__ vals,
                                                      const uint32_t* __restrict__ tstart, const uint32_t* __restrict__ tend,
                                                      const float4* __restrict__ tri_rec, const float4* __restrict__ tri_attr,
                                                      uint32_t* __restrict__ vis, unsigned long long* __restrict__ partial) {
  extern __shared__ prim_lds lds[];
  const int G = P.gt, QW = G / 2;
  const int gtile = order[blockIdx.x];
  const int gx = gtile % P.gtiles_x, gy = gtile / P.gtiles_x;
  const int tid = threadIdx.x, nthreads = blockDim.x;
  const int qx = tid % QW, qy = tid / QW;
  const int px0 = gx * G + qx * 2, py0 = gy * G + qy * 2;  // quad origin
  const int t64x = (gx * G) & ~(TILE - 1), t64y = (gy * G) & ~(TILE - 1);
  const int tw = static_cast<int>(P.target_w), th = static_cast<int>(P.target_h);
  // pixel i of the quad: (px0 + (i & 1), py0 + (i >> 1))
  bool pin[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) pin[i] = (px0 + (i & 1)) < tw && (py0 + (i >> 1)) < th;

  // depth/stencil of the quad's samples, held in registers for the whole tile
  // VIS draws have stencil off, so only depth is kept (and written back)
  float depth[4][S];
  uint32_t stencil[4][VIS ? 1 : S];
  uint32_t winner[4][VIS ? S : 1];
  const bool has_ds = P.ds.data != nullptr;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int s = 0; s < S; ++s) {
      depth[i][s] = 0.0f;
      if constexpr (!VIS) stencil[i][s] = 0;
      if constexpr (VIS) winner[i][s] = NO_WINNER;
      if (has_ds && pin[i]) {
        if constexpr (VIS) {
          depth[i][s] = *reinterpret_cast<const float*>(sample_addr(P.ds, px0 + (i & 1), py0 + (i >> 1), s));
        } else {
          float2 d = *reinterpret_cast<const float2*>(sample_addr(P.ds, px0 + (i & 1), py0 + (i >> 1), s));
          depth[i][s] = d.x;
          stencil[i][s] = __float_as_uint(d.y);
        }
      }
    }
  if constexpr (!VIS) {
    // Stencil changes only where the stencil test passes, so a tile none of whose samples passes it
    // now can never pass it during this draw: nothing in the tile changes (exact early-out, taken
    // only when no statistics are counted).
    if (P.stencil_enable && !P.count_stats) {
      int can = 0;
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int s = 0; s < S; ++s)
          if (pin[i]) {
            const uint32_t os = stencil[i][s] & P.read_mask;
            can |= cmpu(P.st_func[0], P.stencil_ref, os) || cmpu(P.st_func[1], P.stencil_ref, os);
          }
      if (!__syncthreads_or(can)) return;
    }
  }
  const bool any_in = pin[0];
  unsigned long long ps_inv = 0, be_pix = 0;
  const uint32_t full_px = (S == 1) ? 1u : (S == 2 ? 3u : 15u);

  const uint32_t t0 = tstart[gtile], t1 = tend[gtile];
  // 16x16, 4x4 regions of this quad inside its 64x64 tile
  const int rx64 = px0 - t64x, ry64 = py0 - t64y;
  const int ix16 = rx64 >> 4, iy16 = ry64 >> 4;
  const int r16k = ((py0 - gy * G) >> 4) * (G / 16) + ((px0 - gx * G) >> 4);
  const int R16x = t64x + ix16 * 16, R16y = t64y + iy16 * 16;
  const int ix4 = (px0 - R16x) >> 2, iy4 = (py0 - R16y) >> 2;
  const int Bx = R16x + ix4 * 4, By = R16y + iy4 * 4;

  for (uint32_t base = t0; base < t1; base += nthreads) {
    uint32_t nb = min(static_cast<uint32_t>(nthreads), t1 - base);
    if (tid < static_cast<int>(nb)) {
      prim_lds& L = lds[tid];
      uint32_t rec = vals[base + tid];
      uint32_t id = rec >> 1;
      const float4* tr = tri_rec + tri_slot(P, id) * 5;
      float e[9], bb[4];
      read_tri_rec(tr, e, bb);
      float4 d3 = tr[3], d4 = tr[4];
      for (int k = 0; k < 9; ++k) L.e[k] = e[k];
      const float tvx = static_cast<float>(t64x), tvy = static_cast<float>(t64y);
      L.bb[0] = bb[0] - tvx;
      L.bb[1] = bb[1] - tvx;
      L.bb[2] = bb[2] - tvy;
      L.bb[3] = bb[3] - tvy;
      // per-stage constants (rasterize_triangle, rasterizer.cpp:662-675, 704-710)
      for (int k = 0; k < 3; ++k) {
        float ex = e[k * 3], ey = e[k * 3 + 1], ez = e[k * 3 + 2];
        bool mx = ex > 0, my = ey > 0;
        float sx = TILE * ex, sy = TILE * ey;
        float ra = -fabsf(sx) - fabsf(sy);
        float part = static_cast<float>(mx * TILE) * ex + static_cast<float>(my * TILE) * ey;
        sx *= 0.25f; sy *= 0.25f; ra *= 0.25f; part *= 0.25f;
        float ev = ez - part;
        L.s1x[k] = sx; L.s1y[k] = sy; L.ra1[k] = ra;
        L.ev1[k] = ev - ((tvx * ex) + (tvy * ey));
        sx *= 0.25f; sy *= 0.25f; ra *= 0.25f; part *= 0.25f;
        L.s2x[k] = sx; L.s2y[k] = sy; L.ra2[k] = ra;
        L.ev2[k] = ez - part;
      }
      L.zx = d3.y;
      L.zy = d3.z;
      L.z0 = d3.w;
      L.dzx = d4.x;
      L.dzy = d4.y;
      L.front = d4.z != 0.0f;
      for (int s = 0; s < 4; ++s) L.aa[s] = (S > 1 && s < S) ? (P.sp[s][0] - 0.5f) * d4.x + (P.sp[s][1] - 0.5f) * d4.y : 0.0f;
      L.idf = rec;
      // stage-16 decision for each 16x16 region of the GPU tile (subdivide_tile, SSE order)
      uint32_t bits = 0;
      int nreg = G / 16;
      for (int ry = 0; ry < nreg; ++ry)
        for (int rx = 0; rx < nreg; ++rx) {
          int jx = ((gx * G - t64x) >> 4) + rx, jy = ((gy * G - t64y) >> 4) + ry;
          bool rej = false, nacc = false;
          for (int k = 0; k < 3; ++k) {
            float r = (L.s1x[k] * static_cast<float>(jx)) + (L.s1y[k] * static_cast<float>(jy));
            float a = r + L.ra1[k];
            rej |= r < L.ev1[k];
            nacc |= a < L.ev1[k];
          }
          int mxr = jx * 16, myr = jy * 16;
          rej |= L.bb[0] >= static_cast<float>(mxr + 16);
          rej |= L.bb[1] < static_cast<float>(mxr);
          rej |= L.bb[2] >= static_cast<float>(myr + 16);
          rej |= L.bb[3] < static_cast<float>(myr);
          int k = ry * nreg + rx;
          if (rej) bits |= 1u << k;
          else if (!nacc) bits |= 1u << (16 + k);
        }
      L.r16 = bits;
    }
    __syncthreads();
    if (any_in) {
      for (uint32_t j = 0; j < nb; ++j) {
        const prim_lds& L = lds[j];
        uint32_t cov[4];
        if (L.idf & 1) {
          cov[0] = cov[1] = cov[2] = cov[3] = full_px;
        } else {
          uint32_t bits = L.r16;
          if (bits & (1u << r16k)) continue;
          if (bits & (1u << (16 + r16k))) {
            cov[0] = cov[1] = cov[2] = cov[3] = full_px;
          } else {
            // stage 4: child (ix4, iy4) of the 16x16 region at (R16x, R16y)
            bool rej = false, nacc = false;
            const float lf = static_cast<float>(R16x), tf = static_cast<float>(R16y);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              float ev = L.ev2[k] - ((lf * L.e[k * 3]) + (tf * L.e[k * 3 + 1]));
              float r = (L.s2x[k] * static_cast<float>(ix4)) + (L.s2y[k] * static_cast<float>(iy4));
              float a = r + L.ra2[k];
              rej |= r < ev;
              nacc |= a < ev;
            }
            int mxr = ix16 * 16 + ix4 * 4, myr = iy16 * 16 + iy4 * 4;
            rej |= L.bb[0] >= static_cast<float>(mxr + 4);
            rej |= L.bb[1] < static_cast<float>(mxr);
            rej |= L.bb[2] >= static_cast<float>(myr + 4);
            rej |= L.bb[3] < static_cast<float>(myr);
            if (rej) continue;
            if (!nacc) {
              cov[0] = cov[1] = cov[2] = cov[3] = full_px;
            } else {
              // pixel level: draw_partial_tile on the 4x4 block at (Bx, By)
              float evp[3];
              const float bl = static_cast<float>(Bx), bt = static_cast<float>(By);
#pragma unroll
              for (int k = 0; k < 3; ++k) evp[k] = L.e[k * 3 + 2] - ((bl * L.e[k * 3]) + (bt * L.e[k * 3 + 1]));
#pragma unroll
              for (int i = 0; i < 4; ++i) {
                uint32_t m = 0;
                const float fx = static_cast<float>(px0 + (i & 1) - Bx), fy = static_cast<float>(py0 + (i >> 1) - By);
#pragma unroll
                for (int s = 0; s < S; ++s) {
                  float mx = P.sp[s][0] + fx, my = P.sp[s][1] + fy;
                  bool r = false;
#pragma unroll
                  for (int k = 0; k < 3; ++k) r |= ((mx * L.e[k * 3]) + (my * L.e[k * 3 + 1])) < evp[k];
                  if (!r) m |= 1u << s;
                }
                cov[i] = m;
              }
            }
          }
        }
        // pixels outside the render target are never drawn
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (!pin[i]) cov[i] = 0;
        if ((cov[0] | cov[1] | cov[2] | cov[3]) == 0) continue;
        // depth of the quad's pixels (step_2d_unproj_pos_quad, z lane)
        const float dx = (0.5f + static_cast<float>(px0)) - L.zx;
        const float dy = (0.5f + static_cast<float>(py0)) - L.zy;
        float z[4];
        z[0] = L.z0 + ((L.dzx * dx) + (L.dzy * dy));
        z[1] = z[0] + L.dzx;
        z[2] = z[0] + L.dzy;
        z[3] = z[1] + L.dzy;
        const uint32_t id = L.idf >> 1;
        if (P.early_z) {
          // early_z_test (framebuffer.cpp:522-614): depth written at test time
          uint32_t tested[4] = {0, 0, 0, 0};
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int s = 0; s < S; ++s) {
              if (!(cov[i] & (1u << s))) continue;
              float nd = (S > 1) ? (L.aa[s] + z[i]) : z[i];
              float od = P.read_depth ? depth[i][s] : 0.0f;
              bool pass = P.depth_enable ? cmpf(P.depth_func, nd, od) : true;
              if (pass) {
                if (P.write_depth) depth[i][s] = nd;
                tested[i] |= 1u << s;
              }
            }
          if ((tested[0] | tested[1] | tested[2] | tested[3]) == 0) continue;
          ps_inv += 4;
          be_pix += 4;
          if constexpr (VIS) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
              for (int s = 0; s < S; ++s)
                if (tested[i] & (1u << s)) winner[i][s] = id;
          } else {
            shade_store<NA, S, SASL>(P, SM, tri_attr, id, px0, py0,
                                     tested[0] | (tested[1] << 4) | (tested[2] << 8) | (tested[3] << 12));
          }
        } else if constexpr (!VIS) {
          // late depth/stencil (render_sample, framebuffer.cpp:445-479); writes only when both pass.
          // The tests of a sample depend only on that sample's depth/stencil, not on the colour, so
          // they run first and the pixel shader (side-effect free) runs only when some sample passes.
          ps_inv += 4;
          be_pix += 4;
          const int face = L.front ? 0 : 1;
          uint32_t passed[4] = {0, 0, 0, 0};
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int s = 0; s < S; ++s) {
              if (!(cov[i] & (1u << s))) continue;
              float nd = (S > 1) ? (z[i] + L.aa[s]) : z[i];
              float od = P.read_depth ? depth[i][s] : 0.0f;
              uint32_t os = P.stencil_enable ? (stencil[i][s] & P.read_mask) : 0u;
              bool dp = has_ds ? (P.depth_enable ? cmpf(P.depth_func, nd, od) : true) : true;
              bool sp = P.stencil_enable ? cmpu(P.st_func[face], P.stencil_ref, os) : true;
              if (dp && sp) passed[i] |= 1u << s;
            }
          if ((passed[0] | passed[1] | passed[2] | passed[3]) == 0) continue;
          shade_store<NA, S, SASL>(P, SM, tri_attr, id, px0, py0,
                                   passed[0] | (passed[1] << 4) | (passed[2] << 8) | (passed[3] << 12));
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int s = 0; s < S; ++s) {
              if (!(passed[i] & (1u << s))) continue;
              float nd = (S > 1) ? (z[i] + L.aa[s]) : z[i];
              uint32_t os = P.stencil_enable ? (stencil[i][s] & P.read_mask) : 0u;
              uint32_t ns = P.stencil_enable ? stencil_op(P.st_pass_op[face], P.stencil_ref, os) : 0u;
              if (P.write_depth) depth[i][s] = nd;
              if (P.stencil_enable) stencil[i][s] = ns & P.write_mask;
            }
        }
      }
    }
    __syncthreads();
  }

  if constexpr (VIS) {
    // winners for the shading pass, laid out like the colour target's samples
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (pin[i]) {
        size_t base = (static_cast<size_t>(py0 + (i >> 1)) * tw + (px0 + (i & 1))) * S;
#pragma unroll
        for (int s = 0; s < S; ++s) vis[base + s] = winner[i][s];
      }
    if (has_ds && P.write_depth) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (pin[i])
#pragma unroll
          for (int s = 0; s < S; ++s)
            *reinterpret_cast<float*>(sample_addr(P.ds, px0 + (i & 1), py0 + (i >> 1), s)) = depth[i][s];
    }
  } else if (has_ds && (P.write_depth || P.stencil_enable)) {
    // write back depth/stencil
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (pin[i])
#pragma unroll
        for (int s = 0; s < S; ++s)
          *reinterpret_cast<float2*>(sample_addr(P.ds, px0 + (i & 1), py0 + (i >> 1), s)) =
              make_float2(depth[i][s], __uint_as_float(stencil[i][s]));
  }
  // ps_invocations and backend_input_pixels advance together (+4 per shaded quad: the sample shaders
  // never discard), so one partial carries both
  (void)be_pix;
  if (P.count_stats) block_partial(partial, ps_inv);  // uniform over the launch
}

// ---------------------------------------------------------------- order-independent visibility
// For an opaque early-Z draw with blending off, the reference's sequential early-Z test (depth written
// at test time, framebuffer.cpp:522-614) leaves in every sample the result of a running reduction:
//   LESS/GREATER (+write): the first triangle reaching the extreme depth wins;
//   LESS_EQUAL/GREATER_EQUAL (+write): the last triangle reaching the extreme depth wins;
//   ALWAYS, depth off, EQUAL, or any function without depth write: among the triangles that pass
//   against the initial depth, the last one wins.
// Each of these is a minimum over a 64-bit key (ordered depth, primitive order), so lanes can
// rasterise different triangles of a tile at once and merge with LDS atomic-min.
enum { VIS_LESS = 0, VIS_LEQUAL = 1, VIS_GREATER = 2, VIS_GEQUAL = 3, VIS_LATEST = 4, VIS_NONE = 5 };

__device__ __forceinline__ uint32_t ord_depth(float z) {
  uint32_t u = __float_as_uint(z + 0.0f);  // -0 compares equal to +0
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ unsigned long long initial_key(int mode, float d0) {
  if (mode == VIS_LATEST || mode == VIS_NONE) return 0xFFFFFFFFull;
  if (d0 != d0) return 0ull;  // a NaN depth lets nothing pass
  uint32_t zk = ord_depth(d0);
  if (mode == VIS_GREATER || mode == VIS_GEQUAL) zk = ~zk;
  uint32_t lo = (mode == VIS_LESS || mode == VIS_GREATER) ? 0u : 0xFFFFFFFFu;
  return (static_cast<unsigned long long>(zk) << 32) | lo;
}

__device__ __forceinline__ uint32_t key_winner(int mode, unsigned long long k) {
  uint32_t lo = static_cast<uint32_t>(k);
  if (mode == VIS_LESS || mode == VIS_GREATER) return lo == 0u ? NO_WINNER : lo - 1u;
  return lo == 0xFFFFFFFFu ? NO_WINNER : 0xFFFFFFFEu - lo;
}

// Depth the reference's early-Z test wrote for winner w of sample smp of pixel (px, py).
template <int S>
__device__ __forceinline__ float winner_depth(const draw_params& P, int mode, unsigned long long key, uint32_t w,
                                              const float4* __restrict__ tri_rec, int px, int py, int smp) {
  if (mode != VIS_LATEST) {
    // depth modes carry the winning sample depth in the key; only +-0 is ambiguous there
    uint32_t zk = static_cast<uint32_t>(key >> 32);
    if (mode == VIS_GREATER || mode == VIS_GEQUAL) zk = ~zk;
    const uint32_t u = (zk & 0x80000000u) ? (zk & 0x7FFFFFFFu) : ~zk;
    if (u != 0u) return __uint_as_float(u);
  }
  // the winner's depth at this sample, as its quad computed it
  const float4* tr = tri_rec + tri_slot(P, w) * 5;
  const float4 d3 = tr[3], d4 = tr[4];
  const int qx0 = px & ~1, qy0 = py & ~1;
  const float dx = (0.5f + static_cast<float>(qx0)) - d3.y;
  const float dy = (0.5f + static_cast<float>(qy0)) - d3.z;
  float z00 = d3.w + ((d4.x * dx) + (d4.y * dy));
  float zp;
  if ((px & 1) == 0 && (py & 1) == 0) zp = z00;
  else if ((py & 1) == 0) zp = z00 + d4.x;
  else if ((px & 1) == 0) zp = z00 + d4.y;
  else zp = (z00 + d4.x) + d4.y;
  float nd = zp;
  if (S > 1) nd = ((P.sp[smp][0] - 0.5f) * d4.x + (P.sp[smp][1] - 0.5f) * d4.y) + zp;
  return nd;
}

// cmpf as a truth table over the comparison's outcome: bit 0 less, 1 equal, 2 greater, 3 unordered.
__device__ __forceinline__ uint32_t cmpf_table(int f) {
  // nibble f: never, less, equal, less_equal, greater, not_equal, greater_equal, always
  return (f >= 0 && f < 8) ? (0xF6D43210u >> (4 * f)) & 0xFu : 0xFu;
}
__device__ __forceinline__ bool cmpf_by_table(uint32_t table, float a, float b) {
  const uint32_t o = (a < b) ? 0u : (a == b) ? 1u : (a > b) ? 2u : 3u;
  return ((table >> o) & 1u) != 0u;
}

// Phase 1c of vis_body over the chunk's work list: items [0, np) are (record, 4x4 block) pairs the
// rasteriser tests sample by sample (draw_partial_tile, rasterizer.cpp:298-439), items [np, nt)
// blocks it accepts whole (draw_full_tile, :285-296).  A partial item tests, pixel by pixel, only
// the pixels a conservative bound leaves (see below), each sample with the reference's expression
// ((sx + fx) * ex) + ((sy + fy) * ey) < e_value, and every covered sample's key goes to the LDS
// atomic-min.  Depth per sample is the quad chain of
// step_2d_unproj_pos_quad (p00 = v0 + (ddx*dx + ddy*dy), +ddx to the right, +ddy below) plus the
// sample offset term of rasterize_triangle (rasterizer.cpp:678-687).
// Keys: DEPTH (the LESS, LEQUAL, GREATER, GEQUAL modes with depth writes) = the sample depth's
// order-preserving bits (complemented when the greater depth wins) over the record's order term,
// and a NaN depth passes nothing; otherwise (VIS_LATEST) the sample passes against the target's
// initial depth (always without a depth test) and the key is the order term alone.  The per-sample
// predicates are combined with bitwise operators so that each sample costs one masked atomic and
// no branch.
template <int S, int GT, bool DEPTH>
__device__ __forceinline__ void vis_items(const draw_params& P, int mode, const float4* __restrict__ recs,
                                          const uint16_t* __restrict__ work, uint32_t np, uint32_t nt,
                                          unsigned long long* __restrict__ keys, const float* __restrict__ d0s,
                                          int ox, int oy, int tw, int th, int tid, int NT) {
  constexpr int NBS = GT / 4;
  const uint32_t flip = (mode == VIS_GREATER || mode == VIS_GEQUAL) ? 0xFFFFFFFFu : 0u;
  const bool later_wins = mode == VIS_LEQUAL || mode == VIS_GEQUAL || mode == VIS_LATEST;
  const uint32_t table = P.depth_enable ? cmpf_table(P.depth_func) : 0xFu;
  const bool read_depth = P.read_depth != 0;
  // extreme sample positions inside a pixel
  float spx_lo = P.sp[0][0], spx_hi = P.sp[0][0], spy_lo = P.sp[0][1], spy_hi = P.sp[0][1];
#pragma unroll
  for (int s = 1; s < S; ++s) {
    spx_lo = fminf(spx_lo, P.sp[s][0]);
    spx_hi = fmaxf(spx_hi, P.sp[s][0]);
    spy_lo = fminf(spy_lo, P.sp[s][1]);
    spy_hi = fmaxf(spy_hi, P.sp[s][1]);
  }
  for (uint32_t wi = tid; wi < nt; wi += NT) {
    const uint32_t wk = work[wi];
    const int j = static_cast<int>(wk >> 7), b = static_cast<int>((wk >> 1) & 63u);
    const float4 r0 = recs[j * 4 + 0], r1 = recs[j * 4 + 1], r2 = recs[j * 4 + 2], r3 = recs[j * 4 + 3];
    const float e[9] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, r2.x};
    const float z0 = r2.y, zx = r2.z, zy = r2.w, dzx = r3.x, dzy = r3.y;
    const uint32_t id = __float_as_uint(r3.z);
    const uint32_t lo = later_wins ? 0xFFFFFFFEu - id : id + 1u;
    const bool full = wi >= np;
    float aa[S];
#pragma unroll
    for (int s = 0; s < S; ++s) aa[s] = (S > 1) ? (P.sp[s][0] - 0.5f) * dzx + (P.sp[s][1] - 0.5f) * dzy : 0.0f;
    const int bc = b % NBS, br = b / NBS;
    const int Bx = ox + bc * 4, By = oy + br * 4;
    const int lbase = (br * 4 * GT + bc * 4) * S;  // the block's first sample in `keys`
    const int bw = min(4, tw - Bx), bh = min(4, th - By);  // columns and rows inside the target
    // depth at the top-left pixel of each 2x2 quad of the block
    float zq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float dx = (0.5f + static_cast<float>(Bx + (q & 1) * 2)) - zx;
      const float dy = (0.5f + static_cast<float>(By + (q >> 1) * 2)) - zy;
      zq[q] = z0 + ((dzx * dx) + (dzy * dy));
    }
    float evp[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
      evp[k] = e[k * 3 + 2] - ((static_cast<float>(Bx) * e[k * 3]) + (static_cast<float>(By) * e[k * 3 + 1]));
    if (!full) {
      // Pixels that may hold a covered sample.  Float rounding is monotonic, so an edge's value at
      // any sample of a pixel is at most its value at the pixel's extreme sample column and row
      // for that edge's signs, computed with the same operations; a pixel whose extreme value
      // fails an edge has no sample that passes it.  Only the remaining pixels are tested sample
      // by sample.
      uint32_t M = 0;
      {
        float cx[3][4], cy[3][4];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const float sx = e[k * 3] >= 0.0f ? spx_hi : spx_lo, sy = e[k * 3 + 1] >= 0.0f ? spy_hi : spy_lo;
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            cx[k][c] = (sx + static_cast<float>(c)) * e[k * 3];
            cy[k][c] = (sy + static_cast<float>(c)) * e[k * 3 + 1];
          }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const bool maybe = !((cx[0][c] + cy[0][r]) < evp[0]) & !((cx[1][c] + cy[1][r]) < evp[1]) &
                               !((cx[2][c] + cy[2][r]) < evp[2]) & (c < bw) & (r < bh);
            if (maybe) M |= 1u << (r * 4 + c);
          }
      }
      while (M) {
        const int px = __builtin_ctz(M);
        M &= M - 1u;
        const int r = px >> 2, c = px & 3;
        const int q = ((r >> 1) << 1) | (c >> 1);
        float z = (q & 2) ? ((q & 1) ? zq[3] : zq[2]) : ((q & 1) ? zq[1] : zq[0]);
        z = (c & 1) ? z + dzx : z;
        z = (r & 1) ? z + dzy : z;
        const float fx = static_cast<float>(c), fy = static_cast<float>(r);
        const int lp = lbase + (r * GT + c) * S;
#pragma unroll
        for (int s = 0; s < S; ++s) {
          const float mx = P.sp[s][0] + fx, my = P.sp[s][1] + fy;
          const bool covered = !(((mx * e[0]) + (my * e[1])) < evp[0]) & !(((mx * e[3]) + (my * e[4])) < evp[1]) &
                               !(((mx * e[6]) + (my * e[7])) < evp[2]);
          const float nd = (S > 1) ? (aa[s] + z) : z;
          unsigned long long k;
          bool pass;
          if (DEPTH) {
            const uint32_t u = __float_as_uint(nd + 0.0f);  // -0 orders as +0
            const uint32_t ord = u ^ (static_cast<uint32_t>(static_cast<int32_t>(u) >> 31) | 0x80000000u);
            k = (static_cast<unsigned long long>(ord ^ flip) << 32) | lo;
            pass = nd == nd;
          } else {
            k = lo;
            pass = cmpf_by_table(table, nd, read_depth ? d0s[lp + s] : 0.0f);
          }
          if (covered & pass) atomicMin(&keys[lp + s], k);
        }
      }
      continue;
    }
    // a block the rasteriser accepts whole: every sample inside the target
    float zp[16];  // pixel depths (row r, column c at [r * 4 + c]) by the quad chain
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int p0 = (q >> 1) * 8 + (q & 1) * 2;
      zp[p0] = zq[q];
      zp[p0 + 1] = zq[q] + dzx;
      zp[p0 + 4] = zq[q] + dzy;
      zp[p0 + 5] = zp[p0 + 1] + dzy;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int s = 0; s < S; ++s) {
          const float nd = (S > 1) ? (aa[s] + zp[r * 4 + c]) : zp[r * 4 + c];
          const int li = lbase + (r * GT + c) * S + s;
          unsigned long long k;
          bool pass;
          if (DEPTH) {
            const uint32_t u = __float_as_uint(nd + 0.0f);
            const uint32_t ord = u ^ (static_cast<uint32_t>(static_cast<int32_t>(u) >> 31) | 0x80000000u);
            k = (static_cast<unsigned long long>(ord ^ flip) << 32) | lo;
            pass = nd == nd;
          } else {
            k = lo;
            pass = cmpf_by_table(table, nd, read_depth ? d0s[li] : 0.0f);
          }
          if (pass & (c < bw) & (r < bh)) atomicMin(&keys[li], k);
        }
  }
}

// One 16x16 GPU tile per workgroup, in three phases:
//  1. one triangle per lane: coverage with the reference's decision chain, LDS atomic-min of the
//     (ordered depth, primitive order) key of every covered sample;
//  2. one pixel per lane: winners -> depth/stencil target, and one shading item per distinct winner
//     (pixel, winner, sample mask) appended to an LDS list;
//  3. (SHADE) one item per lane: the winner's quad interpolants at the pixel's quad (the values the
//     reference's last passing draw_quad/draw_full_quad computed), the pixel shader for that pixel,
//     and the colour stored to the item's samples.  Items are spread evenly over the lanes however
//     many triangles share a pixel.
#ifndef SRX_VIS_THREADS
#define SRX_VIS_THREADS 256
#endif
// LDS shared by phase 1 (a chunk of NT records + their (record, block) work list) and phase 3
// (shading items), for GT x GT GPU tiles
__host__ __device__ constexpr int vis_union_bytes(int nt, int s, bool shade, int gt = 16) {
  return (nt * 64 + nt * (gt / 4) * (gt / 4) * 2) > (shade ? gt * gt * s * 8 : 0) ? (nt * 64 + nt * (gt / 4) * (gt / 4) * 2)
                                                                                   : gt * gt * s * 8;
}
// Static LDS of vis_body besides the union (keys, initial depths).
__host__ __device__ constexpr int vis_static_bytes(int s, int gt) { return gt * gt * s * 12 + 256; }
// NTC > 0: NTC lanes per workgroup, static LDS.  NTC == 0: blockDim.x lanes (64, 128 or 256) with
// the phase-1/3 union in dynamic LDS of vis_union_bytes(blockDim.x, S, SHADE, GT) bytes.
// GT: GPU tile side, 8, 16 or 32 pixels: a quarter of a 16x16 stage-16 region, one region, or 2x2
// regions of the reference's 64x64 tile; the 4x4 blocks are the reference's stage-4 blocks either way.
template <int NA, int S, bool SHADE, int NTC, int GT>
__device__ __forceinline__ void vis_body(const draw_params& P, int mode, int pass, const uint32_t* __restrict__ order,
                                         const uint4* __restrict__ entries, const uint4* __restrict__ merges,
                                         const uint32_t* __restrict__ counts, unsigned long long* __restrict__ partial,
                                         const uint32_t* __restrict__ vals, const uint32_t* __restrict__ tstart,
                                         const uint32_t* __restrict__ tend, const float4* __restrict__ tri_rec,
                                         const float4* __restrict__ tri_attr, const ps_samplers* __restrict__ SM,
                                         uint32_t bx, uint32_t* __restrict__ wout = nullptr) {
  static_assert(GT == 8 || GT == 16 || GT == 32, "GPU tile side");
  constexpr int PX = GT * GT;       // pixels per GPU tile
  constexpr int NBS = GT / 4;       // 4x4 blocks per side
  constexpr int RPS = GT > 16 ? GT / 16 : 1;  // stage-16 regions per side
  using bmask_t = typename std::conditional<(GT > 16), unsigned long long, uint32_t>::type;
  const int NT = NTC > 0 ? NTC : static_cast<int>(blockDim.x);
  constexpr int UNION_BYTES = NTC > 0 ? vis_union_bytes(NTC, S, SHADE, GT) : 16;
  __shared__ unsigned long long keys[PX * S];
  __shared__ float d0s[PX * S];
  __shared__ __align__(16) unsigned char lds_static[UNION_BYTES];
  extern __shared__ __align__(16) unsigned char lds_dynamic[];
  unsigned char* const lds_union = NTC > 0 ? lds_static : lds_dynamic;
  __shared__ uint32_t wsum[4];
  __shared__ uint32_t hiz[NBS * NBS];
  __shared__ uint32_t nitems;
  float4* recs = reinterpret_cast<float4*>(lds_union);                           // [NT][4]
  uint16_t* work = reinterpret_cast<uint16_t*>(lds_union + NT * 64);             // [NT * NBS^2]
  uint2* items = reinterpret_cast<uint2*>(lds_union);                            // [PX * S]
  // pass 0: one tile, or one part of a split tile; pass 1: merge and finish a split tile
  int gtile;
  uint32_t part = 0, parts = 1, slot = 0;
  if (entries == nullptr) {  // whole tiles in `order`
    gtile = static_cast<int>(order[bx]);
  } else if (pass == 0) {
    if (bx >= counts[0]) return;
    const uint4 en = entries[bx];
    gtile = static_cast<int>(en.x);
    part = en.y;
    parts = en.z;
    slot = en.w + en.y;
  } else {
    if (bx >= counts[1]) return;
    const uint4 mg = merges[bx];
    gtile = static_cast<int>(mg.x);
    parts = mg.y;
    slot = mg.z;
  }
  const int gx = gtile % P.gtiles_x, gy = gtile / P.gtiles_x;
  const int ox = gx * GT, oy = gy * GT;
  const int t64x = ox & ~(TILE - 1), t64y = oy & ~(TILE - 1);
  const int tw = static_cast<int>(P.target_w), th = static_cast<int>(P.target_h);
  const int tid = threadIdx.x;
  const bool has_ds = P.ds.data != nullptr;
  if (tid == 0) nitems = 0;
  if (pass == 1) {
    // the parts' keys, each started from the target's depth: their minimum is the tile's result
    const unsigned long long* src = partial + static_cast<size_t>(slot) * (PX * S);
    for (int i = tid; i < PX * S; i += NT) {
      unsigned long long k = src[i];
      for (uint32_t q = 1; q < parts; ++q) {
        const unsigned long long kq = src[static_cast<size_t>(q) * (PX * S) + i];
        k = kq < k ? kq : k;
      }
      keys[i] = k;
    }
  } else {
    for (int i = tid; i < PX * S; i += NT) {
      int px = ox + ((i / S) % GT), py = oy + ((i / S) / GT), smp = i % S;
      float d0 = 0.0f;
      bool blocked = false;
      if (has_ds && px < tw && py < th) {
        const float2 ds = P.clear_ds ? make_float2(P.clear_depth, __uint_as_float(P.clear_stencil))
                                     : *reinterpret_cast<const float2*>(sample_addr(P.ds, px, py, smp));
        d0 = ds.x;
        if (P.vis_stencil) {
          // framebuffer.cpp:445-479 (both faces share the function here); a failing sample takes
          // no colour and no depth
          const uint32_t os = __float_as_uint(ds.y) & P.read_mask;
          blocked = !cmpu(P.st_func[1], P.stencil_ref, os);
        }
      }
      d0s[i] = d0;
      // a blocked sample's key is below every candidate's and decodes as no winner
      keys[i] = blocked ? ((mode == VIS_LEQUAL || mode == VIS_GEQUAL) ? 0xFFFFFFFFull : 0ull) : initial_key(mode, d0);
    }
  }
  // this workgroup's records: the whole tile, or one part of it (none when merging)
  const uint32_t tr0 = tstart[gtile], tr1 = tend[gtile];
  const uint32_t t0 = pass == 1 ? tr1 : min(tr1, tr0 + part * VIS_PART);
  const uint32_t t1 = (pass == 1 || entries == nullptr) ? tr1 : min(tr1, t0 + VIS_PART);
  const int jx0 = (ox - t64x) >> 4, jy0 = (oy - t64y) >> 4;  // first stage-16 region of the tile
  // an 8x8 tile covers blocks [qx, qx+2) x [qy, qy+2) of its region; larger tiles whole regions
  const int qx = GT < 16 ? ((ox - t64x) & 15) >> 2 : 0, qy = GT < 16 ? ((oy - t64y) & 15) >> 2 : 0;
  constexpr int CW = GT < 16 ? GT / 4 : 4;  // block columns (rows) of a region inside the tile
  const float tvx = static_cast<float>(t64x), tvy = static_cast<float>(t64y);
  // Depth culling from the second chunk on: hiz[b] is the largest key depth word of 4x4 block b's
  // samples after the previous chunks.  Keys only decrease, so a record whose smallest possible
  // depth word over the block exceeds it cannot win any sample there and skips the block.
  const bool depth_keys = mode == VIS_LESS || mode == VIS_LEQUAL || mode == VIS_GREATER || mode == VIS_GEQUAL;
  const bool far_is_min = mode == VIS_GREATER || mode == VIS_GEQUAL;
  for (uint32_t c0 = t0; c0 < t1; c0 += NT) {
    __syncthreads();  // the previous chunk's work list is consumed (and the keys are initialised)
    const bool hiz_on = depth_keys && c0 > t0;
    if (hiz_on) {
      // 16 consecutive lanes take one block's 16 pixels
      for (int u = tid; u < PX; u += NT) {
        const int b = u >> 4, q = u & 15;
        const int pl = ((b / NBS) * 4 + (q >> 2)) * GT + (b % NBS) * 4 + (q & 3);
        uint32_t m = 0;
#pragma unroll
        for (int smp = 0; smp < S; ++smp) m = max(m, static_cast<uint32_t>(keys[pl * S + smp] >> 32));
        m = max(m, static_cast<uint32_t>(__shfl_xor(static_cast<int>(m), 1, 64)));
        m = max(m, static_cast<uint32_t>(__shfl_xor(static_cast<int>(m), 2, 64)));
        m = max(m, static_cast<uint32_t>(__shfl_xor(static_cast<int>(m), 4, 64)));
        m = max(m, static_cast<uint32_t>(__shfl_xor(static_cast<int>(m), 8, 64)));
        if (q == 0) hiz[b] = m;
      }
      __syncthreads();
    }
    // phase 1a: one record per lane -> the 4x4 blocks of this GPU tile it reaches (full or partial)
    bmask_t bm = 0, fm = 0;
    const uint32_t ri = c0 + tid;
    if (ri < t1) {
      const uint32_t rec = vals[ri];
      const uint32_t id = rec >> 1;
      const float4* tr = tri_rec + tri_slot(P, id) * 5;
      float e[9], bb[4];
      read_tri_rec(tr, e, bb);
      const float4 d3 = tr[3], d4 = tr[4];
      const float bx0 = bb[0] - tvx, bx1 = bb[1] - tvx, by0 = bb[2] - tvy, by1 = bb[3] - tvy;
      // per-stage constants (rasterize_triangle, rasterizer.cpp:662-675, 704-710)
      float s1x[3], s1y[3], ra1[3], ev1[3], s2x[3], s2y[3], ra2[3], ev2[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        float ex = e[k * 3], ey = e[k * 3 + 1], ez = e[k * 3 + 2];
        bool mx = ex > 0, my = ey > 0;
        float sx = TILE * ex, sy = TILE * ey;
        float ra = -fabsf(sx) - fabsf(sy);
        float part = static_cast<float>(mx * TILE) * ex + static_cast<float>(my * TILE) * ey;
        sx *= 0.25f; sy *= 0.25f; ra *= 0.25f; part *= 0.25f;
        s1x[k] = sx; s1y[k] = sy; ra1[k] = ra;
        ev1[k] = (ez - part) - ((tvx * ex) + (tvy * ey));
        sx *= 0.25f; sy *= 0.25f; ra *= 0.25f; part *= 0.25f;
        s2x[k] = sx; s2y[k] = sy; ra2[k] = ra;
        ev2[k] = ez - part;
      }
      const bool rec_full = (rec & 1) != 0;  // the record covers its whole 64x64 tile
#pragma unroll
      for (int rr = 0; rr < RPS * RPS; ++rr) {
        const int jx = jx0 + rr % RPS, jy = jy0 + rr / RPS;
        bool region_full = rec_full;
        bool keep = true;
        if (!region_full) {
          // stage 16 (subdivide_tile on the 64x64 tile)
          bool rej = false, nacc = false;
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            float r = (s1x[k] * static_cast<float>(jx)) + (s1y[k] * static_cast<float>(jy));
            float a = r + ra1[k];
            rej |= r < ev1[k];
            nacc |= a < ev1[k];
          }
          rej |= bx0 >= static_cast<float>(jx * 16 + 16);
          rej |= bx1 < static_cast<float>(jx * 16);
          rej |= by0 >= static_cast<float>(jy * 16 + 16);
          rej |= by1 < static_cast<float>(jy * 16);
          keep = !rej;
          region_full = !nacc;
        }
        if (!keep) continue;
        const int rox = t64x + jx * 16, roy = t64y + jy * 16;  // the region's origin
        float ev2q[3];
#pragma unroll
        for (int k = 0; k < 3; ++k)
          ev2q[k] = ev2[k] - ((static_cast<float>(rox) * e[k * 3]) + (static_cast<float>(roy) * e[k * 3 + 1]));
        // candidate 4x4 blocks of the region inside this tile: all of them for an accepted region,
        // else those the stage-4 bounding-box test keeps (bx0 < mxr + 4 and bx1 >= mxr, the same
        // comparisons on the same floats)
        uint32_t cols = 0, rows = 0;
#pragma unroll
        for (int cc = 0; cc < CW; ++cc) {
          const int cx = qx + cc, cy = qy + cc;
          const int mxr = jx * 16 + cx * 4, myr = jy * 16 + cy * 4;
          if (region_full || !(bx0 >= static_cast<float>(mxr + 4) || bx1 < static_cast<float>(mxr))) cols |= 0x1111u << cx;
          if (region_full || !(by0 >= static_cast<float>(myr + 4) || by1 < static_cast<float>(myr))) rows |= 0xFu << (cy * 4);
          if (rox + cx * 4 >= tw) cols &= ~(0x1111u << cx);  // blocks outside the target
          if (roy + cy * 4 >= th) rows &= ~(0xFu << (cy * 4));
        }
        uint32_t blocks = cols & rows;
        while (blocks) {
          const int b16 = __builtin_ctz(blocks);
          blocks &= blocks - 1;
          const int ix4 = b16 & 3, iy4 = b16 >> 2;
          bool block_full = region_full;
          if (!block_full) {
            // stage 4 (subdivide_tile on the 16x16 region); the bounding-box half was applied above
            bool rej = false, nacc = false;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              float r = (s2x[k] * static_cast<float>(ix4)) + (s2y[k] * static_cast<float>(iy4));
              float a = r + ra2[k];
              rej |= r < ev2q[k];
              nacc |= a < ev2q[k];
            }
            if (rej) continue;
            block_full = !nacc;
          }
          // the block's index in this tile
          const int b = ((rr / RPS) * 4 + iy4 - qy) * NBS + (rr % RPS) * 4 + ix4 - qx;
          if (hiz_on) {
            // the depth plane's extreme over the block's sample positions [Bx, Bx+4) x [By, By+4),
            // widened by far more than the rounding of the per-sample evaluation
            const float z0 = d3.w, zx = d3.y, zy = d3.z, dzx = d4.x, dzy = d4.y;
            const float ax = static_cast<float>(rox + ix4 * 4) - zx, ay = static_cast<float>(roy + iy4 * 4) - zy;
            const float tx0 = dzx * ax, tx1 = dzx * (ax + 4.0f), ty0 = dzy * ay, ty1 = dzy * (ay + 4.0f);
            const float margin = (fabsf(z0) + fmaxf(fabsf(tx0), fabsf(tx1)) + fmaxf(fabsf(ty0), fabsf(ty1))) * 1.52587890625e-05f;
            const float zext = far_is_min ? (z0 + fmaxf(tx0, tx1) + fmaxf(ty0, ty1)) + margin
                                          : (z0 + fminf(tx0, tx1) + fminf(ty0, ty1)) - margin;
            if (fabsf(zext) < 3.0e38f) {  // finite: planes with inf/nan terms are never culled
              const uint32_t zk = far_is_min ? ~ord_depth(zext) : ord_depth(zext);
              if (zk > hiz[b]) continue;
            }
          }
          bm |= static_cast<bmask_t>(1) << b;
          if (block_full) fm |= static_cast<bmask_t>(1) << b;
        }
      }
      float4* rs = recs + tid * 4;
      rs[0] = make_float4(e[0], e[1], e[2], e[3]);
      rs[1] = make_float4(e[4], e[5], e[6], e[7]);
      rs[2] = make_float4(e[8], d3.w, d3.y, d3.z);          // z0, zx, zy
      rs[3] = make_float4(d4.x, d4.y, __uint_as_float(id), 0.0f);  // dzx, dzy, id
    }
    // phase 1b: work list of (record, block) pairs spread evenly over the lanes: blocks tested
    // sample by sample first, then blocks accepted whole, so each wave runs one kind
    const bmask_t pm = bm & ~fm;
    const uint32_t np = GT > 16 ? static_cast<uint32_t>(__popcll(static_cast<unsigned long long>(pm))) : __popc(static_cast<uint32_t>(pm));
    const uint32_t nf = GT > 16 ? static_cast<uint32_t>(__popcll(static_cast<unsigned long long>(fm))) : __popc(static_cast<uint32_t>(fm));
    const uint32_t cnt = np | (nf << 16);  // both sums stay below 2^16 (NT * NBS^2 <= 16384)
    const uint32_t incl = wave_incl_scan(cnt);
    if ((tid & 63) == 63) wsum[tid >> 6] = incl;
    __syncthreads();
    uint32_t off = incl - cnt, total = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      if (w >= NT / 64) break;
      if (w < (tid >> 6)) off += wsum[w];
      total += wsum[w];
    }
    const uint32_t total_p = total & 0xFFFFu, total_all = total_p + (total >> 16);
    uint32_t off_p = off & 0xFFFFu, off_f = total_p + (off >> 16);
    bmask_t pr = pm, fr = fm;
    while (pr) {
      const uint32_t b = static_cast<uint32_t>(__builtin_ctzll(static_cast<unsigned long long>(pr)));
      pr &= pr - 1;
      work[off_p++] = static_cast<uint16_t>((tid << 7) | (b << 1));
    }
    while (fr) {
      const uint32_t b = static_cast<uint32_t>(__builtin_ctzll(static_cast<unsigned long long>(fr)));
      fr &= fr - 1;
      work[off_f++] = static_cast<uint16_t>((tid << 7) | (b << 1) | 1u);
    }
    __syncthreads();
    // phase 1c: one (record, block) item per lane: coverage, depth and the key atomic-min
    if (depth_keys)
      vis_items<S, GT, true>(P, mode, recs, work, total_p, total_all, keys, d0s, ox, oy, tw, th, tid, NT);
    else if (mode == VIS_LATEST)
      vis_items<S, GT, false>(P, mode, recs, work, total_p, total_all, keys, d0s, ox, oy, tw, th, tid, NT);
  }
  __syncthreads();
  if (pass == 0 && parts > 1) {  // one part of a split tile: leave the partial keys for the merge
    unsigned long long* dst = partial + static_cast<size_t>(slot) * (PX * S);
    for (int i = tid; i < PX * S; i += NT) dst[i] = keys[i];
    return;
  }
  // phase 2: winners -> depth, shading items
  for (int p = tid; p < PX; p += NT) {
    const int px = ox + (p % GT), py = oy + (p / GT);
    if (px >= tw || py >= th) continue;
    uint32_t w[S];
    uint32_t left = 0;
#pragma unroll
    for (int smp = 0; smp < S; ++smp) {
      const unsigned long long key = keys[p * S + smp];
      w[smp] = key_winner(mode, key);
      const bool won = w[smp] != NO_WINNER;
      if (won) left |= 1u << smp;
      if (has_ds && P.clear_ds) {
        // a cleared target: every sample is written, the stencil keeps its clear value
        const float d = (won && P.write_depth) ? winner_depth<S>(P, mode, key, w[smp], tri_rec, px, py, smp) : P.clear_depth;
        *reinterpret_cast<float2*>(sample_addr(P.ds, px, py, smp)) = make_float2(d, __uint_as_float(P.clear_stencil));
      } else if (won && has_ds && P.write_depth) {
        *reinterpret_cast<float*>(sample_addr(P.ds, px, py, smp)) =
            winner_depth<S>(P, mode, key, w[smp], tri_rec, px, py, smp);
      }
      if (SHADE && !won && P.clear_color) {
        const float4 cc = make_float4(P.clear_rgba[0], P.clear_rgba[1], P.clear_rgba[2], P.clear_rgba[3]);
        store_texel(P.color.pf, sample_addr(P.color, px, py, smp), cc);
      }
      // winners for a separate shading pass (SASL pixel shaders), laid out like the colour samples
      if (!SHADE && wout) wout[(static_cast<size_t>(py) * tw + px) * S + smp] = w[smp];
    }
    if (SHADE) {
#pragma unroll
      for (int it = 0; it < S; ++it) {
        if (!left) break;
        const int s0 = __builtin_ctz(left);
        uint32_t cur = w[0];
#pragma unroll
        for (int smp = 1; smp < S; ++smp)
          if (smp == s0) cur = w[smp];
        uint32_t m = 0;
#pragma unroll
        for (int smp = 0; smp < S; ++smp)
          if (w[smp] == cur) m |= 1u << smp;
        left &= ~m;
        items[atomicAdd(&nitems, 1u)] = make_uint2(cur, static_cast<uint32_t>(p) | (m << 12));
      }
    }
  }
  if (!SHADE) return;
  __syncthreads();
  // phase 3: shading items
  const uint32_t n = nitems;
  for (uint32_t j = tid; j < n; j += NT) {
    const uint2 item = items[j];
    const uint32_t cur = item.x, p = item.y & 4095u, m = item.y >> 12;
    if (cur >= 3u * P.prim_count) continue;  // every winner id is below 3 * prim_count
    const int px = ox + static_cast<int>(p % GT), py = oy + static_cast<int>(p / GT);
    const int pi = (px & 1) | ((py & 1) << 1);
    attr_setup<NA> at;
    load_attr<NA>(P, tri_attr, cur, at);
    const quad_lazy<NA> q(at, P, px & ~1, py & ~1);
    const float4 c = run_ps_pixel<NA>(P, SM, q, pi);
#pragma unroll
    for (int smp = 0; smp < S; ++smp)
      if (m & (1u << smp)) store_texel(P.color.pf, sample_addr(P.color, px, py, smp), c);
  }
}
template <int NA, int S, bool SHADE>
__global__ void __launch_bounds__(SRX_VIS_THREADS, 4) vis_kernel(draw_params P, int mode, int pass,
                                                   const uint32_t* __restrict__ order,
                                                   const uint4* __restrict__ entries, const uint4* __restrict__ merges,
                                                   const uint32_t* __restrict__ counts,
                                                   unsigned long long* __restrict__ partial,
                                                   const uint32_t* __restrict__ vals,
                                                   const uint32_t* __restrict__ tstart, const uint32_t* __restrict__ tend,
                                                   const float4* __restrict__ tri_rec, const float4* __restrict__ tri_attr,
                                                   const ps_samplers* __restrict__ SM, uint32_t* __restrict__ wout) {
  vis_body<NA, S, SHADE, SRX_VIS_THREADS, 16>(P, mode, pass, order, entries, merges, counts, partial, vals, tstart, tend,
                                              tri_rec, tri_attr, SM, blockIdx.x, wout);
}


// Deferred pixel shading of an opaque early-Z draw: one thread per pixel (16x16 pixels per
// workgroup).  For each distinct winning triangle of the pixel's samples, the thread evaluates that
// triangle's quad interpolants (the values the reference's last passing draw_quad/draw_full_quad
// computed for this quad), runs the pixel shader for its own pixel and writes the winning samples.
template <int NA, int S>
__global__ void __launch_bounds__(256) shade_kernel(draw_params P, const ps_samplers* __restrict__ SM,
                                                     const float4* __restrict__ tri_attr, const uint32_t* __restrict__ vis) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
  const int tw = static_cast<int>(P.target_w), th = static_cast<int>(P.target_h);
  if (x >= tw || y >= th) return;
  const int px0 = x & ~1, py0 = y & ~1;
  const int pi = (x & 1) | ((y & 1) << 1);  // position in the quad
  const size_t base = (static_cast<size_t>(y) * tw + x) * S;
  uint32_t w[S];
#pragma unroll
  for (int s = 0; s < S; ++s) w[s] = vis[base + s];
#pragma unroll
  for (int it = 0; it < S; ++it) {
    uint32_t cur = NO_WINNER;
#pragma unroll
    for (int s = 0; s < S; ++s)
      if (cur == NO_WINNER) cur = w[s];
    if (cur >= 3u * P.prim_count) break;  // NO_WINNER (every id is below 3 * prim_count)
    attr_setup<NA> at;
    load_attr<NA>(P, tri_attr, cur, at);
    const quad_lazy<NA> q(at, P, px0, py0);
    const float4 c = run_ps_pixel<NA>(P, SM, q, pi);
#pragma unroll
    for (int s = 0; s < S; ++s)
      if (w[s] == cur) {
        store_texel(P.color.pf, sample_addr(P.color, x, y, s), c);
        w[s] = NO_WINNER;
      }
  }
}

// Deferred SASL shading of an opaque early-Z draw: one thread per 2x2 quad of a 16x16 GPU tile
// (the tiles the visibility pass wrote).  For each distinct winning triangle among the quad's samples
// the program runs once on the whole quad (pixel_shader_unit::execute, shader_unit.cpp:147-176) and
// the samples that triangle won take their pixel's colour: a SASL pixel's colour depends only on the
// triangle and the quad, so this equals shading every early-Z pass in order.
template <int NA, int S>
__global__ void __launch_bounds__(64) sasl_shade_kernel(draw_params P, const ps_samplers* __restrict__ SM,
                                                        const float4* __restrict__ tri_attr, const uint32_t* __restrict__ vis) {
  const int px0 = blockIdx.x * 16 + (threadIdx.x & 7) * 2, py0 = blockIdx.y * 16 + (threadIdx.x >> 3) * 2;
  const int tw = static_cast<int>(P.target_w), th = static_cast<int>(P.target_h);
  if (px0 >= tw || py0 >= th) return;
  uint32_t w[4][S];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int x = px0 + (i & 1), y = py0 + (i >> 1);
    const bool in = x < tw && y < th;
#pragma unroll
    for (int s = 0; s < S; ++s) w[i][s] = in ? vis[(static_cast<size_t>(y) * tw + x) * S + s] : NO_WINNER;
  }
  const sasl_program* __restrict__ prog = P.sasl;
  for (int it = 0; it < 4 * S; ++it) {
    uint32_t cur = NO_WINNER;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int s = 0; s < S; ++s)
        if (cur == NO_WINNER) cur = w[i][s];
    if (cur >= 3u * P.prim_count) break;  // NO_WINNER (every id is below 3 * prim_count)
    attr_setup<NA> at;
    load_attr<NA>(P, tri_attr, cur, at);
    const quad_lazy<NA> q(at, P, px0, py0);
    sasl_regs R;
    const uint32_t nregs = prog->nregs, nin = prog->ninputs;
    for (uint32_t r = 0; r < nregs; ++r)
      for (int px = 0; px < 4; ++px) R.r[r][px] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int px = 0; px < 4; ++px) {
      const float iw = q.inv_w(px);
      for (uint32_t k = 0; k < nin; ++k) {
        const sasl_input inp = prog->inputs[k];
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int a = 0; a < NA; ++a)
          if (a == inp.attr) v = q.reg(a + 1, px, iw);
        R.r[inp.reg][px] = make_float4(v.x, inp.ncomp > 1 ? v.y : 0.0f, inp.ncomp > 2 ? v.z : 0.0f, inp.ncomp > 3 ? v.w : 0.0f);
      }
    }
    sasl_exec(prog, R, [&](int slot, int, float u, float v, float dxu, float dxv, float dyu, float dyv) -> float4 {
      if (!P.has_sampler[slot]) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      const sampler_desc_dev& sp = SM->s[slot];
      return sp.fast_aniso ? sample_2d_grad_fast(sp, u, v, dxu, dxv, dyu, dyv, 0.0f)
                           : sample_2d_grad(sp, u, v, dxu, dxv, dyu, dyv, 0.0f);
    });
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      bool any = false;
#pragma unroll
      for (int s = 0; s < S; ++s) any |= w[i][s] == cur;
      if (!any) continue;
      const float4 c = sasl_color0(prog, R, i);
      const int x = px0 + (i & 1), y = py0 + (i >> 1);
#pragma unroll
      for (int s = 0; s < S; ++s)
        if (w[i][s] == cur) {
          store_texel(P.color.pf, sample_addr(P.color, x, y, s), c);
          w[i][s] = NO_WINNER;
        }
    }
  }
}

// ============================================================== host launchers
hipError_t upload_ewa_table(const float* table256, hipStream_t s) {
  return hipMemcpyToSymbolAsync(HIP_SYMBOL(c_ewa), table256, 256 * sizeof(float), 0, hipMemcpyHostToDevice, s);
}

void launch_vs(const draw_params& P, float4* vbuf, uint32_t* counters, uint32_t* tiles, size_t ntiles, hipStream_t s) {
  uint32_t n = P.nverts + 1;
  hipLaunchKernelGGL(vs_kernel, dim3((n + 255) / 256), dim3(256), 0, s, P, vbuf, counters, tiles,
                     static_cast<uint32_t>(ntiles));
}

// Kernels are instantiated per attribute count: the vertex and setup buffers are strided by it.
#define SRX_NA_DISPATCH(NAV, CALL) \
  switch (NAV) {                   \
  case 1: CALL(1); break;          \
  case 2: CALL(2); break;          \
  case 3: CALL(3); break;          \
  case 4: CALL(4); break;          \
  default: CALL(5); break;         \
  }
// Batched draws are refused above two attributes when they are recorded (batch_max_attrs).
#define SRX_NA_DISPATCH_BATCH(NAV, CALL) \
  switch (NAV) {                         \
  case 1: CALL(1); break;                \
  default: CALL(2); break;               \
  }

void launch_setup(const draw_params& P, const draw_buffers& B, hipStream_t s) {
  if (P.prim_count == 0) return;
  dim3 grid((P.prim_count + 255) / 256);
#define SRX_SETUP(NA) \
  hipLaunchKernelGGL(setup_kernel<NA>, grid, dim3(256), 0, s, P, B.vbuf, B.tri_rec, B.tri_attr, B.cnt, B.clip_flag, B.counters, \
                     B.wide_list, B.partial, B.bsum, B.pairs)
  SRX_NA_DISPATCH(P.nattr, SRX_SETUP)
#undef SRX_SETUP
  if (P.count_stats)
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(256), 0, s, B.partial, grid.x, B.stats, static_cast<int>(ST_CPRIMITIVES));
}

void launch_clip(const draw_params& P, const draw_buffers& B, hipStream_t s) {
  if (P.prim_count == 0) return;
  dim3 grid(std::min<uint32_t>((P.prim_count + 255) / 256, 256));  // up to 1024 waves in 256-lane workgroups
#define SRX_CLIP(NA)                                                                                                     \
  hipLaunchKernelGGL(clip_kernel<NA>, grid, dim3(256), 0, s, P, B.vbuf, B.tri_rec, B.tri_attr, B.cnt, B.clip_flag, B.clip_slot, \
                     B.clip_verts, B.clip_hdr, B.clip_capacity, B.counters, B.stats, B.wide_list, B.bsum)
  SRX_NA_DISPATCH(P.nattr, SRX_CLIP)
#undef SRX_CLIP
}

void launch_emit(const draw_params& P, const draw_buffers& B, uint32_t capacity, hipStream_t s) {
  const uint32_t n = P.prim_count;
  if (!n) return;
  hipLaunchKernelGGL(emit_kernel, dim3((n + 255) / 256), dim3(256), 0, s, P, B.tri_rec, B.cnt, B.bsum, B.pairs,
                     capacity, B.keys[0], B.vals[0]);
}

void launch_raster(const draw_params& P, const draw_buffers& B, int which, hipStream_t s) {
  uint32_t T = static_cast<uint32_t>(P.gtiles_x * P.gtiles_y);
  if (!T) return;
  int threads = (P.gt / 2) * (P.gt / 2);
  size_t lds = sizeof(prim_lds) * threads;
  bool vis = P.early_z && (P.blend_id == BLEND_OFF || P.blend_id == BLEND_DISABLED);
  const bool sasl = !vis && sasl_shades(P);  // in-order shading by the SASL program
#define SRX_RASTER_S(NA, S)                                                                                          \
  do {                                                                                                               \
    if (vis)                                                                                                         \
      hipLaunchKernelGGL((raster_kernel<NA, S, true>), dim3(T), dim3(threads), lds, s, P, B.samplers, B.order, B.vals[which], \
                         B.tstart, B.tend, B.tri_rec, B.tri_attr, B.vis, B.partial);                                 \
    else if (sasl)                                                                                                   \
      hipLaunchKernelGGL((raster_kernel<NA, S, false, true>), dim3(T), dim3(threads), lds, s, P, B.samplers, B.order, \
                         B.vals[which], B.tstart, B.tend, B.tri_rec, B.tri_attr, B.vis, B.partial);                  \
    else                                                                                                             \
      hipLaunchKernelGGL((raster_kernel<NA, S, false>), dim3(T), dim3(threads), lds, s, P, B.samplers, B.order,      \
                         B.vals[which], B.tstart, B.tend, B.tri_rec, B.tri_attr, B.vis, B.partial);                  \
  } while (0)
#define SRX_RASTER(NA)                     \
  switch (P.samples) {                     \
  case 1: SRX_RASTER_S(NA, 1); break;      \
  case 2: SRX_RASTER_S(NA, 2); break;      \
  default: SRX_RASTER_S(NA, 4); break;     \
  }
  SRX_NA_DISPATCH(P.nattr, SRX_RASTER)
#undef SRX_RASTER
#undef SRX_RASTER_S
  if (P.count_stats) {
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(256), 0, s, B.partial, T, B.stats, static_cast<int>(ST_PS_INVOCATIONS));
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(256), 0, s, B.partial, T, B.stats, static_cast<int>(ST_BACKEND_PIXELS));
  }
}

// Visibility mode of an opaque early-Z draw for vis_kernel, or -1 when only the in-order kernel
// reproduces the reference (NOT_EQUAL with depth writes).
int vis_mode(const draw_params& P) {
  if (P.vis_stencil && !(P.depth_enable && P.write_depth && (P.depth_func == 1 || P.depth_func == 3 ||
                                                            P.depth_func == 4 || P.depth_func == 6)))
    return -1;  // a stencil mask is expressed through the depth keys only
  if (!P.depth_enable) return VIS_LATEST;
  if (P.depth_func == 0) return VIS_NONE;
  if (!P.write_depth) return VIS_LATEST;
  switch (P.depth_func) {
  case 1: return VIS_LESS;
  case 3: return VIS_LEQUAL;
  case 4: return VIS_GREATER;
  case 6: return VIS_GEQUAL;
  case 2:
  case 7: return VIS_LATEST;  // EQUAL keeps the depth it compares against; ALWAYS ignores it
  default: return -1;
  }
}

void vis_table_bounds(uint32_t T, uint32_t records, uint32_t* entries, uint32_t* merges, uint32_t* slots) {
  const uint32_t chunks = (records + VIS_PART - 1) / VIS_PART;
  *entries = T + chunks;            // sum over tiles of max(1, ceil(n / VIS_PART))
  *merges = records / (VIS_PART + 1) + 1;  // tiles with more than VIS_PART records
  *slots = 2 * chunks + 1;          // parts of split tiles: ceil(n / VIS_PART) <= 2n / VIS_PART
}

bool launch_vis(const draw_params& P, const draw_buffers& B, int which, bool shade, uint32_t records, hipStream_t s) {
  int mode = vis_mode(P);
  // a SASL pixel shader runs in its own pass (launch_shade) over the winners this pass writes
  uint32_t* const wout = sasl_shades(P) ? B.vis : nullptr;
  if (wout) shade = false;
  uint32_t T = static_cast<uint32_t>(P.gtiles_x * P.gtiles_y);
  if (mode < 0 || P.gt != 16) return false;
  if (!T) return true;
  uint32_t ne = T, nm = 0, nsl;
  if (!shade) {
    vis_table_bounds(T, records, &ne, &nm, &nsl);
    hipLaunchKernelGGL(vis_table_kernel, dim3(1), dim3(1024), 0, s, B.tstart, B.tend, B.order, T, B.vis_entries,
                       B.vis_merges, B.vis_counts);
  }
  const uint4* entries = shade ? nullptr : B.vis_entries;
#define SRX_VIS_LAUNCH(NA, S, SH)                                                                                 \
  for (int pass = 0; pass < (SH ? 1 : 2); ++pass)                                                                 \
    hipLaunchKernelGGL((vis_kernel<NA, S, SH>), dim3(pass == 0 ? ne : nm), dim3(SRX_VIS_THREADS), 0, s, P, mode, pass, \
                       B.order, entries, B.vis_merges, B.vis_counts, B.vis_partial, B.vals[which], B.tstart,      \
                       B.tend, B.tri_rec, B.tri_attr, B.samplers, wout)
#define SRX_VIS_SHADE(NA)                   \
  switch (P.samples) {                      \
  case 1: SRX_VIS_LAUNCH(NA, 1, true); break; \
  case 2: SRX_VIS_LAUNCH(NA, 2, true); break; \
  default: SRX_VIS_LAUNCH(NA, 4, true); break; \
  }
  if (shade) {
    SRX_NA_DISPATCH(P.nattr, SRX_VIS_SHADE)
  } else {
    switch (P.samples) {
    case 1: SRX_VIS_LAUNCH(1, 1, false); break;
    case 2: SRX_VIS_LAUNCH(1, 2, false); break;
    default: SRX_VIS_LAUNCH(1, 4, false); break;
    }
  }
#undef SRX_VIS_SHADE
#undef SRX_VIS_LAUNCH
  return true;
}

bool vis_applicable(const draw_params& P) { return vis_mode(P) >= 0 && P.gt == 16; }

// One thread per quad: pixel_shader_unit::update/execute on caller-supplied inputs (the quad
// evaluation behind srx_ps_execute_quads).  in: [quad][pixel][input] float4, out: [quad][pixel][output].
__global__ void __launch_bounds__(64) sasl_quads_kernel(const sasl_program* __restrict__ prog,
                                                        const ps_samplers* __restrict__ SM, uint32_t has_mask,
                                                        const float4* __restrict__ in, float4* __restrict__ out,
                                                        uint32_t nquads) {
  const uint32_t qi = blockIdx.x * blockDim.x + threadIdx.x;
  if (qi >= nquads) return;
  const uint32_t nin = prog->ninputs, nout = prog->noutputs, nregs = prog->nregs;
  sasl_regs R;
  for (uint32_t r = 0; r < nregs; ++r)
    for (int px = 0; px < 4; ++px) R.r[r][px] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  for (int px = 0; px < 4; ++px)
    for (uint32_t k = 0; k < nin; ++k) {
      const sasl_input inp = prog->inputs[k];
      const float4 v = in[(static_cast<size_t>(qi) * 4 + px) * nin + k];
      R.r[inp.reg][px] = make_float4(v.x, inp.ncomp > 1 ? v.y : 0.0f, inp.ncomp > 2 ? v.z : 0.0f, inp.ncomp > 3 ? v.w : 0.0f);
    }
  sasl_exec(prog, R, [&](int slot, int, float u, float v, float dxu, float dxv, float dyu, float dyv) -> float4 {
    if (!((has_mask >> slot) & 1u)) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const sampler_desc_dev& sp = SM->s[slot];
    return sp.fast_aniso ? sample_2d_grad_fast(sp, u, v, dxu, dxv, dyu, dyv, 0.0f)
                         : sample_2d_grad(sp, u, v, dxu, dxv, dyu, dyv, 0.0f);
  });
  for (int px = 0; px < 4; ++px)
    for (uint32_t o = 0; o < nout; ++o) {
      const float4 v = R.r[prog->outputs[o].reg][px];
      const int nc = prog->outputs[o].ncomp;
      out[(static_cast<size_t>(qi) * 4 + px) * nout + o] =
          make_float4(v.x, nc > 1 ? v.y : 0.0f, nc > 2 ? v.z : 0.0f, nc > 3 ? v.w : 0.0f);
    }
}

void launch_sasl_quads(const sasl_program* prog, const ps_samplers* SM, uint32_t has_mask, const float4* in, float4* out,
                       uint32_t nquads, hipStream_t s) {
  if (!nquads) return;
  hipLaunchKernelGGL(sasl_quads_kernel, dim3((nquads + 63) / 64), dim3(64), 0, s, prog, SM, has_mask, in, out, nquads);
}

bool sasl_shades(const draw_params& P) {
  return P.ps_id == PS_SASL && P.color.data != nullptr && P.blend_id != BLEND_DISABLED;
}

bool needs_shading(const draw_params& P) {
  return (P.early_z || P.vis_stencil) && P.blend_id == BLEND_OFF && P.color.data != nullptr;
}

// Covers exactly the 16x16 GPU tiles vis_kernel wrote: pixels of the target beyond the viewport's
// 64x64 tile grid hold no winners of this draw.
void launch_shade(const draw_params& P, const draw_buffers& B, hipStream_t s) {
  if (P.gtiles_x <= 0 || P.gtiles_y <= 0) return;
  dim3 grid(P.gtiles_x, P.gtiles_y);
  if (P.ps_id == PS_SASL) {
#define SRX_SASL_SHADE_S(NA, S) \
  hipLaunchKernelGGL((sasl_shade_kernel<NA, S>), grid, dim3(64), 0, s, P, B.samplers, B.tri_attr, B.vis)
#define SRX_SASL_SHADE(NA)                    \
  switch (P.samples) {                        \
  case 1: SRX_SASL_SHADE_S(NA, 1); break;     \
  case 2: SRX_SASL_SHADE_S(NA, 2); break;     \
  default: SRX_SASL_SHADE_S(NA, 4); break;    \
  }
    SRX_NA_DISPATCH(P.nattr, SRX_SASL_SHADE)
#undef SRX_SASL_SHADE
#undef SRX_SASL_SHADE_S
    return;
  }
#define SRX_SHADE_S(NA, S) \
  hipLaunchKernelGGL((shade_kernel<NA, S>), grid, dim3(256), 0, s, P, B.samplers, B.tri_attr, B.vis)
#define SRX_SHADE(NA)                    \
  switch (P.samples) {                   \
  case 1: SRX_SHADE_S(NA, 1); break;     \
  case 2: SRX_SHADE_S(NA, 2); break;     \
  default: SRX_SHADE_S(NA, 4); break;    \
  }
  SRX_NA_DISPATCH(P.nattr, SRX_SHADE)
#undef SRX_SHADE
#undef SRX_SHADE_S
}

// ============================================================== clears, resolve, mip generation
// surface::fill with the converted colour: the texel bytes are built once and stored 16 bytes at a
// time (texel sizes 4, 8 and 16 tile a 16-byte word; RGB32F takes the per-texel path).
__global__ void clear_kernel(surface_desc s, float4 v) {
  const size_t n = static_cast<size_t>(s.w) * s.h * s.samples;
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int es = pf_size(s.pf);
  if (es == 12) {
    if (i < n) store_texel(s.pf, s.data + i * 12, v);
    return;
  }
  const uint4 w = fill_words(s.pf, v);
  const size_t words = n * es / 16, tail = n * es % 16;
  if (i < words) reinterpret_cast<uint4*>(s.data)[i] = w;
  if (i == 0 && tail) {  // whole texels (es divides 16): the tail starts a fresh 16-byte word
    const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
    for (size_t b = 0; b < tail; b += 4) *reinterpret_cast<uint32_t*>(s.data + words * 16 + b) = ws[b / 4];
  }
}

// depth and/or stencil of every RG32F sample; both at once are 16-byte stores of two samples
__global__ void clear_ds_kernel(surface_desc s, uint32_t flags, float depth, uint32_t stencil) {
  const size_t n = static_cast<size_t>(s.w) * s.h * s.samples;
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (flags == 3) {
    const uint32_t d = __float_as_uint(depth);
    if (i < n / 2) reinterpret_cast<uint4*>(s.data)[i] = make_uint4(d, stencil, d, stencil);
    if (i == 0 && (n & 1)) {
      float* p = reinterpret_cast<float*>(s.data + (n - 1) * 8);
      p[0] = depth;
      p[1] = __uint_as_float(stencil);
    }
    return;
  }
  if (i >= n) return;
  float* p = reinterpret_cast<float*>(s.data + i * 8);
  if (flags & 1) p[0] = depth;
  if (flags & 2) p[1] = __uint_as_float(stencil);
}

// surface::resolve (surface.cpp:123-140): sum samples in order, then multiply by 1/S
__global__ void resolve_kernel(surface_desc src, surface_desc dst) {
  int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= src.w) return;
  float4 c = make_float4(0, 0, 0, 0);
  for (int s = 0; s < src.samples; ++s) {
    float4 t = load_texel(src.pf, sample_addr(src, x, y, s));
    c = make_float4(c.x + t.x, c.y + t.y, c.z + t.z, c.w + t.w);
  }
  float inv = 1 / static_cast<float>(src.samples);
  store_texel(dst.pf, sample_addr(dst, x, y, 0), make_float4(c.x * inv, c.y * inv, c.z * inv, c.w * inv));
}

// The same for 4-sample RGBA8 into RGBA8: one 16-byte load per pixel.
__global__ void resolve_rgba8x4_kernel(surface_desc src, surface_desc dst) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<size_t>(src.w) * src.h) return;
  const uint4 q = reinterpret_cast<const uint4*>(src.data)[i];
  const uint32_t w[4] = {q.x, q.y, q.z, q.w};
  float4 c = make_float4(0, 0, 0, 0);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const uint32_t u = w[s];
    const float4 t = make_float4(static_cast<float>(u & 0xFF) * (1.0f / 255), static_cast<float>((u >> 8) & 0xFF) * (1.0f / 255),
                                 static_cast<float>((u >> 16) & 0xFF) * (1.0f / 255), static_cast<float>(u >> 24) * (1.0f / 255));
    c = make_float4(c.x + t.x, c.y + t.y, c.z + t.z, c.w + t.w);
  }
  const float inv = 1 / 4.0f;
  store_texel(PF_RGBA8, dst.data + i * 4, make_float4(c.x * inv, c.y * inv, c.z * inv, c.w * inv));
}

// surface::make_mip_surface (surface.cpp:53-92); odd-size edge reads clamp to the last texel
__global__ void mip_kernel(surface_desc src, surface_desc dst, int filter) {
  int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= dst.w) return;
  for (int s = 0; s < dst.samples; ++s) {
    int x0 = x * 2, y0 = y * 2, x1 = min(x0 + 1, src.w - 1), y1 = min(y0 + 1, src.h - 1);
    float4 c;
    if (filter == FILTER_POINT) {
      c = load_texel(src.pf, sample_addr(src, x0, y0, s));
    } else {
      float4 c0 = load_texel(src.pf, sample_addr(src, x0, y0, s)), c1 = load_texel(src.pf, sample_addr(src, x1, y0, s));
      float4 c2 = load_texel(src.pf, sample_addr(src, x0, y1, s)), c3 = load_texel(src.pf, sample_addr(src, x1, y1, s));
      c = make_float4(((c0.x + c1.x) + c2.x) + c3.x, ((c0.y + c1.y) + c2.y) + c3.y, ((c0.z + c1.z) + c2.z) + c3.z,
                      ((c0.w + c1.w) + c2.w) + c3.w);
      c = make_float4(c.x * 0.25f, c.y * 0.25f, c.z * 0.25f, c.w * 0.25f);
    }
    store_texel(dst.pf, sample_addr(dst, x, y, s), c);
  }
}

// ============================================================== batched jobs
// Every stage of a batch of independent draws is one launch: each workgroup looks up its job and
// its index within the job's grid (stage_map) and runs the single-draw kernel body on the job's
// parameters and buffers.  Record counts are read on the device (from the batch-wide scan of the
// per-triangle counts), clamped to the capacities the batch was prepared with, so a replay whose
// geometry changed stays in bounds and reports the overflow through the status word.
__device__ __forceinline__ uint32_t stage_job(const stage_map& m, uint32_t& local, uint32_t& blocks) {
  const uint32_t j = m.job[blockIdx.x];
  const uint2 r = m.range[j];
  local = blockIdx.x - r.x;
  blocks = r.y;
  return j;
}

__global__ void batch_reset_kernel(uint32_t* __restrict__ counters, uint32_t n, uint32_t* __restrict__ cnt_tail,
                                   uint32_t* __restrict__ hist_tail, uint32_t* __restrict__ status) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) counters[i] = 0;
  if (i == 0) {
    *cnt_tail = 0;
    *hist_tail = 0;
    *status = 0;
  }
}

__global__ void __launch_bounds__(256) fill_batch_kernel(const fill_item* __restrict__ items, stage_map m) {
  uint32_t lb, nb;
  const fill_item& it = items[stage_job(m, lb, nb)];
  const size_t i = static_cast<size_t>(lb) * 256 + threadIdx.x;
  const uint4 w = fill_words(it.pf, it.value);
  if (i < it.words) reinterpret_cast<uint4*>(it.data)[i] = w;
  if (i == 0 && it.tail) {
    const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
    for (uint32_t b = 0; b < it.tail; b += 4) *reinterpret_cast<uint32_t*>(it.data + it.words * 16 + b) = ws[b / 4];
  }
}

__global__ void __launch_bounds__(256) resolve_batch_kernel(const resolve_item* __restrict__ items, stage_map m) {
  uint32_t lb, nb;
  const resolve_item& it = items[stage_job(m, lb, nb)];
  const surface_desc src = it.src, dst = it.dst;
  const size_t i = static_cast<size_t>(lb) * 256 + threadIdx.x;
  if (i >= static_cast<size_t>(src.w) * src.h) return;
  const int x = static_cast<int>(i % src.w), y = static_cast<int>(i / src.w);
  float4 c = make_float4(0, 0, 0, 0);
  for (int smp = 0; smp < src.samples; ++smp) {
    const float4 t = load_texel(src.pf, sample_addr(src, x, y, smp));
    c = make_float4(c.x + t.x, c.y + t.y, c.z + t.z, c.w + t.w);
  }
  const float inv = 1 / static_cast<float>(src.samples);
  store_texel(dst.pf, sample_addr(dst, x, y, 0), make_float4(c.x * inv, c.y * inv, c.z * inv, c.w * inv));
}

__global__ void __launch_bounds__(256) vs_batch_kernel(const batch_job* __restrict__ jobs, stage_map m) {
  uint32_t lb, nb;
  const batch_job& J = jobs[stage_job(m, lb, nb)];
  vs_body(J.P, J.B.vbuf, lb);
}

template <int NA>
__global__ void __launch_bounds__(256) setup_batch_kernel(const batch_job* __restrict__ jobs, stage_map m) {
  uint32_t lb, nb;
  const batch_job& J = jobs[stage_job(m, lb, nb)];
  setup_body<NA>(J.P, J.B.vbuf, J.B.tri_rec, J.B.tri_attr, J.B.cnt, J.B.clip_flag, J.B.counters, J.B.wide_list, J.B.partial,
                 nullptr, nullptr, lb);
}

template <int NA>
__global__ void clip_batch_kernel(const batch_job* __restrict__ jobs, stage_map m) {
  uint32_t lb, nb;
  const batch_job& J = jobs[stage_job(m, lb, nb)];
  clip_body<NA>(J.P, J.B.vbuf, J.B.tri_rec, J.B.tri_attr, J.B.cnt, J.B.clip_flag, J.B.clip_slot, J.B.clip_verts,
                J.B.clip_hdr, J.B.clip_capacity, J.B.counters, J.B.stats, J.B.wide_list, nullptr, lb, nb);
}

// The job's counts are followed by the next job's first scanned count (or the batch total): the
// end of its last triangle's records.
__global__ void __launch_bounds__(256) emit_batch_kernel(const batch_job* __restrict__ jobs, stage_map m, uint32_t capacity) {
  uint32_t lb, nb;
  const batch_job& J = jobs[stage_job(m, lb, nb)];
  emit_body(J.P, J.B.tri_rec, J.B.cnt, J.B.cnt + 3 * J.P.prim_count, capacity, J.B.keys[0], J.B.vals[0], lb);
}

// The job's records [begin, begin + n): its scanned counts clamped to the record capacity and to the
// records its binning workgroups cover.
__device__ __forceinline__ void job_records(const batch_job& J, uint32_t capacity, uint32_t items, uint32_t& begin,
                                            uint32_t& n) {
  const uint32_t b = min(J.B.cnt[0], capacity), e = min(J.B.cnt[3 * J.P.prim_count], capacity);
  begin = b;
  n = min(e - b, J.bucket_blocks * items);
}

template <int PT>
__global__ void __launch_bounds__(256) bucket_hist_batch_kernel(const batch_job* __restrict__ jobs, stage_map m,
                                                                uint32_t capacity) {
  uint32_t lb, nb, begin, n;
  const batch_job& J = jobs[stage_job(m, lb, nb)];
  job_records(J, capacity, 256 * PT, begin, n);
  bucket_hist_body<PT>(J.B.keys[0] + begin, n, J.tiles, J.bucket_blocks, J.B.hist, lb);
}

template <int PT>
__global__ void __launch_bounds__(256) bucket_scatter_batch_kernel(const batch_job* __restrict__ jobs, stage_map m,
                                                                   uint32_t capacity) {
  uint32_t lb, nb, begin, n;
  const batch_job& J = jobs[stage_job(m, lb, nb)];
  job_records(J, capacity, 256 * PT, begin, n);
  bucket_scatter_body<PT>(J.B.keys[0] + begin, J.B.vals[0] + begin, n, J.tiles, J.bucket_blocks, J.B.hist, J.B.vals[1],
                          capacity, lb);
}

// The job's histogram is followed by the next job's first scanned entry (or the batch total).
__global__ void __launch_bounds__(256) bucket_range_batch_kernel(const batch_job* __restrict__ jobs, stage_map m) {
  uint32_t lb, nb;
  const batch_job& J = jobs[stage_job(m, lb, nb)];
  bucket_range_body(J.B.hist, J.B.hist[J.tiles * J.bucket_blocks], J.tiles, J.bucket_blocks, J.B.tstart, J.B.tend,
                    lb * 256 + threadIdx.x);
}

__global__ void __launch_bounds__(1024) tile_order_batch_kernel(const batch_job* __restrict__ jobs) {
  const batch_job& J = jobs[blockIdx.x];
  tile_order_body(J.B.tstart, J.B.tend, J.tiles, const_cast<uint32_t*>(J.B.order));
}

template <int NA, int S, bool SHADE, int GT>
__global__ void __launch_bounds__(256, 4) vis_batch_kernel(const batch_job* __restrict__ jobs, stage_map m) {
  uint32_t lb, nb;
  const batch_job& J = jobs[stage_job(m, lb, nb)];
  vis_body<NA, S, SHADE, 0, GT>(J.P, J.vis_mode, 0, J.B.order, nullptr, nullptr, nullptr, nullptr, J.B.vals[1], J.B.tstart,
                            J.B.tend, J.B.tri_rec, J.B.tri_attr, J.B.samplers, lb);
}

// status bit 0: a job's records exceeded the capacity it was prepared with; bit 1: clip overflow
__global__ void batch_check_kernel(const batch_job* __restrict__ jobs, uint32_t njobs, uint32_t capacity, uint32_t items,
                                   uint32_t* __restrict__ status) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= njobs) return;
  const batch_job& J = jobs[j];
  const uint32_t b = J.B.cnt[0], e = J.B.cnt[3 * J.P.prim_count];
  uint32_t st = 0;
  if (e > capacity || e - b > J.bucket_blocks * items) st |= 1u;
  if (J.B.counters[CT_CLIP_OVERFLOW]) st |= 2u;
  if (st) atomicOr(status, st);
}

// records[j] of every job (unclamped), for sizing a batch
__global__ void batch_counts_kernel(const batch_job* __restrict__ jobs, uint32_t njobs, uint32_t* __restrict__ out) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= njobs) return;
  const batch_job& J = jobs[j];
  out[j] = J.B.cnt[3 * J.P.prim_count] - J.B.cnt[0];
  out[njobs + j] = J.B.counters[CT_CLIP_OVERFLOW];
}
void launch_batch_counts(const batch_job* jobs, uint32_t njobs, uint32_t* out, hipStream_t s) {
  if (njobs) hipLaunchKernelGGL(batch_counts_kernel, dim3((njobs + 255) / 256), dim3(256), 0, s, jobs, njobs, out);
}

void launch_batch_reset(uint32_t* counters, uint32_t n, uint32_t* cnt_tail, uint32_t* hist_tail, uint32_t* status,
                        hipStream_t s) {
  hipLaunchKernelGGL(batch_reset_kernel, dim3((std::max<uint32_t>(n, 1) + 255) / 256), dim3(256), 0, s, counters, n,
                     cnt_tail, hist_tail, status);
}
void launch_fill_batch(const fill_item* items, stage_map m, uint32_t blocks, hipStream_t s) {
  if (blocks) hipLaunchKernelGGL(fill_batch_kernel, dim3(blocks), dim3(256), 0, s, items, m);
}
void launch_resolve_batch(const resolve_item* items, stage_map m, uint32_t blocks, hipStream_t s) {
  if (blocks) hipLaunchKernelGGL(resolve_batch_kernel, dim3(blocks), dim3(256), 0, s, items, m);
}
void launch_vs_batch(const batch_job* jobs, stage_map m, uint32_t blocks, hipStream_t s) {
  if (blocks) hipLaunchKernelGGL(vs_batch_kernel, dim3(blocks), dim3(256), 0, s, jobs, m);
}
void launch_setup_batch(int na, const batch_job* jobs, stage_map m, uint32_t blocks, hipStream_t s) {
  if (!blocks) return;
#define SRX_SETUP_B(NA) hipLaunchKernelGGL(setup_batch_kernel<NA>, dim3(blocks), dim3(256), 0, s, jobs, m)
  SRX_NA_DISPATCH_BATCH(na, SRX_SETUP_B)
#undef SRX_SETUP_B
}
void launch_clip_batch(int na, const batch_job* jobs, stage_map m, uint32_t blocks, hipStream_t s) {
  if (!blocks) return;
#define SRX_CLIP_B(NA) hipLaunchKernelGGL(clip_batch_kernel<NA>, dim3(blocks), dim3(64), 0, s, jobs, m)
  SRX_NA_DISPATCH_BATCH(na, SRX_CLIP_B)
#undef SRX_CLIP_B
}
void launch_emit_batch(const batch_job* jobs, stage_map m, uint32_t blocks, uint32_t capacity, hipStream_t s) {
  if (blocks) hipLaunchKernelGGL(emit_batch_kernel, dim3(blocks), dim3(256), 0, s, jobs, m, capacity);
}
bool batch_bin_items_valid(int pt) { return pt == 8 || pt == 16 || pt == 32 || pt == 64; }
void launch_bucket_batch(int pt, int phase, const batch_job* jobs, stage_map m, uint32_t blocks, uint32_t capacity,
                         hipStream_t s) {
  if (!blocks) return;
#define SRX_BUCKET_B(PT)                                                                                          \
  if (phase == 0)                                                                                                  \
    hipLaunchKernelGGL(bucket_hist_batch_kernel<PT>, dim3(blocks), dim3(256), 0, s, jobs, m, capacity);          \
  else                                                                                                             \
    hipLaunchKernelGGL(bucket_scatter_batch_kernel<PT>, dim3(blocks), dim3(256), 0, s, jobs, m, capacity)
  switch (pt) {
  case 8: SRX_BUCKET_B(8); break;
  case 32: SRX_BUCKET_B(32); break;
  case 64: SRX_BUCKET_B(64); break;
  default: SRX_BUCKET_B(16); break;
  }
#undef SRX_BUCKET_B
}
void launch_range_batch(const batch_job* jobs, stage_map m, uint32_t blocks, hipStream_t s) {
  if (blocks) hipLaunchKernelGGL(bucket_range_batch_kernel, dim3(blocks), dim3(256), 0, s, jobs, m);
}
void launch_tile_order_batch(const batch_job* jobs, uint32_t njobs, hipStream_t s) {
  if (njobs) hipLaunchKernelGGL(tile_order_batch_kernel, dim3(njobs), dim3(1024), 0, s, jobs);
}
// GPU tiles of 8 and 32 pixels are built for single-sample draws; every variant must fit 64 KB of LDS.
bool vis_batch_supported(int samples, bool shade, int threads, int gt) {
  if (!(threads == 64 || threads == 128 || threads == 256)) return false;
  if (gt != 16 && !(gt == 8 || gt == 32)) return false;
  if (gt != 16 && samples != 1) return false;
  return vis_static_bytes(samples, gt) + vis_union_bytes(threads, samples, shade, gt) <= 65536;
}
bool launch_vis_batch(int na, int samples, bool shade, int threads, int gt, const batch_job* jobs, stage_map m,
                      uint32_t blocks, hipStream_t s) {
  if (!vis_batch_supported(samples, shade, threads, gt)) return false;
  if (!blocks) return true;
  const size_t lds = static_cast<size_t>(vis_union_bytes(threads, samples, shade, gt));
#define SRX_VIS_B(NA, S, SH, GT) \
  hipLaunchKernelGGL((vis_batch_kernel<NA, S, SH, GT>), dim3(blocks), dim3(threads), lds, s, jobs, m)
#define SRX_VIS_BS(NA)                       \
  switch (samples) {                         \
  case 1: SRX_VIS_B(NA, 1, true, 16); break; \
  case 2: SRX_VIS_B(NA, 2, true, 16); break; \
  default: SRX_VIS_B(NA, 4, true, 16); break; \
  }
#define SRX_VIS_B8(NA) SRX_VIS_B(NA, 1, true, 8)
#define SRX_VIS_B32(NA) SRX_VIS_B(NA, 1, true, 32)
  if (gt == 8) {
    if (shade) {
      SRX_NA_DISPATCH_BATCH(na, SRX_VIS_B8)
    } else {
      SRX_VIS_B(1, 1, false, 8);
    }
  } else if (gt == 32) {
    if (shade) {
      SRX_NA_DISPATCH_BATCH(na, SRX_VIS_B32)
    } else {
      SRX_VIS_B(1, 1, false, 32);
    }
  } else if (shade) {
    SRX_NA_DISPATCH_BATCH(na, SRX_VIS_BS)
  } else {
    switch (samples) {
    case 1: SRX_VIS_B(1, 1, false, 16); break;
    case 2: SRX_VIS_B(1, 2, false, 16); break;
    default: SRX_VIS_B(1, 4, false, 16); break;
    }
  }
#undef SRX_VIS_B32
#undef SRX_VIS_B8
#undef SRX_VIS_BS
#undef SRX_VIS_B
  return true;
}
void launch_batch_check(const batch_job* jobs, uint32_t njobs, uint32_t capacity, int pt, uint32_t* status, hipStream_t s) {
  if (njobs)
    hipLaunchKernelGGL(batch_check_kernel, dim3((njobs + 255) / 256), dim3(256), 0, s, jobs, njobs, capacity,
                       static_cast<uint32_t>(256 * pt), status);
}

void launch_clear(const surface_desc& s, float4 v, hipStream_t st) {
  const size_t n = static_cast<size_t>(s.w) * s.h * s.samples;
  const size_t threads = pf_size(s.pf) == 12 ? n : std::max<size_t>(n * pf_size(s.pf) / 16, 1);
  if (n) hipLaunchKernelGGL(clear_kernel, dim3(static_cast<unsigned>((threads + 255) / 256)), dim3(256), 0, st, s, v);
}
void launch_clear_ds(const surface_desc& s, uint32_t flags, float depth, uint32_t stencil, hipStream_t st) {
  const size_t n = static_cast<size_t>(s.w) * s.h * s.samples;
  const size_t threads = flags == 3 ? std::max<size_t>(n / 2, 1) : n;
  if (n) hipLaunchKernelGGL(clear_ds_kernel, dim3(static_cast<unsigned>((threads + 255) / 256)), dim3(256), 0, st, s, flags, depth, stencil);
}
void launch_resolve(const surface_desc& src, const surface_desc& dst, hipStream_t st) {
  if (src.pf == PF_RGBA8 && dst.pf == PF_RGBA8 && src.samples == 4 && dst.samples == 1 && src.w == dst.w && src.h == dst.h) {
    const size_t n = static_cast<size_t>(src.w) * src.h;
    if (n) hipLaunchKernelGGL(resolve_rgba8x4_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, src, dst);
    return;
  }
  hipLaunchKernelGGL(resolve_kernel, dim3((src.w + 255) / 256, src.h), dim3(256), 0, st, src, dst);
}
void launch_mip(const surface_desc& src, const surface_desc& dst, int filter, hipStream_t st) {
  hipLaunchKernelGGL(mip_kernel, dim3((dst.w + 255) / 256, dst.h), dim3(256), 0, st, src, dst, filter);
}

}  // namespace srx
