// Host runtime of the HIP backend: the srx C ABI (include/srx.h) over device resources.
//
// One HIP stream per renderer carries every command in submission order (the reference's
// async renderer queue, async_renderer.cpp:17-110); srx_flush() waits for it.  Resources live
// in HBM in the reference's linear layout ((y*w + x)*samples + s)*elem_size (surface.cpp:277-295),
// so map/read-back bytes are identical.  State validation follows renderer_impl.cpp.
#include "../../include/srx.h"
#include "srx_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

using namespace srx;

namespace {

thread_local int32_t g_last_error = 0;
thread_local std::string g_last_error_str;

int32_t fail(int32_t code, const char* msg) {
  g_last_error = 0;
  g_last_error_str = msg;
  return code;
}
int32_t hip_fail(hipError_t e, const char* where) {
  g_last_error = static_cast<int32_t>(e);
  g_last_error_str = std::string(where) + ": " + hipGetErrorString(e);
  return SRX_E_HIP;
}
#define HIPCHK(expr)                                 \
  do {                                               \
    hipError_t _e = (expr);                          \
    if (_e != hipSuccess) return hip_fail(_e, #expr); \
  } while (0)

// ---------------------------------------------------------------- objects
enum obj_kind { K_RENDERER = 1, K_BUFFER, K_TEXTURE, K_SURFACE, K_SAMPLER, K_LAYOUT, K_RASTER, K_DS, K_QUERY, K_BATCH, K_SERVER };

struct object {
  std::atomic<int> refs{1};
  int kind;
  explicit object(int k) : kind(k) {}
  virtual ~object() = default;
};
void retain(object* o) {
  if (o) o->refs.fetch_add(1);
}
void release(object* o) {
  if (o && o->refs.fetch_sub(1) == 1) delete o;
}
template <class T>
struct ref {
  T* p = nullptr;
  ref() = default;
  explicit ref(T* q) : p(q) { retain(p); }
  ref(const ref& o) : p(o.p) { retain(p); }
  ref& operator=(const ref& o) {
    if (this != &o) {
      retain(o.p);
      release(p);
      p = o.p;
    }
    return *this;
  }
  ~ref() { release(p); }
  T* operator->() const { return p; }
  T* get() const { return p; }
  explicit operator bool() const { return p != nullptr; }
};

int pf_bytes(int pf) {
  switch (pf) {
  case PF_RGBA32F: return 16;
  case PF_RGB32F: return 12;
  case PF_RG32F: return 8;
  default: return 4;
  }
}

struct buffer_obj : object {
  uint8_t* d = nullptr;
  size_t size = 0;
  buffer_obj() : object(K_BUFFER) {}
  ~buffer_obj() override {
    if (d) (void)hipFree(d);
  }
};

struct surface_obj : object {
  int w, h, samples, pf;
  uint8_t* d = nullptr;
  size_t bytes = 0;
  surface_obj(int w_, int h_, int s_, int pf_) : object(K_SURFACE), w(w_), h(h_), samples(s_), pf(pf_) {}
  ~surface_obj() override {
    if (d) (void)hipFree(d);
  }
  surface_desc desc() const { return surface_desc{d, w, h, samples, pf}; }
};

int32_t make_surface(int w, int h, int s, int pf, surface_obj** out) {
  auto* o = new surface_obj(w, h, s, pf);
  o->bytes = static_cast<size_t>(w) * h * s * pf_bytes(pf);
  hipError_t e = hipMalloc(&o->d, o->bytes ? o->bytes : 1);
  if (e != hipSuccess) {
    delete o;
    return hip_fail(e, "hipMalloc(surface)");
  }
  e = hipMemset(o->d, 0, o->bytes);
  if (e != hipSuccess) {
    delete o;
    return hip_fail(e, "hipMemset(surface)");
  }
  *out = o;
  return SRX_OK;
}

struct texture_obj : object {
  int pf, samples;
  int max_lod = 0, min_lod = 0;
  uint32_t size[3];
  std::vector<ref<surface_obj>> surfs;
  texture_obj() : object(K_TEXTURE) {}
};

struct sampler_obj : object {
  srx_sampler_desc desc;
  ref<texture_obj> tex;
  sampler_obj() : object(K_SAMPLER) {}
};

struct layout_elem {
  std::string sem;
  uint32_t index, fmt, slot, offset;
};
struct layout_obj : object {
  std::vector<layout_elem> elems;
  layout_obj() : object(K_LAYOUT) {}
};
struct raster_obj : object {
  srx_raster_desc desc;
  raster_obj() : object(K_RASTER) {}
};
struct ds_obj : object {
  srx_depth_stencil_desc desc;
  ds_obj() : object(K_DS) {}
};
// pipeline_profiles slots (async_object.h:187-196): gather_vtx, vtx_proc, clipping, compact_clip,
// vp_trans, tri_dispatch, ras; slot 7 counts the profiled draws.  Slots 8 and 9 split ras into
// its visibility/in-order raster kernel and the deferred shading kernel (returned when the caller's
// buffer holds them).
enum { PROF_GATHER = 0, PROF_VTX, PROF_CLIP, PROF_COMPACT, PROF_VP, PROF_DISPATCH, PROF_RAS, PROF_DRAWS,
       PROF_RAS_VIS, PROF_RAS_SHADE, PROF_COUNT };

struct query_obj : object {
  uint32_t id;
  bool running = false, has_data = false;
  uint64_t begin_snap[ST_COUNT] = {};
  srx_pipeline_statistics stats{};
  srx_internal_statistics istats{};
  uint64_t prof[PROF_COUNT] = {};
  query_obj() : object(K_QUERY) {}
};

// HIP events bracketing the stages of one profiled draw, on the renderer's stream:
// [0,1) vertex shader, [1,2) setup + clipping (viewport transform fused), [2,3) compaction scan,
// [3,4) record emission + tile sort, [4,5) visibility or in-order raster, [5,6) deferred shading.
struct draw_events {
  hipEvent_t ev[7];
};

std::string lower(const char* s) {
  std::string r(s ? s : "");
  for (auto& c : r)
    if ('A' <= c && c <= 'Z') c = static_cast<char>(c - 'A' + 'a');
  return r;
}
// semantic_value name normalisation (shader/constants.h:52-83)
std::string canon_sem(const std::string& l) {
  if (l == "position" || l == "sv_position") return "position";
  if (l == "color" || l == "sv_target") return "color";
  return l;
}

struct vs_binding {
  const char* sem;
  uint32_t index;
  int reg;
};
int vs_bindings(int id, vs_binding* out) {
  switch (id) {
  case SRX_VS_PASSTHROUGH_COLOR:
  case SRX_VS_WVP_COLOR:
  case SRX_VS_WVP_COLOR_LIGHT:
    out[0] = {"position", 0, 0};
    out[1] = {"color", 0, 1};
    return 2;
  case SRX_VS_WVP_NORMAL_UV:
    out[0] = {"position", 0, 0};
    out[1] = {"normal", 0, 1};
    out[2] = {"texcoord", 0, 2};
    return 3;
  }
  return -1;
}
int vs_output_count(int id) {
  switch (id) {
  case SRX_VS_PASSTHROUGH_COLOR:
  case SRX_VS_WVP_COLOR: return 1;
  case SRX_VS_WVP_NORMAL_UV:
  case SRX_VS_WVP_COLOR_LIGHT: return 2;
  }
  return 0;
}
uint32_t fmt_bytes(uint32_t fmt) {
  switch (fmt) {
  case SRX_FMT_R32_FLOAT: return 4;
  case SRX_FMT_R32G32_FLOAT: return 8;
  case SRX_FMT_R32G32B32_FLOAT: return 12;
  case SRX_FMT_R32G32B32A32_FLOAT:
  case SRX_FMT_R32G32B32A32_SINT: return 16;
  }
  return 0;
}

// Blender's EWA weights (exp(a r) - exp(a)) / (1 - exp(a)), a = -2, r = i/255, printed with six
// significant digits -- the values of the reference's table (sampler.cpp:618-677).
void make_ewa(float* t) {
  for (int i = 0; i < 256; ++i) {
    double r = i / 255.0;
    double w = (std::exp(-2.0 * r) - std::exp(-2.0)) / (1.0 - std::exp(-2.0));
    char buf[64];
    std::snprintf(buf, sizeof buf, "%g", w);
    t[i] = std::strtof(buf, nullptr);
  }
}

// ---------------------------------------------------------------- renderer
struct vb_slot {
  ref<buffer_obj> buf;
  size_t stride = 0, offset = 0;
};

struct render_state {
  vb_slot vbs[SRX_MAX_INPUT_SLOTS];
  ref<buffer_obj> ib;
  uint32_t index_fmt = SRX_FMT_R16_UINT;
  ref<layout_obj> layout;
  uint32_t topo = SRX_TOPOLOGY_TRIANGLE_LIST;
  int vs_id = 0, ps_id = 0, blend_id = SRX_BLEND_OFF;
  float vs_consts[32] = {}, ps_consts[16] = {};
  ref<sampler_obj> samplers[SRX_MAX_SAMPLER_SLOTS];
  srx_raster_desc raster;
  srx_depth_stencil_desc ds;
  int32_t stencil_ref = 0;
  std::vector<ref<surface_obj>> color_targets;
  ref<surface_obj> ds_target;
  srx_viewport target_vp{0, 0, 0, 0, 0, 0};
  size_t target_samples = 0;
  srx_viewport vp{0, 0, 0, 0, 0, 1};
};

template <class T>
struct dev_array {
  T* p = nullptr;
  size_t cap = 0;
};

struct batch_obj;

struct renderer_obj : object {
  render_state st;
  batch_obj* recording = nullptr;  // srx_batch_begin .. srx_batch_end: commands are recorded
  hipStream_t stream = nullptr;
  std::vector<ref<query_obj>> active;
  // per-draw device scratch, grown on demand
  dev_array<float4> vbuf, tri_rec, tri_attr, clip_verts;
  dev_array<ps_samplers> samplers_d;
  ps_samplers* samplers_h = nullptr;  // pinned staging of the per-draw sampler table
  dev_array<uint32_t> cnt, keys0, keys1, vals0, vals1, tstart, tend, hist, scan_tmp, clip_hdr, counters, vis;
  dev_array<uint32_t> order;  // GPU tile scheduling order
  dev_array<uint4> vis_entries, vis_merges;  // visibility work table
  dev_array<uint32_t> vis_counts;
  dev_array<unsigned long long> vis_partial;
  dev_array<uint8_t> clip_flag;
  dev_array<int32_t> clip_slot;
  dev_array<unsigned long long> partial;
  unsigned long long* stats_d = nullptr;
  uint32_t* pinned = nullptr;  // host staging for counter read-back
  uint32_t last_records = 0;   // binned records of the previous draw (sizes the next emission)
  uint64_t host_stats[ST_COUNT] = {};
  std::vector<draw_events> pending_prof;  // recorded while a pipeline_profiles query runs
  std::vector<hipEvent_t> event_pool;
  bool profiling() const {
    for (auto& q : active)
      if (q->id == SRX_QUERY_PIPELINE_PROFILES) return true;
    return false;
  }
  hipEvent_t take_event() {
    if (!event_pool.empty()) {
      hipEvent_t e = event_pool.back();
      event_pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
  renderer_obj() : object(K_RENDERER) {
    srx_default_raster_desc(&st.raster);
    srx_default_depth_stencil_desc(&st.ds);
  }
  ~renderer_obj() override;
};

template <class T>
int32_t ensure(renderer_obj* r, dev_array<T>& a, size_t n) {
  if (n <= a.cap && a.p) return SRX_OK;
  size_t cap = std::max<size_t>(n + n / 2, 1024);
  if (a.p) {
    HIPCHK(hipStreamSynchronize(r->stream));
    HIPCHK(hipFree(a.p));
    a.p = nullptr;
    a.cap = 0;
  }
  HIPCHK(hipMalloc(&a.p, cap * sizeof(T)));
  a.cap = cap;
  return SRX_OK;
}

std::once_flag g_ewa_once;
int32_t g_ewa_status = SRX_OK;

int gpu_tile_size() {
  static int g = [] {
    const char* e = std::getenv("SRX_GPU_TILE");
    int v = e ? std::atoi(e) : 16;
    return (v == 32) ? 32 : 16;
  }();
  return g;
}

void snapshot_stats(renderer_obj* r, uint64_t* out) {
  unsigned long long dev[ST_COUNT] = {};
  (void)hipStreamSynchronize(r->stream);
  (void)hipMemcpy(dev, r->stats_d, sizeof dev, hipMemcpyDeviceToHost);
  for (int i = 0; i < ST_COUNT; ++i) out[i] = r->host_stats[i] + dev[i];
}

// Reads the draw's device counters (waits for the queued work).
int32_t read_counters(renderer_obj* r, uint32_t* out) {
  HIPCHK(hipMemcpyAsync(r->pinned, r->counters.p, CT_COUNT * sizeof(uint32_t), hipMemcpyDeviceToHost, r->stream));
  HIPCHK(hipStreamSynchronize(r->stream));
  for (int i = 0; i < CT_COUNT; ++i) out[i] = r->pinned[i];
  return SRX_OK;
}

#define RCHK(expr)                     \
  do {                                 \
    int32_t _rc = (expr);              \
    if (_rc != SRX_OK) return _rc;     \
  } while (0)

// The draw's parameters and sampler table from the renderer's state (renderer_impl.cpp:159-238
// validation); skip = nothing to draw into.
int32_t build_draw(renderer_obj* r, bool indexed, uint32_t start, uint32_t primcnt, int32_t basevert, draw_params& P,
                   ps_samplers& SM, bool& skip) {
  render_state& st = r->st;
  skip = st.color_targets.empty() && !st.ds_target;  // render_core.cpp:44-46
  if (skip) return SRX_OK;
  if (st.topo != SRX_TOPOLOGY_TRIANGLE_LIST && st.topo != SRX_TOPOLOGY_TRIANGLE_STRIP)
    return fail(SRX_E_UNSUPPORTED, "only triangle lists and strips are rasterised (lines are unimplemented upstream)");
  if (st.raster.fill_mode != SRX_FILL_SOLID) return fail(SRX_E_UNSUPPORTED, "wireframe fill is unimplemented upstream");
  int nout = vs_output_count(st.vs_id);
  if (nout <= 0) return fail(SRX_INVALID_PARAMETER, "no vertex shader bound");
  if (!st.layout) return fail(SRX_INVALID_PARAMETER, "no input layout");
  if (indexed && !st.ib) return fail(SRX_INVALID_PARAMETER, "no index buffer");
  int samples = static_cast<int>(st.target_samples);
  if (samples != 1 && samples != 2 && samples != 4) return fail(SRX_E_UNSUPPORTED, "sample count must be 1, 2 or 4");

  std::memset(&P, 0, sizeof P);
  P.vs_id = st.vs_id;
  P.nattr = nout;
  for (int i = 0; i < MAX_ATTRS; ++i) P.attr_mod[i] = SRX_AM_LINEAR;
  std::memcpy(P.vs_consts, st.vs_consts, sizeof P.vs_consts);

  // stream assembler register map (stream_assembler.cpp:57-93)
  vs_binding binds[8];
  int nb = vs_bindings(st.vs_id, binds);
  uint64_t count_max = 0;
  for (int i = 0; i < nb; ++i) {
    const layout_elem* le = nullptr;
    for (auto& e : st.layout->elems)
      if (e.sem == binds[i].sem && e.index == binds[i].index) le = &e;
    if (!le) return fail(SRX_INVALID_PARAMETER, "input layout lacks a semantic the vertex shader reads");
    if (le->slot >= SRX_MAX_INPUT_SLOTS || !st.vbs[le->slot].buf) return fail(SRX_INVALID_PARAMETER, "vertex buffer slot unbound");
    const vb_slot& vb = st.vbs[le->slot];
    uint32_t need = fmt_bytes(le->fmt);
    if (!need) return fail(SRX_E_UNSUPPORTED, "vertex element format");
    if (vb.stride == 0) return fail(SRX_E_UNSUPPORTED, "zero vertex stride");
    size_t off = vb.offset + le->offset;
    input_elem_dev& d = P.in[P.nin++];
    d.base = vb.buf->d + std::min(off, vb.buf->size);
    d.avail = off < vb.buf->size ? vb.buf->size - off : 0;
    d.stride = static_cast<uint32_t>(vb.stride);
    d.fmt = le->fmt;
    d.reg = binds[i].reg;
    d.default_w = (std::string(binds[i].sem) == "position") ? 1.0f : 0.0f;
    uint64_t c = d.avail >= need ? (d.avail - need) / d.stride + 1 : 0;
    count_max = std::max(count_max, c);
  }
  uint64_t needed = count_max;
  if (!indexed) needed = std::min<uint64_t>(count_max, static_cast<uint64_t>(primcnt) * 3 + 2);
  if (needed > 0x7FFFFFFFull) return fail(SRX_E_UNSUPPORTED, "vertex count");
  P.nverts = static_cast<uint32_t>(needed);

  P.ib = indexed ? st.ib->d : nullptr;
  P.ib_size = indexed ? st.ib->size : 0;
  P.istride = st.index_fmt == SRX_FMT_R16_UINT ? 2 : 4;
  P.istart_bytes = start * P.istride;
  P.indexed = indexed ? 1 : 0;
  P.topo = st.topo;
  P.basevert = basevert;
  P.prim_count = primcnt;

  P.cull_mode = st.raster.cull_mode;
  P.front_ccw = st.raster.front_ccw;
  P.vp[0] = st.vp.x; P.vp[1] = st.vp.y; P.vp[2] = st.vp.w; P.vp[3] = st.vp.h; P.vp[4] = st.vp.minz; P.vp[5] = st.vp.maxz;
  P.target_w = st.target_vp.w;
  P.target_h = st.target_vp.h;
  P.samples = samples;
  static const float p1[1][2] = {{0.5f, 0.5f}};
  static const float p2[2][2] = {{0.25f, 0.25f}, {0.75f, 0.75f}};
  static const float p4[4][2] = {{0.375f, 0.125f}, {0.875f, 0.375f}, {0.125f, 0.625f}, {0.625f, 0.875f}};
  const float(*pat)[2] = samples == 1 ? p1 : samples == 2 ? p2 : p4;
  for (int s = 0; s < samples; ++s) {
    P.sp[s][0] = pat[s][0];
    P.sp[s][1] = pat[s][1];
  }
  // rasterizer.cpp:1106-1108
  P.tiles_x = static_cast<int32_t>(static_cast<size_t>(st.vp.w + 64 - 1) / 64);
  P.tiles_y = static_cast<int32_t>(static_cast<size_t>(st.vp.h + 64 - 1) / 64);
  P.gt = gpu_tile_size();
  float cw = std::min(st.target_vp.w, static_cast<float>(P.tiles_x * 64));
  float ch = std::min(st.target_vp.h, static_cast<float>(P.tiles_y * 64));
  P.gtiles_x = static_cast<int32_t>(std::ceil(std::max(cw, 0.0f) / P.gt));
  P.gtiles_y = static_cast<int32_t>(std::ceil(std::max(ch, 0.0f) / P.gt));

  surface_obj* ct = nullptr;
  for (auto& c : st.color_targets)
    if (c) {
      ct = c.get();
      break;
    }
  if (!st.color_targets.empty()) ct = st.color_targets[0].get();
  if (ct) P.color = ct->desc();
  surface_obj* ds = st.ds_target.get();
  if (ds) P.ds = ds->desc();
  const srx_depth_stencil_desc& d = st.ds;
  P.depth_enable = d.depth_enable;
  P.depth_func = d.depth_func;
  P.stencil_enable = ds && d.stencil_enable;
  P.read_depth = ds && d.depth_enable && d.depth_func != SRX_CMP_NEVER && d.depth_func != SRX_CMP_ALWAYS;
  P.write_depth = ds && d.depth_enable && d.depth_write_mask && d.depth_func != SRX_CMP_NEVER;
  P.read_mask = d.stencil_read_mask & 0xFF;
  P.write_mask = d.stencil_write_mask & 0xFF;
  P.stencil_ref = P.stencil_enable ? (static_cast<uint32_t>(st.stencil_ref) & P.read_mask) : 0;
  P.st_func[0] = d.front_face.stencil_func;
  P.st_func[1] = d.back_face.stencil_func;
  P.st_pass_op[0] = d.front_face.stencil_pass_op;
  P.st_pass_op[1] = d.back_face.stencil_pass_op;
  // early-Z: stencil off and no depth-writing pixel shader (framebuffer.cpp:424), per draw
  P.early_z = ds && !d.stencil_enable;
  P.count_stats = 0;
  for (auto& q : r->active)
    if (q->id != SRX_QUERY_PIPELINE_PROFILES) P.count_stats = 1;
  P.blend_id = st.blend_id;
  P.ps_id = st.ps_id;
  std::memcpy(P.ps_consts, st.ps_consts, sizeof P.ps_consts);

  std::memset(&SM, 0, sizeof SM);
  for (int i = 0; i < MAX_SAMPLERS; ++i) {
    sampler_obj* so = st.samplers[i].get();
    if (!so || !so->tex) continue;
    texture_obj* t = so->tex.get();
    sampler_desc_dev& sd = SM.s[i];
    sd.min_filter = so->desc.min_filter;
    sd.mag_filter = so->desc.mag_filter;
    sd.mip_filter = so->desc.mip_filter;
    sd.mip_qual = so->desc.mip_qual;
    sd.addr_u = so->desc.addr_mode_u;
    sd.addr_v = so->desc.addr_mode_v;
    sd.lod_bias = so->desc.mip_lod_bias;
    sd.max_aniso = so->desc.max_anisotropy;
    for (int k = 0; k < 4; ++k) sd.border[k] = so->desc.border_color[k];
    sd.nlevels = static_cast<int>(std::min<size_t>(t->surfs.size(), MAX_LEVELS));
    sd.max_lod = t->max_lod;
    sd.min_lod = std::min(t->min_lod, sd.nlevels - 1);
    sd.size[0] = t->size[0];
    sd.size[1] = t->size[1];
    sd.size[2] = t->size[2];
    for (int l = 0; l < sd.nlevels; ++l) sd.levels[l] = t->surfs[l]->desc();
    bool rgba8 = sd.nlevels > 0;
    for (int l = 0; l < sd.nlevels; ++l) rgba8 &= sd.levels[l].pf == SRX_PF_RGBA8 && sd.levels[l].samples == 1;
    sd.fast_aniso = rgba8 && sd.mip_filter == SRX_FILTER_ANISOTROPIC && sd.max_aniso > 1 &&
                    sd.min_filter == SRX_FILTER_LINEAR && sd.mag_filter == SRX_FILTER_LINEAR &&
                    sd.addr_u == SRX_ADDRESS_WRAP && sd.addr_v == SRX_ADDRESS_WRAP;
    P.has_sampler[i] = 1;
  }
  // Opaque draws whose samples keep or lose a colour in one depth reduction: early-Z draws, and
  // draws whose stencil never changes (KEEP on pass; fail ops never run), whose stencil test is
  // then a fixed per-sample mask.
  P.vis_stencil = P.stencil_enable && P.st_pass_op[0] == SRX_SOP_KEEP && P.st_pass_op[1] == SRX_SOP_KEEP &&
                  P.st_func[0] == P.st_func[1];
  return SRX_OK;
}

int32_t record_draw(renderer_obj* r, const draw_params& P, const ps_samplers& SM);

int32_t execute_draw(renderer_obj* r, bool indexed, uint32_t start, uint32_t primcnt, int32_t basevert) {
  draw_params P;
  ps_samplers SM;
  bool skip = false;
  RCHK(build_draw(r, indexed, start, primcnt, basevert, P, SM, skip));
  if (skip) return SRX_OK;
  if (r->recording) return record_draw(r, P, SM);
  const int nout = P.nattr;
  const int samples = P.samples;
  hipStream_t s = r->stream;
  uint32_t P3 = 3 * primcnt;
  uint64_t& hs_vs = r->host_stats[ST_VS_INVOCATIONS];
  r->host_stats[ST_IA_VERTICES] += static_cast<uint64_t>(primcnt) * 3;
  r->host_stats[ST_IA_PRIMITIVES] += primcnt;
  r->host_stats[ST_CINVOCATIONS] += primcnt;
  hs_vs += P.nverts;
  if (primcnt == 0) return SRX_OK;

  RCHK(ensure(r, r->vbuf, (static_cast<size_t>(P.nverts) + 1) * (nout + 1)));
  RCHK(ensure(r, r->tri_rec, static_cast<size_t>(P3) * 5));
  RCHK(ensure(r, r->tri_attr, static_cast<size_t>(P3) * 3 * (nout + 1)));
  bool vis_pass = (P.early_z || P.vis_stencil) && (P.blend_id == SRX_BLEND_OFF || P.blend_id == SRX_BLEND_DISABLED);
  if (vis_pass)
    RCHK(ensure(r, r->vis, static_cast<size_t>(std::max(P.target_w, 1.0f)) * static_cast<size_t>(std::max(P.target_h, 1.0f)) * samples));
  RCHK(ensure(r, r->samplers_d, 1));
  RCHK(ensure(r, r->cnt, P3));
  RCHK(ensure(r, r->clip_flag, primcnt));
  RCHK(ensure(r, r->clip_slot, primcnt));
  RCHK(ensure(r, r->scan_tmp, scan_tmp_size(P3)));
  uint32_t T = static_cast<uint32_t>(P.gtiles_x * P.gtiles_y);
  RCHK(ensure(r, r->partial, std::max<size_t>(T, (primcnt + 255) / 256) + 1));
  RCHK(ensure(r, r->tstart, std::max<uint32_t>(T, 1)));
  RCHK(ensure(r, r->tend, std::max<uint32_t>(T, 1)));

  draw_buffers B{};
  B.vbuf = r->vbuf.p;
  B.tri_rec = r->tri_rec.p;
  B.tri_attr = r->tri_attr.p;
  B.vis = r->vis.p;
  B.samplers = r->samplers_d.p;
  B.cnt = r->cnt.p;
  B.clip_flag = r->clip_flag.p;
  B.clip_slot = r->clip_slot.p;
  B.counters = r->counters.p;
  B.stats = r->stats_d;
  B.partial = r->partial.p;
  B.scan_tmp = r->scan_tmp.p;
  B.tstart = r->tstart.p;
  B.tend = r->tend.p;

  bool prof = r->profiling();
  draw_events de{};
  if (prof)
    for (auto& e : de.ev) {
      e = r->take_event();
      if (!e) return fail(SRX_E_HIP, "hipEventCreate");
    }
  auto mark = [&](int i) {
    if (prof) (void)hipEventRecord(de.ev[i], s);
  };
  // sampler table: staged in pinned memory; the stream syncs below before the staging is reused
  *r->samplers_h = SM;
  HIPCHK(hipMemcpyAsync(r->samplers_d.p, r->samplers_h, sizeof(ps_samplers), hipMemcpyHostToDevice, s));
  launch_fill_u32(B.counters, CT_COUNT, 0u, s);
  mark(0);
  launch_vs(P, B.vbuf, s);
  mark(1);
  // clipping capacity for every primitive, so no read-back decides whether clipping runs
  RCHK(ensure(r, r->clip_verts, static_cast<size_t>(primcnt) * 5 * (nout + 1)));
  RCHK(ensure(r, r->clip_hdr, primcnt));
  B.clip_verts = r->clip_verts.p;
  B.clip_hdr = r->clip_hdr.p;
  B.clip_capacity = primcnt;
  launch_setup(P, B, s);
  launch_clip(P, B, s);  // returns at once on the device when setup flagged no primitive
  mark(2);
  launch_exclusive_scan(B.cnt, P3, B.scan_tmp, B.counters + CT_TOTAL, s);
  mark(3);
  // Records are emitted into a grow-only buffer before their count is read back, so the read-back
  // overlaps the emission; a buffer found too small is grown and the emission repeated.
  const uint32_t guess = std::max<uint32_t>(std::max<uint32_t>(r->last_records + r->last_records / 4, P3 + T), 1u << 16);
  if (r->keys0.cap < guess) {
    RCHK(ensure(r, r->keys0, guess));
    RCHK(ensure(r, r->vals0, guess));
  }
  B.keys[0] = r->keys0.p;
  B.vals[0] = r->vals0.p;
  launch_emit(P, B, static_cast<uint32_t>(std::min<size_t>(r->keys0.cap, r->vals0.cap)), s);
  uint32_t ctr[CT_COUNT];
  RCHK(read_counters(r, ctr));
  if (ctr[CT_CLIP_OVERFLOW]) return fail(SRX_OUT_OF_MEMORY, "clip buffer overflow");
  const uint32_t total = ctr[CT_TOTAL];
  r->last_records = total;
  if (total > std::min<size_t>(r->keys0.cap, r->vals0.cap)) {
    RCHK(ensure(r, r->keys0, total));
    RCHK(ensure(r, r->vals0, total));
    B.keys[0] = r->keys0.p;
    B.vals[0] = r->vals0.p;
    launch_emit(P, B, total, s);
  }
  // statistics queries need the in-order walk (ps_invocations counts early-Z passes as they happen)
  const bool counting = P.count_stats != 0;
  const bool order_free = vis_pass && !counting && vis_applicable(P);  // tile lists need no primitive order
  uint32_t nblocks = (std::max(total, T) + 4095) / 4096;
  const size_t hist_n = std::max<size_t>(static_cast<size_t>(256) * std::max<uint32_t>(nblocks, 1), bucket_hist_size(T, total));
  RCHK(ensure(r, r->keys1, std::max<uint32_t>(total, 1)));
  RCHK(ensure(r, r->vals1, std::max<uint32_t>(total, 1)));
  RCHK(ensure(r, r->hist, hist_n));
  RCHK(ensure(r, r->scan_tmp, std::max(scan_tmp_size(P3), scan_tmp_size(static_cast<uint32_t>(hist_n)))));
  B.keys[1] = r->keys1.p;
  B.vals[1] = r->vals1.p;
  B.hist = r->hist.p;
  B.scan_tmp = r->scan_tmp.p;
  int which = 1;
  if (!(order_free && launch_bucket_sort(P, B, total, s))) {
    int bits = 1;
    while ((1u << bits) < T) ++bits;
    which = launch_sort(B, total, bits, s);
    launch_tile_ranges(P, B, which, total, s);
  }
  RCHK(ensure(r, r->order, std::max<uint32_t>(T, 1)));
  launch_tile_order(P, B, r->order.p, s);
  B.order = r->order.p;
  mark(4);
  if (vis_pass && !counting) {
    uint32_t ne, nm, nsl;
    vis_table_bounds(T, total, &ne, &nm, &nsl);
    RCHK(ensure(r, r->vis_entries, ne));
    RCHK(ensure(r, r->vis_merges, nm));
    RCHK(ensure(r, r->vis_counts, 4));
    RCHK(ensure(r, r->vis_partial, static_cast<size_t>(nsl) * 256 * samples));
    B.vis_entries = r->vis_entries.p;
    B.vis_merges = r->vis_merges.p;
    B.vis_counts = r->vis_counts.p;
    B.vis_partial = r->vis_partial.p;
  }
  const bool fused = vis_pass && !counting && launch_vis(P, B, which, needs_shading(P), total, s);
  if (!fused) launch_raster(P, B, which, s);
  mark(5);
  // the in-order kernel defers shading (to shade_kernel) only for early-Z draws
  if (!fused && vis_pass && P.early_z && needs_shading(P)) launch_shade(P, B, s);
  mark(6);
  if (prof) r->pending_prof.push_back(de);
  HIPCHK(hipGetLastError());
  return SRX_OK;
}

// ---------------------------------------------------------------- batches
// Recorded clears, draws and resolves of independent render jobs, replayed as one launch per
// pipeline stage (all jobs' workgroups in one grid), captured in a HIP graph.  The draws take the
// order-independent visibility path, so every job's output equals its one-by-one render.
enum { BS_DRAWN = 1, BS_FILLED = 2, BS_RESOLVE_SRC = 4, BS_RESOLVE_DST = 8, BS_SAMPLED = 16 };

struct batch_draw {
  draw_params P;
  ps_samplers SM;
  render_state keep;  // holds the draw's buffers, targets and textures
};
struct batch_fill {
  fill_item it;
  ref<surface_obj> s;
};
struct batch_resolve {
  resolve_item it;
  ref<surface_obj> a, b;
};

// One device image of stage maps: per stage, job[] then range[] (uint2, 8-byte aligned).
struct map_image {
  std::vector<uint32_t> words;
  struct entry {
    size_t job_off, range_off;
    uint32_t blocks;
  };
  // order: jobs in launch order; blocks[j]: workgroups of job j (0: none)
  entry add(const std::vector<uint32_t>& blocks, const std::vector<uint32_t>& order) {
    std::vector<uint32_t> first(blocks.size(), 0);
    uint32_t total = 0;
    for (uint32_t j : order) {
      first[j] = total;
      total += blocks[j];
    }
    entry e;
    if (words.size() & 1) words.push_back(0);
    e.range_off = words.size();
    for (size_t j = 0; j < blocks.size(); ++j) {
      words.push_back(first[j]);
      words.push_back(blocks[j]);
    }
    e.job_off = words.size();
    for (uint32_t j : order)
      for (uint32_t b = 0; b < blocks[j]; ++b) words.push_back(j);
    e.blocks = total;
    return e;
  }
  static stage_map at(const uint32_t* base, const entry& e) {
    return stage_map{base + e.job_off, reinterpret_cast<const uint2*>(base + e.range_off)};
  }
};

struct batch_obj : object {
  renderer_obj* r = nullptr;  // retained from srx_batch_end on
  bool owns_renderer = false;
  std::vector<batch_draw> draws;
  std::vector<batch_fill> fills;
  std::vector<batch_resolve> resolves;
  std::vector<std::pair<surface_obj*, uint32_t>> marks;  // recording hazards per surface
  int32_t error = SRX_OK;  // first recording error, reported by srx_batch_end
  // tuning
  int tile = 16, bin_pt = 16, vis_threads = 256;
  std::map<int, std::vector<uint32_t>> counted;  // records per draw, by GPU tile size
  // sizes
  std::vector<uint32_t> records;  // per draw, from the counting pass
  uint32_t capacity = 0;
  uint64_t primitives = 0, device_bytes = 0;
  // device memory: `fixed` lives as long as the batch, `tuned` is rebuilt by srx_batch_configure
  std::vector<void*> fixed, tuned;
  batch_job* jobs_d = nullptr;
  std::vector<batch_job> jobs_h;
  uint32_t *cnt_all = nullptr, *status_d = nullptr, *counts_d = nullptr, *scan_total = nullptr, *status_h = nullptr;
  uint32_t cnt_n = 0;  // per-triangle count slots + 1 trailing
  uint32_t* counters_all = nullptr;
  uint32_t* hist_all = nullptr;
  uint32_t hist_n = 0;
  uint32_t* scan_tmp = nullptr;
  uint32_t *maps_fixed = nullptr, *maps_tuned = nullptr;
  fill_item* fills_d = nullptr;
  resolve_item* resolves_d = nullptr;
  map_image::entry m_vs, m_emit, m_clip_all;
  std::vector<std::pair<int, map_image::entry>> m_setup, m_clip;  // per attribute count
  map_image::entry m_fill, m_resolve, m_bucket, m_range;
  struct vis_class {
    int na, samples;
    bool shade;
    map_image::entry m;
  };
  std::vector<vis_class> m_vis;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  batch_obj() : object(K_BATCH) {}
  uint64_t tuned_bytes = 0;
  void free_tuned() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    for (void* p : tuned) (void)hipFree(p);
    tuned.clear();
    device_bytes -= tuned_bytes;
    tuned_bytes = 0;
  }
  ~batch_obj() override {
    if (r && r->stream) (void)hipStreamSynchronize(r->stream);
    free_tuned();
    for (void* p : fixed) (void)hipFree(p);
    if (status_h) (void)hipHostFree(status_h);
    if (owns_renderer) release(r);
  }
  uint32_t& mark(surface_obj* s) {
    for (auto& m : marks)
      if (m.first == s) return m.second;
    marks.emplace_back(s, 0u);
    return marks.back().second;
  }
  int32_t refuse(int32_t code, const char* msg) {
    if (error == SRX_OK) error = fail(code, msg);
    return code;
  }
};

void drop_recording(renderer_obj* r) {
  if (r->recording) release(r->recording);
  r->recording = nullptr;
}

renderer_obj::~renderer_obj() {
  drop_recording(this);
  if (stream) (void)hipStreamSynchronize(stream);

  auto fr = [](auto& a) {
    if (a.p) (void)hipFree(a.p);
  };
  fr(vbuf); fr(tri_rec); fr(tri_attr); fr(clip_verts); fr(samplers_d); fr(vis);
  if (samplers_h) (void)hipHostFree(samplers_h); fr(cnt); fr(keys0); fr(keys1); fr(vals0); fr(vals1);
  fr(tstart); fr(tend); fr(hist); fr(scan_tmp); fr(clip_hdr); fr(counters); fr(clip_flag); fr(clip_slot); fr(partial); fr(order); fr(vis_entries); fr(vis_merges); fr(vis_counts); fr(vis_partial);
  if (stats_d) (void)hipFree(stats_d);
  if (pinned) (void)hipHostFree(pinned);
  for (auto& d : pending_prof)
    for (auto e : d.ev)
      if (e) (void)hipEventDestroy(e);
  for (auto e : event_pool) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
}

template <class T>
int32_t batch_alloc(batch_obj* b, std::vector<void*>& owner, T** p, size_t n) {
  const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  void* q = nullptr;
  HIPCHK(hipMalloc(&q, bytes));
  owner.push_back(q);
  b->device_bytes += bytes;
  if (&owner == &b->tuned) b->tuned_bytes += bytes;
  *p = static_cast<T*>(q);
  return SRX_OK;
}

int32_t record_draw(renderer_obj* r, const draw_params& P, const ps_samplers& SM) {
  batch_obj* b = r->recording;
  if (P.prim_count == 0) return SRX_OK;
  const uint32_t T = static_cast<uint32_t>(P.gtiles_x * P.gtiles_y);
  if (T == 0) return SRX_OK;
  if (!r->active.empty()) return b->refuse(SRX_E_UNSUPPORTED, "queries do not observe batched draws");
  if (!(P.early_z || P.vis_stencil) || !(P.blend_id == SRX_BLEND_OFF || P.blend_id == SRX_BLEND_DISABLED) ||
      !vis_applicable(P))
    return b->refuse(SRX_E_UNSUPPORTED, "a batched draw must be opaque with an order-independent depth test");
  if (T > BATCH_MAX_TILES) return b->refuse(SRX_E_UNSUPPORTED, "batched draw target larger than 8192 16x16 tiles");
  if (3ull * P.prim_count >= 0x40000000ull) return b->refuse(SRX_E_UNSUPPORTED, "batched draw primitive count");
  surface_obj* ts[2] = {r->st.color_targets.empty() ? nullptr : r->st.color_targets[0].get(), r->st.ds_target.get()};
  for (surface_obj* t : ts) {
    if (!t) continue;
    uint32_t& m = b->mark(t);
    if (m & (BS_DRAWN | BS_RESOLVE_SRC | BS_RESOLVE_DST | BS_SAMPLED))
      return b->refuse(SRX_E_UNSUPPORTED, "batched jobs must not share or reuse targets");
    m |= BS_DRAWN;
  }
  for (int i = 0; i < MAX_SAMPLERS; ++i) {
    sampler_obj* so = r->st.samplers[i].get();
    if (!so || !so->tex) continue;
    for (auto& sf : so->tex->surfs) {
      uint32_t& m = b->mark(sf.get());
      if (m & (BS_DRAWN | BS_FILLED | BS_RESOLVE_DST))
        return b->refuse(SRX_E_UNSUPPORTED, "a batched draw samples a surface the batch writes");
      m |= BS_SAMPLED;
    }
  }
  b->draws.push_back(batch_draw{P, SM, r->st});
  b->primitives += P.prim_count;
  return SRX_OK;
}

int32_t record_fill(renderer_obj* r, surface_obj* o, float4 v) {
  batch_obj* b = r->recording;
  const int es = pf_bytes(o->pf);
  if (es == 12) return b->refuse(SRX_E_UNSUPPORTED, "batched clears of 12-byte texels");
  uint32_t& m = b->mark(o);
  if (m & (BS_DRAWN | BS_RESOLVE_SRC | BS_RESOLVE_DST | BS_SAMPLED))
    return b->refuse(SRX_E_UNSUPPORTED, "a batched clear follows a use of its surface");
  m |= BS_FILLED;
  const size_t bytes = static_cast<size_t>(o->w) * o->h * o->samples * es;
  fill_item it{o->d, bytes / 16, static_cast<uint32_t>(bytes % 16), o->pf, v};
  for (auto& f : b->fills)
    if (f.s.get() == o) {  // a later clear of the same surface replaces the earlier one
      f.it = it;
      return SRX_OK;
    }
  b->fills.push_back(batch_fill{it, ref<surface_obj>(o)});
  return SRX_OK;
}

int32_t record_resolve(renderer_obj* r, surface_obj* a, surface_obj* d) {
  batch_obj* b = r->recording;
  uint32_t& ma = b->mark(a);
  if (ma & BS_RESOLVE_DST) return b->refuse(SRX_E_UNSUPPORTED, "batched resolve chains");
  ma |= BS_RESOLVE_SRC;
  uint32_t& md = b->mark(d);
  if (md & (BS_DRAWN | BS_RESOLVE_SRC | BS_RESOLVE_DST | BS_SAMPLED))
    return b->refuse(SRX_E_UNSUPPORTED, "a batched resolve target is used elsewhere in the batch");
  md |= BS_RESOLVE_DST;
  b->resolves.push_back(batch_resolve{resolve_item{a->desc(), d->desc()}, ref<surface_obj>(a), ref<surface_obj>(d)});
  return SRX_OK;
}

// GPU tiles of a draw's viewport for tiles of side gt (as build_draw derives them for 16).
void gpu_tiles(const draw_params& P, int gt, int32_t* gx, int32_t* gy) {
  const float cw = std::min(P.target_w, static_cast<float>(P.tiles_x * 64));
  const float ch = std::min(P.target_h, static_cast<float>(P.tiles_y * 64));
  *gx = static_cast<int32_t>(std::ceil(std::max(cw, 0.0f) / gt));
  *gy = static_cast<int32_t>(std::ceil(std::max(ch, 0.0f) / gt));
}

// Sets every job to GPU tiles of side b->tile and, unless done before for that size, runs the
// counting pass (vertex shader, setup, clipping, scan of the per-triangle record counts; one
// host synchronisation) whose per-draw record counts size the binning.
int32_t batch_count(batch_obj* b) {
  hipStream_t s = b->r->stream;
  const uint32_t J = static_cast<uint32_t>(b->draws.size());
  for (uint32_t j = 0; j < J; ++j) {
    batch_job& jb = b->jobs_h[j];
    jb.P.gt = b->tile;
    gpu_tiles(jb.P, b->tile, &jb.P.gtiles_x, &jb.P.gtiles_y);
    jb.tiles = static_cast<uint32_t>(jb.P.gtiles_x * jb.P.gtiles_y);
  }
  if (b->counted.count(b->tile)) {
    b->records = b->counted[b->tile];
  } else {
    RCHK(batch_alloc(b, b->tuned, &b->scan_tmp, scan_tmp_size(b->cnt_n)));
    RCHK(batch_alloc(b, b->tuned, &b->hist_all, 1));
    HIPCHK(hipMemcpyAsync(b->jobs_d, b->jobs_h.data(), sizeof(batch_job) * J, hipMemcpyHostToDevice, s));
    const uint32_t* mf = b->maps_fixed;
    launch_batch_reset(b->counters_all, J * CT_COUNT, b->cnt_all + (b->cnt_n - 1), b->hist_all, b->status_d, s);
    launch_vs_batch(b->jobs_d, map_image::at(mf, b->m_vs), b->m_vs.blocks, s);
    for (auto& e : b->m_setup) launch_setup_batch(e.first, b->jobs_d, map_image::at(mf, e.second), e.second.blocks, s);
    for (auto& e : b->m_clip) launch_clip_batch(e.first, b->jobs_d, map_image::at(mf, e.second), e.second.blocks, s);
    launch_exclusive_scan(b->cnt_all, b->cnt_n, b->scan_tmp, b->scan_total, s);
    launch_batch_counts(b->jobs_d, J, b->counts_d, s);
    HIPCHK(hipMemcpyAsync(b->status_h, b->counts_d, 2 * sizeof(uint32_t) * J, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipGetLastError());
    for (uint32_t j = 0; j < J; ++j)
      if (b->status_h[J + j]) return fail(SRX_OUT_OF_MEMORY, "clip buffer overflow");
    b->records.assign(b->status_h, b->status_h + J);
    b->counted[b->tile] = b->records;
  }
  uint64_t total = 0;
  for (uint32_t j = 0; j < J; ++j) total += b->records[j];
  if (total >= 0xFFFFFFFFull) return fail(SRX_E_UNSUPPORTED, "batch record count");
  b->capacity = static_cast<uint32_t>(total);
  return SRX_OK;
}

// Per-draw buffers, the job table, the fixed stage maps and the counting pass (vertex shader,
// setup, clipping and the scan of the per-triangle record counts), whose per-job record counts size
// the binning buffers.
int32_t batch_prepare(batch_obj* b) {
  renderer_obj* r = b->r;
  hipStream_t s = r->stream;
  const uint32_t J = static_cast<uint32_t>(b->draws.size());
  size_t n_vbuf = 0, n_rec = 0, n_attr = 0, n_clipv = 0, n_prims = 0, n_tiles = 0;
  uint64_t n_cnt = 0;
  for (auto& d : b->draws) {
    const draw_params& P = d.P;
    const size_t regs = static_cast<size_t>(P.nattr) + 1;
    n_vbuf += (static_cast<size_t>(P.nverts) + 1) * regs;
    n_rec += 15ull * P.prim_count;
    n_attr += 9ull * P.prim_count * regs;
    n_clipv += 5ull * P.prim_count * regs;
    n_prims += P.prim_count;
    n_cnt += 3ull * P.prim_count;
    int32_t gx8, gy8;  // tile arrays sized for the smallest GPU tiles
    gpu_tiles(P, 8, &gx8, &gy8);
    n_tiles += static_cast<size_t>(gx8) * gy8;
  }
  if (n_cnt + 1 >= 0xFFFFFFFFull) return fail(SRX_E_UNSUPPORTED, "batch primitive count");
  b->cnt_n = static_cast<uint32_t>(n_cnt + 1);
  float4 *vbuf, *tri_rec, *tri_attr, *clip_verts;
  uint8_t* clip_flag;
  int32_t* clip_slot;
  uint32_t *clip_hdr, *tiles3;
  ps_samplers* samplers;
  auto& F = b->fixed;
  RCHK(batch_alloc(b, F, &vbuf, n_vbuf));
  RCHK(batch_alloc(b, F, &tri_rec, n_rec));
  RCHK(batch_alloc(b, F, &tri_attr, n_attr));
  RCHK(batch_alloc(b, F, &clip_verts, n_clipv));
  RCHK(batch_alloc(b, F, &clip_flag, n_prims));
  RCHK(batch_alloc(b, F, &clip_slot, n_prims));
  RCHK(batch_alloc(b, F, &clip_hdr, n_prims));
  RCHK(batch_alloc(b, F, &tiles3, 3 * n_tiles));
  RCHK(batch_alloc(b, F, &samplers, J));
  RCHK(batch_alloc(b, F, &b->cnt_all, b->cnt_n));
  RCHK(batch_alloc(b, F, &b->counters_all, static_cast<size_t>(J) * CT_COUNT));
  RCHK(batch_alloc(b, F, &b->counts_d, 2 * static_cast<size_t>(J)));
  RCHK(batch_alloc(b, F, &b->status_d, 4));
  RCHK(batch_alloc(b, F, &b->jobs_d, J));
  b->scan_total = b->status_d + 1;
  HIPCHK(hipHostMalloc(&b->status_h, std::max<size_t>(2 * static_cast<size_t>(J), 4) * sizeof(uint32_t), hipHostMallocDefault));

  std::vector<ps_samplers> sm(J);
  b->jobs_h.assign(J, batch_job{});
  size_t o_vbuf = 0, o_rec = 0, o_attr = 0, o_clipv = 0, o_prims = 0, o_tiles = 0, o_cnt = 0;
  std::vector<uint32_t> vs_blocks(J), setup_blocks(J), clip_blocks(J);
  for (uint32_t j = 0; j < J; ++j) {
    const draw_params& P = b->draws[j].P;
    batch_job& jb = b->jobs_h[j];
    std::memset(&jb, 0, sizeof jb);
    jb.P = P;
    sm[j] = b->draws[j].SM;
    const size_t regs = static_cast<size_t>(P.nattr) + 1;
    int32_t gx8, gy8;
    gpu_tiles(P, 8, &gx8, &gy8);
    const uint32_t T = static_cast<uint32_t>(gx8 * gy8);
    draw_buffers& B = jb.B;
    B.vbuf = vbuf + o_vbuf;
    B.tri_rec = tri_rec + o_rec;
    B.tri_attr = tri_attr + o_attr;
    B.clip_verts = clip_verts + o_clipv;
    B.clip_flag = clip_flag + o_prims;
    B.clip_slot = clip_slot + o_prims;
    B.clip_hdr = clip_hdr + o_prims;
    B.clip_capacity = P.prim_count;
    B.cnt = b->cnt_all + o_cnt;
    B.counters = b->counters_all + static_cast<size_t>(j) * CT_COUNT;
    B.stats = r->stats_d;
    B.samplers = samplers + j;
    B.tstart = tiles3 + o_tiles;
    B.tend = tiles3 + n_tiles + o_tiles;
    B.order = tiles3 + 2 * n_tiles + o_tiles;
    jb.vis_mode = vis_mode(P);
    jb.tiles = static_cast<uint32_t>(P.gtiles_x * P.gtiles_y);
    o_vbuf += (static_cast<size_t>(P.nverts) + 1) * regs;
    o_rec += 15ull * P.prim_count;
    o_attr += 9ull * P.prim_count * regs;
    o_clipv += 5ull * P.prim_count * regs;
    o_prims += P.prim_count;
    o_tiles += T;
    o_cnt += 3ull * P.prim_count;
    vs_blocks[j] = (P.nverts + 1 + 255) / 256;
    setup_blocks[j] = (P.prim_count + 255) / 256;
    clip_blocks[j] = std::min<uint32_t>((P.prim_count + 63) / 64, 16);
  }
  HIPCHK(hipMemcpyAsync(samplers, sm.data(), sizeof(ps_samplers) * J, hipMemcpyHostToDevice, s));

  // fixed stage maps: jobs in recorded order
  std::vector<uint32_t> order(J);
  for (uint32_t j = 0; j < J; ++j) order[j] = j;
  map_image img;
  b->m_vs = img.add(vs_blocks, order);
  b->m_emit = img.add(setup_blocks, order);
  for (int na = 1; na <= 3; ++na) {
    std::vector<uint32_t> sb(J, 0), cb(J, 0);
    bool any = false;
    for (uint32_t j = 0; j < J; ++j)
      if (std::min(std::max(b->draws[j].P.nattr, 1), 3) == na) {
        sb[j] = setup_blocks[j];
        cb[j] = clip_blocks[j];
        any = true;
      }
    if (!any) continue;
    b->m_setup.emplace_back(na, img.add(sb, order));
    b->m_clip.emplace_back(na, img.add(cb, order));
  }
  std::vector<uint32_t> fill_blocks(b->fills.size()), fill_order(b->fills.size());
  std::vector<fill_item> fi(b->fills.size());
  for (size_t i = 0; i < b->fills.size(); ++i) {
    fi[i] = b->fills[i].it;
    fill_blocks[i] = static_cast<uint32_t>((std::max<size_t>(fi[i].words, 1) + 255) / 256);
    fill_order[i] = static_cast<uint32_t>(i);
  }
  b->m_fill = img.add(fill_blocks, fill_order);
  std::vector<uint32_t> res_blocks(b->resolves.size()), res_order(b->resolves.size());
  std::vector<resolve_item> ri(b->resolves.size());
  for (size_t i = 0; i < b->resolves.size(); ++i) {
    ri[i] = b->resolves[i].it;
    res_blocks[i] = static_cast<uint32_t>((static_cast<size_t>(ri[i].src.w) * ri[i].src.h + 255) / 256);
    res_order[i] = static_cast<uint32_t>(i);
  }
  b->m_resolve = img.add(res_blocks, res_order);
  RCHK(batch_alloc(b, F, &b->maps_fixed, img.words.size()));
  HIPCHK(hipMemcpyAsync(b->maps_fixed, img.words.data(), img.words.size() * 4, hipMemcpyHostToDevice, s));
  RCHK(batch_alloc(b, F, &b->fills_d, fi.size()));
  RCHK(batch_alloc(b, F, &b->resolves_d, ri.size()));
  if (!fi.empty()) HIPCHK(hipMemcpyAsync(b->fills_d, fi.data(), fi.size() * sizeof(fill_item), hipMemcpyHostToDevice, s));
  if (!ri.empty()) HIPCHK(hipMemcpyAsync(b->resolves_d, ri.data(), ri.size() * sizeof(resolve_item), hipMemcpyHostToDevice, s));
  b->tile = 16;  // the recorded draws' GPU tiles
  return batch_count(b);
}

// Every stage of the batch on stream s (captured into the batch's graph).
void batch_enqueue(batch_obj* b, hipStream_t s) {
  const uint32_t J = static_cast<uint32_t>(b->draws.size());
  const uint32_t* mf = b->maps_fixed;
  const uint32_t* mt = b->maps_tuned;
  launch_batch_reset(b->counters_all, J * CT_COUNT, b->cnt_all + (b->cnt_n - 1), b->hist_all + (b->hist_n - 1),
                     b->status_d, s);
  launch_fill_batch(b->fills_d, map_image::at(mf, b->m_fill), b->m_fill.blocks, s);
  launch_vs_batch(b->jobs_d, map_image::at(mf, b->m_vs), b->m_vs.blocks, s);
  for (auto& e : b->m_setup) launch_setup_batch(e.first, b->jobs_d, map_image::at(mf, e.second), e.second.blocks, s);
  for (auto& e : b->m_clip) launch_clip_batch(e.first, b->jobs_d, map_image::at(mf, e.second), e.second.blocks, s);
  launch_exclusive_scan(b->cnt_all, b->cnt_n, b->scan_tmp, b->scan_total, s);
  launch_emit_batch(b->jobs_d, map_image::at(mf, b->m_emit), b->m_emit.blocks, b->capacity, s);
  launch_bucket_batch(b->bin_pt, 0, b->jobs_d, map_image::at(mt, b->m_bucket), b->m_bucket.blocks, b->capacity, s);
  launch_exclusive_scan(b->hist_all, b->hist_n, b->scan_tmp, nullptr, s);
  launch_bucket_batch(b->bin_pt, 1, b->jobs_d, map_image::at(mt, b->m_bucket), b->m_bucket.blocks, b->capacity, s);
  launch_range_batch(b->jobs_d, map_image::at(mt, b->m_range), b->m_range.blocks, s);
  launch_tile_order_batch(b->jobs_d, J, s);
  for (auto& c : b->m_vis)
    launch_vis_batch(c.na, c.samples, c.shade, b->vis_threads, b->tile, b->jobs_d, map_image::at(mt, c.m), c.m.blocks, s);
  launch_resolve_batch(b->resolves_d, map_image::at(mf, b->m_resolve), b->m_resolve.blocks, s);
  launch_batch_check(b->jobs_d, J, b->capacity, b->bin_pt, b->status_d, s);
}

// Binning buffers and maps for the tuning (bin_pt, vis_threads), then the graph.
int32_t batch_configure(batch_obj* b, int tile, int bin_pt, int vis_threads) {
  if (!batch_bin_items_valid(bin_pt) || !(vis_threads == 64 || vis_threads == 128 || vis_threads == 256) ||
      !(tile == 8 || tile == 16 || tile == 32))
    return fail(SRX_INVALID_PARAMETER, "batch tuning");
  for (auto& d : b->draws) {
    int32_t gx, gy;
    gpu_tiles(d.P, tile, &gx, &gy);
    if (!vis_batch_supported(d.P.samples, needs_shading(d.P), vis_threads, tile) ||
        static_cast<uint64_t>(gx) * gy > BATCH_MAX_TILES)
      return fail(SRX_E_UNSUPPORTED, "GPU tile size not available for every job of the batch");
  }
  renderer_obj* r = b->r;
  hipStream_t s = r->stream;
  HIPCHK(hipStreamSynchronize(s));
  b->free_tuned();
  b->tile = tile;
  RCHK(batch_count(b));
  b->free_tuned();
  b->bin_pt = bin_pt;
  b->vis_threads = vis_threads;
  const uint32_t J = static_cast<uint32_t>(b->draws.size());
  const uint32_t items = 256u * bin_pt;
  std::vector<uint32_t> bucket_blocks(J), range_blocks(J), order(J);
  uint64_t hist = 0;
  for (uint32_t j = 0; j < J; ++j) {
    batch_job& jb = b->jobs_h[j];
    jb.bucket_blocks = std::max<uint32_t>((b->records[j] + items - 1) / items, 1);
    bucket_blocks[j] = jb.bucket_blocks;
    range_blocks[j] = (jb.tiles + 255) / 256;
    hist += static_cast<uint64_t>(jb.tiles) * jb.bucket_blocks;
    order[j] = j;
  }
  if (hist + 1 >= 0xFFFFFFFFull) return fail(SRX_E_UNSUPPORTED, "batch histogram size");
  b->hist_n = static_cast<uint32_t>(hist + 1);
  uint32_t *keys0, *vals0, *vals1;
  RCHK(batch_alloc(b, b->tuned, &keys0, b->capacity));
  RCHK(batch_alloc(b, b->tuned, &vals0, b->capacity));
  RCHK(batch_alloc(b, b->tuned, &vals1, b->capacity));
  RCHK(batch_alloc(b, b->tuned, &b->hist_all, b->hist_n));
  RCHK(batch_alloc(b, b->tuned, &b->scan_tmp, std::max(scan_tmp_size(b->cnt_n), scan_tmp_size(b->hist_n))));
  uint64_t h = 0;
  for (uint32_t j = 0; j < J; ++j) {
    batch_job& jb = b->jobs_h[j];
    jb.B.keys[0] = keys0;
    jb.B.vals[0] = vals0;
    jb.B.keys[1] = nullptr;
    jb.B.vals[1] = vals1;
    jb.B.hist = b->hist_all + h;
    h += static_cast<uint64_t>(jb.tiles) * jb.bucket_blocks;
  }
  // visibility: one launch per kernel variant, the jobs with the most records per tile first
  std::vector<uint32_t> heavy(J);
  for (uint32_t j = 0; j < J; ++j) heavy[j] = j;
  std::stable_sort(heavy.begin(), heavy.end(), [&](uint32_t x, uint32_t y) {
    return static_cast<uint64_t>(b->records[x]) * b->jobs_h[y].tiles > static_cast<uint64_t>(b->records[y]) * b->jobs_h[x].tiles;
  });
  map_image img;
  b->m_bucket = img.add(bucket_blocks, order);
  b->m_range = img.add(range_blocks, order);
  b->m_vis.clear();
  for (uint32_t j : heavy) {
    const draw_params& P = b->draws[j].P;
    const int na = needs_shading(P) ? std::min(std::max(P.nattr, 1), 3) : 1;
    const bool shade = needs_shading(P);
    bool found = false;
    for (auto& c : b->m_vis) found |= c.na == na && c.samples == P.samples && c.shade == shade;
    if (!found) b->m_vis.push_back(batch_obj::vis_class{na, P.samples, shade, {}});
  }
  for (auto& c : b->m_vis) {
    std::vector<uint32_t> vb(J, 0);
    for (uint32_t j = 0; j < J; ++j) {
      const draw_params& P = b->draws[j].P;
      const bool shade = needs_shading(P);
      const int na = shade ? std::min(std::max(P.nattr, 1), 3) : 1;
      if (na == c.na && P.samples == c.samples && shade == c.shade) vb[j] = b->jobs_h[j].tiles;
    }
    c.m = img.add(vb, heavy);
  }
  RCHK(batch_alloc(b, b->tuned, &b->maps_tuned, img.words.size()));
  HIPCHK(hipMemcpyAsync(b->maps_tuned, img.words.data(), img.words.size() * 4, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(b->jobs_d, b->jobs_h.data(), sizeof(batch_job) * J, hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  batch_enqueue(b, s);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  b->graph = g;
  HIPCHK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return SRX_OK;
}

// Folds the recorded stage events into the running pipeline_profiles queries (ns).
int32_t collect_profiles(renderer_obj* r) {
  if (r->pending_prof.empty()) return SRX_OK;
  HIPCHK(hipStreamSynchronize(r->stream));
  uint64_t add[PROF_COUNT] = {};
  for (auto& d : r->pending_prof) {
    float ms[6];
    for (int i = 0; i < 6; ++i) HIPCHK(hipEventElapsedTime(&ms[i], d.ev[i], d.ev[i + 1]));
    auto ns = [](float m) { return static_cast<uint64_t>(static_cast<double>(m) * 1.0e6); };
    add[PROF_VTX] += ns(ms[0]);
    add[PROF_CLIP] += ns(ms[1]);
    add[PROF_COMPACT] += ns(ms[2]);
    add[PROF_DISPATCH] += ns(ms[3]);
    add[PROF_RAS] += ns(ms[4]) + ns(ms[5]);
    add[PROF_RAS_VIS] += ns(ms[4]);
    add[PROF_RAS_SHADE] += ns(ms[5]);
    add[PROF_DRAWS] += 1;
    for (auto e : d.ev) r->event_pool.push_back(e);
  }
  r->pending_prof.clear();
  for (auto& q : r->active)
    if (q->id == SRX_QUERY_PIPELINE_PROFILES)
      for (int i = 0; i < PROF_COUNT; ++i) q->prof[i] += add[i];
  return SRX_OK;
}

template <class T>
T* as(void* h, int kind) {
  auto* o = static_cast<object*>(h);
  return (o && o->kind == kind) ? static_cast<T*>(o) : nullptr;
}

// ---------------------------------------------------------------- request server
// Serving loop over the batched kernels: requests (a resident mesh drawn with the pass-through
// vertex shader, Gouraud pixel shader and depth LESS into a fresh target of the requested size and
// sample count, then resolved) are admitted by their device footprint -- geometry scratch, bins
// and targets -- into batches that a dispatcher thread launches on several streams as earlier
// batches are still running (continuous batching).  Meshes and textures stay resident in an LRU
// pool.  Every request's result equals its one-by-one render.

// First-fit allocator over one allocation (offsets 256-byte aligned, free ranges coalesced).
struct range_arena {
  uint8_t* base = nullptr;
  size_t size = 0, used = 0, peak = 0;
  std::map<size_t, size_t> free_;  // offset -> bytes
  void init(uint8_t* b, size_t n) {
    base = b;
    size = n;
    free_.clear();
    if (n) free_[0] = n;
  }
  static size_t round(size_t n) { return (std::max<size_t>(n, 1) + 255) & ~static_cast<size_t>(255); }
  size_t largest() const {
    size_t m = 0;
    for (auto& f : free_) m = std::max(m, f.second);
    return m;
  }
  bool alloc(size_t n, size_t* off) {
    n = round(n);
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= n) {
        *off = it->first;
        const size_t rest = it->second - n, at = it->first + n;
        free_.erase(it);
        if (rest) free_[at] = rest;
        used += n;
        peak = std::max(peak, used);
        return true;
      }
    return false;
  }
  void release(size_t off, size_t n) {
    n = round(n);
    used -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct srv_asset {
  uint8_t* d = nullptr;
  size_t bytes = 0;
  uint64_t stamp = 0;  // last use (LRU)
  uint32_t users = 0;  // queued or running requests
  uint32_t n_tris = 0, size = 0;
};

struct srv_job {
  uint64_t ticket = 0;
  srx_request req{};
  int64_t admitted_ns = 0, done_ns = 0;
  uint32_t est_records = 0, batch_jobs = 0;
  bool exact = false;  // est_records is a measured count
  int32_t status = SRX_OK;
  size_t stage_off = 0, stage_bytes = 0;  // host staging of the outputs
  size_t color_bytes = 0, depth_bytes = 0;
};

struct srv_batch {
  std::vector<srv_job> jobs;
  size_t dev_off = 0, dev_bytes = 0, host_off = 0, host_bytes = 0;
  uint32_t* counts_h = nullptr;  // pinned: records[J], clip overflow[J]
  uint32_t capacity = 0;
  std::vector<uint32_t> bucket_blocks;
  int stream = 0;
  hipEvent_t done = nullptr;
};

struct server_obj : object {
  renderer_obj* r = nullptr;  // retained
  srx_server_desc desc{};
  std::vector<hipStream_t> streams;
  std::vector<bool> stream_busy;
  hipStream_t upload = nullptr;
  uint8_t* dev = nullptr;
  uint8_t* host = nullptr;
  range_arena dev_arena, host_arena;
  std::mutex mu;
  std::condition_variable cv;
  std::thread worker;
  bool stop = false;
  int32_t fatal = SRX_OK;
  std::string fatal_msg;
  std::chrono::steady_clock::time_point epoch = std::chrono::steady_clock::now();
  uint64_t next_ticket = 1, clock = 0;
  std::unordered_map<uint64_t, srv_asset> meshes, textures;
  size_t pool_used = 0;
  std::deque<srv_job> queue;
  std::vector<srv_batch> running;
  std::unordered_map<uint64_t, srv_job> finished;  // awaiting srx_server_release
  std::vector<uint64_t> unreported;
  std::unordered_map<uint64_t, uint32_t> records_cache;  // (mesh, tiles) -> exact records
  srx_server_stats stats{};
  std::vector<hipEvent_t> events;
  server_obj() : object(K_SERVER) {}
  ~server_obj() override;
  int64_t now() const {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - epoch).count();
  }
};

// Records depend on the mesh and the target size (the sample count is kept apart too).
uint64_t cache_key(const srx_request& q) {
  return (q.mesh_key * 0x9E3779B97F4A7C15ull) ^ (static_cast<uint64_t>(q.width) << 36 | static_cast<uint64_t>(q.height) << 8 | q.samples);
}

// draw_params of a served request: the state build_draw derives for this pipeline (pass-through
// vertex shader over POSITION/COLOR float4 pairs, Gouraud, cull none, depth LESS with writes,
// stencil off, blend off, full-target viewport).
void serve_params(draw_params& P, const uint8_t* mesh, uint32_t n, uint32_t w, uint32_t h, uint32_t samples,
                  uint8_t* color, uint8_t* ds) {
  std::memset(&P, 0, sizeof P);
  P.vs_id = SRX_VS_PASSTHROUGH_COLOR;
  P.nattr = 1;
  for (int i = 0; i < MAX_ATTRS; ++i) P.attr_mod[i] = SRX_AM_LINEAR;
  const uint64_t bytes = 96ull * n;
  for (int i = 0; i < 2; ++i) {
    input_elem_dev& e = P.in[P.nin++];
    e.base = mesh + 16 * i;
    e.avail = bytes - 16 * i;
    e.stride = 32;
    e.fmt = SRX_FMT_R32G32B32A32_FLOAT;
    e.reg = i;
    e.default_w = i == 0 ? 1.0f : 0.0f;
  }
  P.nverts = 3 * n;
  P.topo = SRX_TOPOLOGY_TRIANGLE_LIST;
  P.prim_count = n;
  P.cull_mode = SRX_CULL_NONE;
  P.vp[2] = static_cast<float>(w);
  P.vp[3] = static_cast<float>(h);
  P.vp[5] = 1.0f;
  P.target_w = static_cast<float>(w);
  P.target_h = static_cast<float>(h);
  P.samples = static_cast<int32_t>(samples);
  static const float p1[1][2] = {{0.5f, 0.5f}};
  static const float p2[2][2] = {{0.25f, 0.25f}, {0.75f, 0.75f}};
  static const float p4[4][2] = {{0.375f, 0.125f}, {0.875f, 0.375f}, {0.125f, 0.625f}, {0.625f, 0.875f}};
  const float(*pat)[2] = samples == 1 ? p1 : samples == 2 ? p2 : p4;
  for (uint32_t s = 0; s < samples; ++s) {
    P.sp[s][0] = pat[s][0];
    P.sp[s][1] = pat[s][1];
  }
  P.tiles_x = static_cast<int32_t>((w + 63) / 64);
  P.tiles_y = static_cast<int32_t>((h + 63) / 64);
  P.gt = 16;
  P.gtiles_x = static_cast<int32_t>((w + 15) / 16);
  P.gtiles_y = static_cast<int32_t>((h + 15) / 16);
  P.color = surface_desc{color, static_cast<int32_t>(w), static_cast<int32_t>(h), static_cast<int32_t>(samples), PF_RGBA8};
  P.ds = surface_desc{ds, static_cast<int32_t>(w), static_cast<int32_t>(h), static_cast<int32_t>(samples), PF_RG32F};
  P.depth_enable = 1;
  P.depth_func = SRX_CMP_LESS;
  P.read_depth = 1;
  P.write_depth = 1;
  P.read_mask = 0xFF;
  P.write_mask = 0xFF;
  P.st_func[0] = P.st_func[1] = SRX_CMP_ALWAYS;
  P.st_pass_op[0] = P.st_pass_op[1] = SRX_SOP_KEEP;
  P.early_z = 1;
  P.blend_id = SRX_BLEND_OFF;
  P.ps_id = SRX_PS_GOURAUD;
}

// Device bytes of one job with `records` bin entries and its share of the batch-wide arrays.
struct job_layout {
  size_t color, ds, resolved, vbuf, tri_rec, tri_attr, clip_verts, clip_flag, clip_slot, clip_hdr, tiles, hist, total;
};
job_layout layout_of(const srx_request& q, uint32_t records, uint32_t bin_items) {
  const size_t px = static_cast<size_t>(q.width) * q.height, n = q.n_tris, T = ((q.width + 15) / 16) * ((q.height + 15) / 16);
  const size_t nb = std::max<size_t>((records + bin_items - 1) / bin_items, 1);
  job_layout L;
  size_t o = 0;
  auto put = [&](size_t bytes) {
    const size_t at = o;
    o += range_arena::round(bytes);
    return at;
  };
  L.color = put(px * q.samples * 4);
  L.ds = put(px * q.samples * 8);
  L.resolved = put(q.samples > 1 ? px * 4 : 0);
  L.vbuf = put((3 * n + 1) * 2 * 16);
  L.tri_rec = put(15 * n * 16);
  L.tri_attr = put(9 * n * 2 * 16);
  L.clip_verts = put(5 * n * 2 * 16);
  L.clip_flag = put(n);
  L.clip_slot = put(4 * n);
  L.clip_hdr = put(4 * n);
  L.tiles = put(12 * T);
  L.hist = put(4 * T * nb);
  L.total = o;
  return L;
}

uint32_t estimate_records(server_obj* S, const srx_request& q, bool* exact) {
  auto it = S->records_cache.find(cache_key(q));
  *exact = it != S->records_cache.end();
  if (*exact) return it->second;
  const uint64_t T = ((q.width + 15) / 16) * ((q.height + 15) / 16);
  return static_cast<uint32_t>(std::min<uint64_t>(16ull * q.n_tris + T, 0x7FFFFFFFull));
}

void unuse_assets(server_obj* S, const srx_request& q) {
  auto m = S->meshes.find(q.mesh_key);
  if (m != S->meshes.end() && m->second.users) --m->second.users;
  if (q.texture_key) {
    auto t = S->textures.find(q.texture_key);
    if (t != S->textures.end() && t->second.users) --t->second.users;
  }
}

// Frees least recently used assets nobody waits for until `need` more bytes fit the pool.
int32_t make_room(server_obj* S, size_t need) {
  while (S->pool_used + need > S->desc.pool_bytes) {
    std::unordered_map<uint64_t, srv_asset>* from = nullptr;
    uint64_t key = 0, oldest = UINT64_MAX;
    for (auto* m : {&S->meshes, &S->textures})
      for (auto& a : *m)
        if (!a.second.users && a.second.stamp < oldest) {
          oldest = a.second.stamp;
          key = a.first;
          from = m;
        }
    if (!from) return fail(SRX_OUT_OF_MEMORY, "resident pool full of assets in use");
    srv_asset& a = (*from)[key];
    HIPCHK(hipStreamSynchronize(S->upload));
    HIPCHK(hipFree(a.d));
    S->pool_used -= a.bytes;
    ++S->stats.evictions;
    from->erase(key);
  }
  return SRX_OK;
}

int32_t put_mesh(server_obj* S, uint64_t key, const void* data, uint32_t n) {
  auto it = S->meshes.find(key);
  if (it != S->meshes.end()) {
    if (it->second.n_tris != n) return fail(SRX_INVALID_PARAMETER, "mesh key reused with another triangle count");
    it->second.stamp = ++S->clock;
    return SRX_OK;
  }
  if (!data) return fail(SRX_INVALID_PARAMETER, "mesh not resident and no data given");
  const size_t bytes = 96ull * n;
  RCHK(make_room(S, bytes));
  srv_asset a;
  a.bytes = bytes;
  a.n_tris = n;
  a.stamp = ++S->clock;
  HIPCHK(hipMalloc(&a.d, std::max<size_t>(bytes, 1)));
  HIPCHK(hipMemcpyAsync(a.d, data, bytes, hipMemcpyHostToDevice, S->upload));
  HIPCHK(hipStreamSynchronize(S->upload));
  S->pool_used += bytes;
  ++S->stats.uploads;
  S->meshes[key] = a;
  return SRX_OK;
}

// RGBA8 level 0 of size^2 texels; the box-filtered chain is generated on the device.
int32_t put_texture(server_obj* S, uint64_t key, const void* data, uint32_t size) {
  auto it = S->textures.find(key);
  if (it != S->textures.end()) {
    it->second.stamp = ++S->clock;
    return SRX_OK;
  }
  if (!data || size == 0) return fail(SRX_INVALID_PARAMETER, "texture not resident and no data given");
  size_t bytes = 0;
  for (uint32_t s = size; s; s >>= 1) bytes += static_cast<size_t>(s) * s * 4;
  RCHK(make_room(S, bytes));
  srv_asset a;
  a.bytes = bytes;
  a.size = size;
  a.stamp = ++S->clock;
  HIPCHK(hipMalloc(&a.d, bytes));
  HIPCHK(hipMemcpyAsync(a.d, data, static_cast<size_t>(size) * size * 4, hipMemcpyHostToDevice, S->upload));
  size_t off = 0;
  for (uint32_t s = size; s > 1; s >>= 1) {
    const uint32_t h = s / 2;
    surface_desc src{a.d + off, static_cast<int32_t>(s), static_cast<int32_t>(s), 1, PF_RGBA8};
    off += static_cast<size_t>(s) * s * 4;
    surface_desc dst{a.d + off, static_cast<int32_t>(h), static_cast<int32_t>(h), 1, PF_RGBA8};
    launch_mip(src, dst, SRX_FILTER_LINEAR, S->upload);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(S->upload));
  S->pool_used += bytes;
  ++S->stats.uploads;
  S->textures[key] = a;
  return SRX_OK;
}

// Stage-map words one job adds to a batch: a (first, count) pair per stage and the job id of each
// of its workgroups (vertex, setup/emit, clip, binning, ranges, visibility, two clears, resolve).
uint64_t job_map_words(const srx_request& q, uint32_t est, uint32_t items) {
  const uint64_t n = q.n_tris, T = ((q.width + 15) / 16) * ((q.height + 15) / 16), px = static_cast<uint64_t>(q.width) * q.height;
  const uint64_t nb = std::max<uint64_t>((est + items - 1) / items, 1);
  const uint64_t fill_c = (std::max<uint64_t>(px * q.samples * 4 / 16, 1) + 255) / 256;
  const uint64_t fill_d = (std::max<uint64_t>(px * q.samples * 8 / 16, 1) + 255) / 256;
  return 2 * 12 + (3 * n + 1 + 255) / 256 + (n + 255) / 256 + std::min<uint64_t>((n + 63) / 64, 16) + nb +
         (T + 255) / 256 + T + fill_c + fill_d + (q.samples > 1 ? (px + 255) / 256 : 0);
}

// Device bytes a batch needs besides its jobs' own ranges.
size_t batch_shared_bytes(size_t jobs, uint64_t cnt_n, uint64_t capacity, uint64_t hist_n, uint64_t map_words) {
  const size_t scan = std::max(scan_tmp_size(static_cast<uint32_t>(cnt_n)), scan_tmp_size(static_cast<uint32_t>(hist_n)));
  return range_arena::round(4 * cnt_n) + range_arena::round(16 * jobs) + range_arena::round(8 * jobs + 16) +
         3 * range_arena::round(4 * capacity) + range_arena::round(4 * hist_n) + range_arena::round(4 * scan) +
         range_arena::round(sizeof(batch_job) * jobs) + range_arena::round(sizeof(ps_samplers) * jobs) +
         range_arena::round(4 * (map_words + 16)) + range_arena::round(sizeof(fill_item) * 2 * jobs) +
         range_arena::round(sizeof(resolve_item) * jobs);
}
// Pinned bytes of a batch: the tables it uploads, the counts it reads back, and the outputs.
size_t batch_host_bytes(size_t jobs, uint64_t map_words, size_t outputs) {
  return range_arena::round(sizeof(batch_job) * jobs) + range_arena::round(4 * (map_words + 16)) +
         range_arena::round(sizeof(fill_item) * 2 * jobs) + range_arena::round(sizeof(resolve_item) * jobs) +
         range_arena::round(8 * jobs) + outputs;
}
size_t output_bytes(const srx_request& q) {
  const size_t px = static_cast<size_t>(q.width) * q.height;
  return range_arena::round(px * 4) + (q.want_depth ? range_arena::round(px * q.samples * 8) : 0);
}

// Builds and launches one batch from the queue head on stream si (called with the lock held):
// arrived requests are admitted in order while batch slots and arena space last.
int32_t launch_served_batch(server_obj* S, int si) {
  const int64_t now = S->now();
  const uint32_t items = static_cast<uint32_t>(S->desc.bin_items);
  std::vector<srv_job> take;
  std::vector<job_layout> lays;
  size_t dev_jobs = 0, host_out = 0;
  uint64_t cnt_n = 1, cap = 0, hist_n = 1, words = 0;
  const size_t dev_room = S->dev_arena.largest(), host_room = S->host_arena.largest();
  while (!S->queue.empty() && take.size() < S->desc.max_batch_jobs) {
    srv_job& q = S->queue.front();
    if (q.req.arrival_ns > now) break;
    if (!q.exact) q.est_records = estimate_records(S, q.req, &q.exact);  // measured since submission?
    const uint32_t est = q.est_records;
    const job_layout L = layout_of(q.req, est, items);
    const uint64_t n = q.req.n_tris, T = ((q.req.width + 15) / 16) * ((q.req.height + 15) / 16);
    const uint64_t nb = std::max<uint64_t>((est + items - 1) / items, 1);
    const uint64_t w2 = words + job_map_words(q.req, est, items);
    const size_t J2 = take.size() + 1;
    const size_t dev_need = dev_jobs + L.total + batch_shared_bytes(J2, cnt_n + 3 * n, cap + est, hist_n + T * nb, w2);
    const size_t host_need = batch_host_bytes(J2, w2, host_out + output_bytes(q.req));
    if (dev_need > dev_room || host_need > host_room || cap + est > 0xFFFFFFF0ull || cnt_n + 3 * n > 0xFFFFFFF0ull ||
        hist_n + T * nb > 0xFFFFFFF0ull) {
      if (take.empty() && S->running.empty()) {  // can never fit
        q.status = fail(SRX_OUT_OF_MEMORY, "request larger than the server's arenas");
        q.done_ns = now;
        unuse_assets(S, q.req);
        S->unreported.push_back(q.ticket);
        S->finished[q.ticket] = q;
        S->queue.pop_front();
        continue;
      }
      break;
    }
    dev_jobs += L.total;
    host_out += output_bytes(q.req);
    cnt_n += 3 * n;
    cap += est;
    hist_n += T * nb;
    words = w2;
    take.push_back(q);
    lays.push_back(L);
    S->queue.pop_front();
  }
  if (take.empty()) return SRX_OK;
  // Records are laid out job after job, so a job whose estimate falls short pushes the later ones
  // out of the batch's capacity: measured jobs go first, estimated ones last.
  {
    std::vector<size_t> idx(take.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
    std::stable_partition(idx.begin(), idx.end(), [&](size_t i) { return take[i].exact; });
    std::vector<srv_job> t2;
    std::vector<job_layout> l2;
    for (size_t i : idx) {
      t2.push_back(take[i]);
      l2.push_back(lays[i]);
    }
    take.swap(t2);
    lays.swap(l2);
  }
  const uint32_t J = static_cast<uint32_t>(take.size());
  auto give_back = [&] {
    for (auto it = take.rbegin(); it != take.rend(); ++it) S->queue.push_front(*it);
  };
  srv_batch b;
  b.stream = si;
  b.capacity = static_cast<uint32_t>(cap);
  b.dev_bytes = dev_jobs + batch_shared_bytes(J, cnt_n, cap, hist_n, words);
  b.host_bytes = batch_host_bytes(J, words, host_out);
  if (!S->dev_arena.alloc(b.dev_bytes, &b.dev_off)) {
    give_back();
    return SRX_OK;
  }
  if (!S->host_arena.alloc(b.host_bytes, &b.host_off)) {
    S->dev_arena.release(b.dev_off, b.dev_bytes);
    give_back();
    return SRX_OK;
  }
  hipStream_t s = S->streams[si];
  uint8_t* D = S->dev + b.dev_off;
  size_t off = 0;
  auto take_dev = [&](size_t bytes) {
    uint8_t* p = D + off;
    off += range_arena::round(bytes);
    return p;
  };
  // per-job ranges
  std::vector<batch_job> jobs(J);
  std::vector<fill_item> fills;
  std::vector<resolve_item> resolves;
  std::vector<const uint8_t*> out_color(J), out_ds(J);
  for (uint32_t j = 0; j < J; ++j) {
    const srx_request& q = take[j].req;
    const job_layout& L = lays[j];
    uint8_t* base = take_dev(L.total);
    batch_job& jb = jobs[j];
    std::memset(&jb, 0, sizeof jb);
    serve_params(jb.P, S->meshes[q.mesh_key].d, q.n_tris, q.width, q.height, q.samples, base + L.color, base + L.ds);
    jb.tiles = static_cast<uint32_t>(jb.P.gtiles_x * jb.P.gtiles_y);
    jb.vis_mode = vis_mode(jb.P);
    jb.bucket_blocks = std::max<uint32_t>((take[j].est_records + items - 1) / items, 1);
    draw_buffers& B = jb.B;
    B.vbuf = reinterpret_cast<float4*>(base + L.vbuf);
    B.tri_rec = reinterpret_cast<float4*>(base + L.tri_rec);
    B.tri_attr = reinterpret_cast<float4*>(base + L.tri_attr);
    B.clip_verts = reinterpret_cast<float4*>(base + L.clip_verts);
    B.clip_flag = base + L.clip_flag;
    B.clip_slot = reinterpret_cast<int32_t*>(base + L.clip_slot);
    B.clip_hdr = reinterpret_cast<uint32_t*>(base + L.clip_hdr);
    B.clip_capacity = q.n_tris;
    uint32_t* t3 = reinterpret_cast<uint32_t*>(base + L.tiles);
    B.tstart = t3;
    B.tend = t3 + jb.tiles;
    B.order = t3 + 2 * jb.tiles;
    B.stats = S->r->stats_d;
    const size_t px = static_cast<size_t>(q.width) * q.height;
    uint32_t zero = 0;
    float stencil_bits;
    std::memcpy(&stencil_bits, &zero, 4);
    fills.push_back(fill_item{base + L.color, px * q.samples * 4 / 16, static_cast<uint32_t>(px * q.samples * 4 % 16), PF_RGBA8,
                              make_float4(q.clear_color[0], q.clear_color[1], q.clear_color[2], q.clear_color[3])});
    fills.push_back(fill_item{base + L.ds, px * q.samples * 8 / 16, static_cast<uint32_t>(px * q.samples * 8 % 16), PF_RG32F,
                              make_float4(1.0f, stencil_bits, 0.0f, 0.0f)});
    if (q.samples > 1)
      resolves.push_back(resolve_item{jb.P.color, surface_desc{base + L.resolved, static_cast<int32_t>(q.width),
                                                               static_cast<int32_t>(q.height), 1, PF_RGBA8}});
    out_color[j] = base + (q.samples > 1 ? L.resolved : L.color);
    out_ds[j] = base + L.ds;
  }
  // batch-wide arrays
  uint32_t* cnt_all = reinterpret_cast<uint32_t*>(take_dev(4 * cnt_n));
  uint32_t* counters_all = reinterpret_cast<uint32_t*>(take_dev(16 * J));
  uint32_t* counts_d = reinterpret_cast<uint32_t*>(take_dev(8 * J + 16));
  uint32_t* keys0 = reinterpret_cast<uint32_t*>(take_dev(4 * cap));
  uint32_t* vals0 = reinterpret_cast<uint32_t*>(take_dev(4 * cap));
  uint32_t* vals1 = reinterpret_cast<uint32_t*>(take_dev(4 * cap));
  uint32_t* hist_all = reinterpret_cast<uint32_t*>(take_dev(4 * hist_n));
  uint32_t* scan_tmp = reinterpret_cast<uint32_t*>(
      take_dev(4 * std::max(scan_tmp_size(static_cast<uint32_t>(cnt_n)), scan_tmp_size(static_cast<uint32_t>(hist_n)))));
  batch_job* jobs_d = reinterpret_cast<batch_job*>(take_dev(sizeof(batch_job) * J));
  ps_samplers* sm_d = reinterpret_cast<ps_samplers*>(take_dev(sizeof(ps_samplers) * J));
  std::vector<uint32_t> vs_b(J), setup_b(J), clip_b(J), bucket_b(J), range_b(J), order(J);
  uint64_t co = 0, ho = 0;
  for (uint32_t j = 0; j < J; ++j) {
    batch_job& jb = jobs[j];
    jb.B.cnt = cnt_all + co;
    jb.B.counters = counters_all + 4 * j;
    jb.B.samplers = sm_d + j;
    jb.B.keys[0] = keys0;
    jb.B.vals[0] = vals0;
    jb.B.vals[1] = vals1;
    jb.B.hist = hist_all + ho;
    co += 3ull * jb.P.prim_count;
    ho += static_cast<uint64_t>(jb.tiles) * jb.bucket_blocks;
    vs_b[j] = (jb.P.nverts + 1 + 255) / 256;
    setup_b[j] = (jb.P.prim_count + 255) / 256;
    clip_b[j] = std::min<uint32_t>((jb.P.prim_count + 63) / 64, 16);
    bucket_b[j] = jb.bucket_blocks;
    range_b[j] = (jb.tiles + 255) / 256;
    order[j] = j;
  }
  std::vector<uint32_t> heavy(order);
  std::stable_sort(heavy.begin(), heavy.end(), [&](uint32_t x, uint32_t y) {
    return static_cast<uint64_t>(take[x].est_records) * jobs[y].tiles > static_cast<uint64_t>(take[y].est_records) * jobs[x].tiles;
  });
  map_image img;
  const map_image::entry m_vs = img.add(vs_b, order), m_setup = img.add(setup_b, order), m_clip = img.add(clip_b, order);
  const map_image::entry m_bucket = img.add(bucket_b, order), m_range = img.add(range_b, order);
  std::vector<std::pair<int, map_image::entry>> m_vis;
  for (int smp : {1, 2, 4}) {
    std::vector<uint32_t> vb(J, 0);
    bool any = false;
    for (uint32_t j = 0; j < J; ++j)
      if (jobs[j].P.samples == smp) {
        vb[j] = jobs[j].tiles;
        any = true;
      }
    if (any) m_vis.emplace_back(smp, img.add(vb, heavy));
  }
  std::vector<uint32_t> fill_b(fills.size()), fill_o(fills.size()), res_b(resolves.size()), res_o(resolves.size());
  for (size_t i = 0; i < fills.size(); ++i) {
    fill_b[i] = static_cast<uint32_t>((std::max<size_t>(fills[i].words, 1) + 255) / 256);
    fill_o[i] = static_cast<uint32_t>(i);
  }
  for (size_t i = 0; i < resolves.size(); ++i) {
    res_b[i] = static_cast<uint32_t>((static_cast<size_t>(resolves[i].src.w) * resolves[i].src.h + 255) / 256);
    res_o[i] = static_cast<uint32_t>(i);
  }
  const map_image::entry m_fill = img.add(fill_b, fill_o), m_res = img.add(res_b, res_o);
  const size_t nwords = img.words.size();
  uint32_t* maps_d = reinterpret_cast<uint32_t*>(take_dev(4 * (words + 16)));
  fill_item* fills_d = reinterpret_cast<fill_item*>(take_dev(sizeof(fill_item) * 2 * J));
  resolve_item* res_d = reinterpret_cast<resolve_item*>(take_dev(sizeof(resolve_item) * J));
  if (off > b.dev_bytes || nwords > words + 16) {  // the reservation is an upper bound by construction
    S->dev_arena.release(b.dev_off, b.dev_bytes);
    S->host_arena.release(b.host_off, b.host_bytes);
    give_back();
    return fail(SRX_FAILED, "served batch layout exceeds its reservation");
  }
  // pinned staging: tables, read-back counts, outputs
  uint8_t* H = S->host + b.host_off;
  size_t hoff = 0;
  auto take_host = [&](size_t bytes) {
    uint8_t* p = H + hoff;
    hoff += range_arena::round(bytes);
    return p;
  };
  uint8_t* h_jobs = take_host(sizeof(batch_job) * J);
  uint8_t* h_maps = take_host(4 * (words + 16));
  uint8_t* h_fills = take_host(sizeof(fill_item) * 2 * J);
  uint8_t* h_res = take_host(sizeof(resolve_item) * J);
  b.counts_h = reinterpret_cast<uint32_t*>(take_host(8 * J));
  std::memcpy(h_jobs, jobs.data(), sizeof(batch_job) * J);
  std::memcpy(h_maps, img.words.data(), 4 * nwords);
  std::memcpy(h_fills, fills.data(), sizeof(fill_item) * fills.size());
  if (!resolves.empty()) std::memcpy(h_res, resolves.data(), sizeof(resolve_item) * resolves.size());
  HIPCHK(hipMemcpyAsync(jobs_d, h_jobs, sizeof(batch_job) * J, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(maps_d, h_maps, 4 * nwords, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(fills_d, h_fills, sizeof(fill_item) * fills.size(), hipMemcpyHostToDevice, s));
  if (!resolves.empty()) HIPCHK(hipMemcpyAsync(res_d, h_res, sizeof(resolve_item) * resolves.size(), hipMemcpyHostToDevice, s));
  // the pipeline of batch_enqueue, launched directly (every served batch differs, so no graph)
  const uint32_t* mf = maps_d;
  launch_batch_reset(counters_all, 4 * J, cnt_all + (cnt_n - 1), hist_all + (hist_n - 1), counts_d + 2 * J, s);
  launch_fill_batch(fills_d, map_image::at(mf, m_fill), m_fill.blocks, s);
  launch_vs_batch(jobs_d, map_image::at(mf, m_vs), m_vs.blocks, s);
  launch_setup_batch(1, jobs_d, map_image::at(mf, m_setup), m_setup.blocks, s);
  launch_clip_batch(1, jobs_d, map_image::at(mf, m_clip), m_clip.blocks, s);
  launch_exclusive_scan(cnt_all, static_cast<uint32_t>(cnt_n), scan_tmp, nullptr, s);
  launch_emit_batch(jobs_d, map_image::at(mf, m_setup), m_setup.blocks, b.capacity, s);
  const int pt = static_cast<int>(items / 256);
  launch_bucket_batch(pt, 0, jobs_d, map_image::at(mf, m_bucket), m_bucket.blocks, b.capacity, s);
  launch_exclusive_scan(hist_all, static_cast<uint32_t>(hist_n), scan_tmp, nullptr, s);
  launch_bucket_batch(pt, 1, jobs_d, map_image::at(mf, m_bucket), m_bucket.blocks, b.capacity, s);
  launch_range_batch(jobs_d, map_image::at(mf, m_range), m_range.blocks, s);
  launch_tile_order_batch(jobs_d, J, s);
  for (auto& v : m_vis)
    launch_vis_batch(1, v.first, true, S->desc.vis_threads, 16, jobs_d, map_image::at(mf, v.second), v.second.blocks, s);
  launch_resolve_batch(res_d, map_image::at(mf, m_res), m_res.blocks, s);
  launch_batch_counts(jobs_d, J, counts_d, s);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(b.counts_h, counts_d, 8 * J, hipMemcpyDeviceToHost, s));
  for (uint32_t j = 0; j < J; ++j) {
    srv_job& t = take[j];
    const srx_request& q = t.req;
    const size_t px = static_cast<size_t>(q.width) * q.height;
    t.color_bytes = px * 4;
    t.depth_bytes = q.want_depth ? px * q.samples * 8 : 0;
    t.stage_off = b.host_off + hoff;
    HIPCHK(hipMemcpyAsync(take_host(t.color_bytes), out_color[j], t.color_bytes, hipMemcpyDeviceToHost, s));
    if (t.depth_bytes) HIPCHK(hipMemcpyAsync(take_host(t.depth_bytes), out_ds[j], t.depth_bytes, hipMemcpyDeviceToHost, s));
    t.stage_bytes = b.host_off + hoff - t.stage_off;
    t.admitted_ns = now;
    t.batch_jobs = J;
  }
  if (S->events.empty()) {
    hipEvent_t e;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    S->events.push_back(e);
  }
  b.done = S->events.back();
  S->events.pop_back();
  HIPCHK(hipEventRecord(b.done, s));
  b.bucket_blocks = bucket_b;
  b.jobs = std::move(take);
  S->stream_busy[si] = true;
  S->running.push_back(std::move(b));
  ++S->stats.batches;
  S->stats.max_batch_jobs = std::max<uint64_t>(S->stats.max_batch_jobs, J);
  return SRX_OK;
}

// A finished batch: results of complete jobs are reported, jobs whose records outgrew their
// estimate go back to the head of the queue with the exact count.
void finish_batch(server_obj* S, srv_batch& b) {
  const uint32_t J = static_cast<uint32_t>(b.jobs.size());
  const uint32_t items = static_cast<uint32_t>(S->desc.bin_items);
  const int64_t now = S->now();
  uint64_t end = 0;
  std::vector<srv_job> retry;
  for (uint32_t j = 0; j < J; ++j) {
    srv_job& t = b.jobs[j];
    const uint32_t rec = b.counts_h[j];
    end += rec;
    const bool ok = end <= b.capacity && rec <= b.bucket_blocks[j] * items && !b.counts_h[J + j];
    S->records_cache[cache_key(t.req)] = rec;
    if (!ok && !b.counts_h[J + j]) {
      S->host_arena.release(t.stage_off, t.stage_bytes);
      t.est_records = rec;
      t.exact = true;
      ++S->stats.retries;
      retry.push_back(t);
      continue;
    }
    t.status = b.counts_h[J + j] ? SRX_OUT_OF_MEMORY : SRX_OK;
    t.done_ns = now;
    unuse_assets(S, t.req);
    ++S->stats.jobs;
    S->unreported.push_back(t.ticket);
    S->finished[t.ticket] = t;
  }
  for (auto it = retry.rbegin(); it != retry.rend(); ++it) S->queue.push_front(*it);
  // the tables and counts precede the outputs in the host region; outputs stay until released
  const size_t outs_begin = J ? b.jobs[0].stage_off - b.host_off : 0;
  S->host_arena.release(b.host_off, outs_begin);
  S->dev_arena.release(b.dev_off, b.dev_bytes);
  S->events.push_back(b.done);
  S->stream_busy[b.stream] = false;
}

void serve_loop(server_obj* S) {
  std::unique_lock<std::mutex> lk(S->mu);
  while (!S->stop) {
    bool progress = false;
    for (size_t i = 0; i < S->running.size();) {
      hipError_t e = hipEventQuery(S->running[i].done);
      if (e == hipSuccess) {
        finish_batch(S, S->running[i]);
        S->running.erase(S->running.begin() + static_cast<long>(i));
        progress = true;
      } else if (e == hipErrorNotReady) {
        ++i;
      } else {
        S->fatal = hip_fail(e, "served batch");
        S->fatal_msg = g_last_error_str;
        S->stop = true;
        break;
      }
    }
    if (progress) S->cv.notify_all();
    for (size_t si = 0; si < S->streams.size() && !S->stop; ++si) {
      if (S->stream_busy[si] || S->queue.empty() || S->queue.front().req.arrival_ns > S->now()) continue;
      const size_t before = S->queue.size();
      int32_t rc = launch_served_batch(S, static_cast<int>(si));
      if (rc != SRX_OK) {
        S->fatal = rc;
        S->fatal_msg = g_last_error_str;
        S->stop = true;
      }
      if (S->queue.size() != before) progress = true;
      if (!S->stream_busy[si]) break;  // nothing admitted: wait for space
    }
    if (!progress) {
      // sleep until a batch may have finished or the next request arrives
      int64_t wait_ns = S->running.empty() ? 2000000 : 20000;
      if (!S->queue.empty() && S->running.size() < S->streams.size())
        wait_ns = std::min<int64_t>(wait_ns, std::max<int64_t>(S->queue.front().req.arrival_ns - S->now(), 0));
      S->cv.wait_for(lk, std::chrono::nanoseconds(std::max<int64_t>(wait_ns, 5000)));
    }
  }
  S->cv.notify_all();
}

server_obj::~server_obj() {
  {
    std::lock_guard<std::mutex> lk(mu);
    stop = true;
  }
  cv.notify_all();
  if (worker.joinable()) worker.join();
  for (auto s : streams) (void)hipStreamSynchronize(s);
  if (upload) (void)hipStreamSynchronize(upload);
  for (auto* m : {&meshes, &textures})
    for (auto& a : *m) (void)hipFree(a.second.d);
  for (auto e : events) (void)hipEventDestroy(e);
  for (auto& b : running)
    if (b.done) (void)hipEventDestroy(b.done);
  if (dev) (void)hipFree(dev);
  if (host) (void)hipHostFree(host);
  for (auto s : streams) (void)hipStreamDestroy(s);
  if (upload) (void)hipStreamDestroy(upload);
  release(r);
}

}  // namespace

// ==================================================================== C ABI
extern "C" {

uint32_t srx_abi_version(void) { return SRX_ABI_VERSION; }
const char* srx_backend_name(void) { return "hip"; }
int32_t srx_last_error(void) { return g_last_error; }
const char* srx_last_error_string(void) { return g_last_error_str.c_str(); }
void srx_retain(void* h) { retain(static_cast<object*>(h)); }
void srx_release(void* h) { release(static_cast<object*>(h)); }

void srx_default_raster_desc(srx_raster_desc* d) {
  *d = srx_raster_desc{SRX_FILL_SOLID, SRX_CULL_BACK, 0, 0, 0.0f, 0.0f, 1, 0, 1, 0};
}
void srx_default_depth_stencil_desc(srx_depth_stencil_desc* d) {
  srx_depth_stencil_op_desc op{SRX_SOP_KEEP, SRX_SOP_KEEP, SRX_SOP_KEEP, SRX_CMP_ALWAYS};
  *d = srx_depth_stencil_desc{1, 1, SRX_CMP_LESS, 0, 0xFF, 0xFF, op, op};
}
void srx_default_sampler_desc(srx_sampler_desc* d) {
  *d = srx_sampler_desc{SRX_FILTER_POINT, SRX_FILTER_POINT, SRX_FILTER_POINT, SRX_MIP_MI, SRX_ADDRESS_WRAP,
                        SRX_ADDRESS_WRAP, SRX_ADDRESS_WRAP, 0.0f, 0, SRX_CMP_ALWAYS, {0, 0, 0, 0}, -1e20f, 1e20f};
}

int32_t srx_create_renderer(uint32_t kind, srx_renderer* out) {
  (void)kind;
  if (!out) return SRX_INVALID_PARAMETER;
  std::call_once(g_ewa_once, [] {
    float t[256];
    make_ewa(t);
    hipError_t e = upload_ewa_table(t, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) g_ewa_status = hip_fail(e, "EWA table upload");
  });
  if (g_ewa_status != SRX_OK) return g_ewa_status;
  auto* r = new renderer_obj();
  hipError_t e = hipStreamCreateWithFlags(&r->stream, hipStreamDefault);
  if (e == hipSuccess) e = hipMalloc(&r->counters.p, CT_COUNT * sizeof(uint32_t));
  if (e == hipSuccess) r->counters.cap = CT_COUNT;
  if (e == hipSuccess) e = hipMalloc(&r->stats_d, ST_COUNT * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMemset(r->stats_d, 0, ST_COUNT * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipHostMalloc(&r->pinned, 64, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc(&r->samplers_h, sizeof(ps_samplers), hipHostMallocDefault);
  if (e != hipSuccess) {
    delete r;
    return hip_fail(e, "srx_create_renderer");
  }
  *out = reinterpret_cast<srx_renderer>(static_cast<object*>(r));
  return SRX_OK;
}

int32_t srx_create_buffer(srx_renderer r, size_t size, srx_buffer* out) {
  if (!as<renderer_obj>(r, K_RENDERER) || !out) return SRX_INVALID_PARAMETER;
  auto* b = new buffer_obj();
  b->size = size;
  hipError_t e = hipMalloc(&b->d, size ? size : 1);
  if (e == hipSuccess) e = hipMemset(b->d, 0, size ? size : 1);
  if (e != hipSuccess) {
    delete b;
    return hip_fail(e, "srx_create_buffer");
  }
  *out = reinterpret_cast<srx_buffer>(static_cast<object*>(b));
  return SRX_OK;
}
size_t srx_buffer_size(srx_buffer b) {
  auto* o = as<buffer_obj>(b, K_BUFFER);
  return o ? o->size : 0;
}
// Host copies go through the null stream, which waits for every renderer stream.
int32_t srx_buffer_write(srx_buffer b, size_t off, const void* data, size_t size) {
  auto* o = as<buffer_obj>(b, K_BUFFER);
  if (!o || off + size > o->size || (!data && size)) return SRX_INVALID_PARAMETER;
  if (size) HIPCHK(hipMemcpy(o->d + off, data, size, hipMemcpyHostToDevice));
  return SRX_OK;
}
int32_t srx_buffer_read(srx_buffer b, size_t off, void* data, size_t size) {
  auto* o = as<buffer_obj>(b, K_BUFFER);
  if (!o || off + size > o->size || (!data && size)) return SRX_INVALID_PARAMETER;
  if (size) HIPCHK(hipMemcpy(data, o->d + off, size, hipMemcpyDeviceToHost));
  return SRX_OK;
}

int32_t srx_create_tex2d(srx_renderer r, size_t w, size_t h, size_t ns, int32_t fmt, srx_texture* out) {
  if (!as<renderer_obj>(r, K_RENDERER) || !out || w == 0 || h == 0 || ns == 0 || fmt < 0 || fmt > 6 ||
      w > SRX_MAX_RENDER_TARGET_SIZE * 2 || h > SRX_MAX_RENDER_TARGET_SIZE * 2 || ns > SRX_MAX_SAMPLE_COUNT)
    return SRX_INVALID_PARAMETER;
  auto* t = new texture_obj();
  t->pf = fmt;
  t->samples = static_cast<int>(ns);
  t->size[0] = static_cast<uint32_t>(w);
  t->size[1] = static_cast<uint32_t>(h);
  t->size[2] = 1;
  surface_obj* s = nullptr;
  int32_t rc = make_surface(static_cast<int>(w), static_cast<int>(h), static_cast<int>(ns), fmt, &s);
  if (rc != SRX_OK) {
    delete t;
    return rc;
  }
  t->surfs.emplace_back(s);
  release(s);
  *out = reinterpret_cast<srx_texture>(static_cast<object*>(t));
  return SRX_OK;
}
int32_t srx_texture_surface(srx_texture t, size_t level, srx_surface* out) {
  auto* o = as<texture_obj>(t, K_TEXTURE);
  if (!o || !out || level >= o->surfs.size()) return SRX_INVALID_PARAMETER;
  if (static_cast<int>(level) < o->max_lod || static_cast<int>(level) > o->min_lod) return SRX_INVALID_PARAMETER;
  retain(o->surfs[level].get());
  *out = reinterpret_cast<srx_surface>(static_cast<object*>(o->surfs[level].get()));
  return SRX_OK;
}
// texture_2d::gen_mipmap (texture2d.cpp:25-36) with the box filter on the device
int32_t srx_texture_gen_mipmap(srx_texture t, int32_t filter, int32_t auto_gen) {
  auto* o = as<texture_obj>(t, K_TEXTURE);
  if (!o || (filter != SRX_FILTER_POINT && filter != SRX_FILTER_LINEAR)) return SRX_INVALID_PARAMETER;
  if (auto_gen) {
    uint32_t m = std::max(o->size[0], std::max(o->size[1], o->size[2]));
    int n = 0;
    while (m > 0) {
      m >>= 1;
      ++n;
    }
    o->max_lod = 0;
    o->min_lod = n - 1;
  }
  o->surfs.resize(1);
  for (int l = o->max_lod; l < o->min_lod; ++l) {
    surface_obj* src = o->surfs.back().get();
    surface_obj* dst = nullptr;
    RCHK(make_surface((src->w + 1) / 2, (src->h + 1) / 2, src->samples, src->pf, &dst));
    launch_mip(src->desc(), dst->desc(), filter, nullptr);
    o->surfs.emplace_back(dst);
    release(dst);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return SRX_OK;
}
size_t srx_texture_max_lod(srx_texture t) {
  auto* o = as<texture_obj>(t, K_TEXTURE);
  return o ? o->max_lod : 0;
}
size_t srx_texture_min_lod(srx_texture t) {
  auto* o = as<texture_obj>(t, K_TEXTURE);
  return o ? o->min_lod : 0;
}
size_t srx_surface_width(srx_surface s) { auto* o = as<surface_obj>(s, K_SURFACE); return o ? o->w : 0; }
size_t srx_surface_height(srx_surface s) { auto* o = as<surface_obj>(s, K_SURFACE); return o ? o->h : 0; }
size_t srx_surface_sample_count(srx_surface s) { auto* o = as<surface_obj>(s, K_SURFACE); return o ? o->samples : 0; }
int32_t srx_surface_format(srx_surface s) { auto* o = as<surface_obj>(s, K_SURFACE); return o ? o->pf : -1; }
size_t srx_surface_size_bytes(srx_surface s) { auto* o = as<surface_obj>(s, K_SURFACE); return o ? o->bytes : 0; }
int32_t srx_surface_write(srx_surface s, const void* d, size_t n) {
  auto* o = as<surface_obj>(s, K_SURFACE);
  if (!o || n != o->bytes || !d) return SRX_INVALID_PARAMETER;
  HIPCHK(hipMemcpy(o->d, d, n, hipMemcpyHostToDevice));
  return SRX_OK;
}
int32_t srx_surface_read(srx_surface s, void* d, size_t n) {
  auto* o = as<surface_obj>(s, K_SURFACE);
  if (!o || n != o->bytes || !d) return SRX_INVALID_PARAMETER;
  HIPCHK(hipMemcpy(d, o->d, n, hipMemcpyDeviceToHost));
  return SRX_OK;
}
int32_t srx_surface_resolve(srx_renderer rr, srx_surface src, srx_surface dst) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* a = as<surface_obj>(src, K_SURFACE);
  auto* b = as<surface_obj>(dst, K_SURFACE);
  if (!r || !a || !b || b->samples != 1 || a->w != b->w || a->h != b->h) return SRX_INVALID_PARAMETER;
  if (r->recording) return record_resolve(r, a, b);
  launch_resolve(a->desc(), b->desc(), r->stream);
  HIPCHK(hipGetLastError());
  return SRX_OK;
}

int32_t srx_create_sampler(srx_renderer r, const srx_sampler_desc* d, srx_texture t, srx_sampler* out) {
  auto* tex = as<texture_obj>(t, K_TEXTURE);
  if (!as<renderer_obj>(r, K_RENDERER) || !d || !tex || !out) return SRX_INVALID_PARAMETER;
  auto* s = new sampler_obj();
  s->desc = *d;
  s->tex = ref<texture_obj>(tex);
  *out = reinterpret_cast<srx_sampler>(static_cast<object*>(s));
  return SRX_OK;
}
int32_t srx_create_query(srx_renderer r, uint32_t id, srx_query* out) {
  if (!as<renderer_obj>(r, K_RENDERER) || !out) return SRX_INVALID_PARAMETER;
  if (id != SRX_QUERY_PIPELINE_STATISTICS && id != SRX_QUERY_INTERNAL_STATISTICS && id != SRX_QUERY_PIPELINE_PROFILES)
    return SRX_INVALID_PARAMETER;
  auto* q = new query_obj();
  q->id = id;
  *out = reinterpret_cast<srx_query>(static_cast<object*>(q));
  return SRX_OK;
}
int32_t srx_create_input_layout(srx_renderer r, const srx_input_element_desc* d, size_t n, int32_t vs_id,
                                srx_input_layout* out) {
  (void)vs_id;
  if (!as<renderer_obj>(r, K_RENDERER) || (!d && n) || !out) return SRX_INVALID_PARAMETER;
  auto* l = new layout_obj();
  for (size_t i = 0; i < n; ++i)
    l->elems.push_back(layout_elem{canon_sem(lower(d[i].semantic_name)), d[i].semantic_index, d[i].data_format,
                                   d[i].input_slot, d[i].aligned_byte_offset});
  *out = reinterpret_cast<srx_input_layout>(static_cast<object*>(l));
  return SRX_OK;
}
int32_t srx_create_raster_state(srx_renderer r, const srx_raster_desc* d, srx_raster_state* out) {
  if (!as<renderer_obj>(r, K_RENDERER) || !d || !out) return SRX_INVALID_PARAMETER;
  if (d->cull_mode < 0 || d->cull_mode > 2) return SRX_INVALID_PARAMETER;
  auto* o = new raster_obj();
  o->desc = *d;
  *out = reinterpret_cast<srx_raster_state>(static_cast<object*>(o));
  return SRX_OK;
}
int32_t srx_create_depth_stencil_state(srx_renderer r, const srx_depth_stencil_desc* d, srx_ds_state* out) {
  if (!as<renderer_obj>(r, K_RENDERER) || !d || !out) return SRX_INVALID_PARAMETER;
  auto* o = new ds_obj();
  o->desc = *d;
  *out = reinterpret_cast<srx_ds_state>(static_cast<object*>(o));
  return SRX_OK;
}

int32_t srx_set_vertex_buffers(srx_renderer rr, size_t start, size_t n, const srx_buffer* bufs, const size_t* strides,
                               const size_t* offsets) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || start + n > SRX_MAX_INPUT_SLOTS || (n && (!bufs || !strides))) return SRX_INVALID_PARAMETER;
  for (size_t i = 0; i < n; ++i) {
    r->st.vbs[start + i].buf = ref<buffer_obj>(as<buffer_obj>(bufs[i], K_BUFFER));
    r->st.vbs[start + i].stride = strides[i];
    r->st.vbs[start + i].offset = offsets ? offsets[i] : 0;
  }
  return SRX_OK;
}
int32_t srx_set_index_buffer(srx_renderer rr, srx_buffer b, uint32_t fmt) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  if (fmt != SRX_FMT_R16_UINT && fmt != SRX_FMT_R32_UINT) return SRX_FAILED;  // renderer_impl.cpp:49-53
  r->st.ib = ref<buffer_obj>(as<buffer_obj>(b, K_BUFFER));
  r->st.index_fmt = fmt;
  return SRX_OK;
}
int32_t srx_set_input_layout(srx_renderer rr, srx_input_layout l) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  r->st.layout = ref<layout_obj>(as<layout_obj>(l, K_LAYOUT));
  return SRX_OK;
}
int32_t srx_set_primitive_topology(srx_renderer rr, uint32_t t) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  switch (t) {  // renderer_impl.cpp:70-82
  case SRX_TOPOLOGY_LINE_LIST:
  case SRX_TOPOLOGY_LINE_STRIP:
  case SRX_TOPOLOGY_TRIANGLE_LIST:
  case SRX_TOPOLOGY_TRIANGLE_STRIP: r->st.topo = t; return SRX_OK;
  }
  return SRX_FAILED;
}
int32_t srx_set_vertex_shader(srx_renderer rr, int32_t id, const void* consts, size_t size) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || vs_output_count(id) <= 0) return SRX_INVALID_PARAMETER;
  if ((id == SRX_VS_WVP_NORMAL_UV || id == SRX_VS_WVP_COLOR) && size < 64) return SRX_INVALID_PARAMETER;
  if (id == SRX_VS_WVP_COLOR_LIGHT && size < 128) return SRX_INVALID_PARAMETER;
  r->st.vs_id = id;
  std::memset(r->st.vs_consts, 0, sizeof r->st.vs_consts);
  if (consts && size) std::memcpy(r->st.vs_consts, consts, std::min<size_t>(size, sizeof r->st.vs_consts));
  return SRX_OK;
}
int32_t srx_set_pixel_shader(srx_renderer rr, int32_t id, const void* consts, size_t size) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || id < SRX_PS_GOURAUD || id > SRX_PS_SHADOW_PCF) return SRX_INVALID_PARAMETER;
  if ((id == SRX_PS_LAMBERT_TEX || id == SRX_PS_CONSTANT || id == SRX_PS_SHADOW_PCF) && size < 16) return SRX_INVALID_PARAMETER;
  r->st.ps_id = id;
  std::memset(r->st.ps_consts, 0, sizeof r->st.ps_consts);
  if (consts && size) std::memcpy(r->st.ps_consts, consts, std::min<size_t>(size, 64));
  return SRX_OK;
}
int32_t srx_set_ps_sampler(srx_renderer rr, uint32_t slot, srx_sampler s) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || slot >= SRX_MAX_SAMPLER_SLOTS) return SRX_INVALID_PARAMETER;
  r->st.samplers[slot] = ref<sampler_obj>(as<sampler_obj>(s, K_SAMPLER));
  return SRX_OK;
}
int32_t srx_set_blend_shader(srx_renderer rr, int32_t id) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || id < 0 || id > 2) return SRX_INVALID_PARAMETER;
  r->st.blend_id = id;
  return SRX_OK;
}
int32_t srx_set_rasterizer_state(srx_renderer rr, srx_raster_state rs) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* o = as<raster_obj>(rs, K_RASTER);
  if (!r || !o) return SRX_INVALID_PARAMETER;
  r->st.raster = o->desc;
  return SRX_OK;
}
int32_t srx_set_depth_stencil_state(srx_renderer rr, srx_ds_state d, int32_t ref_) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* o = as<ds_obj>(d, K_DS);
  if (!r || !o) return SRX_INVALID_PARAMETER;
  r->st.ds = o->desc;
  r->st.stencil_ref = ref_;
  return SRX_OK;
}
// renderer_impl::set_render_targets (renderer_impl.cpp:159-238)
int32_t srx_set_render_targets(srx_renderer rr, size_t n, const srx_surface* cts, srx_surface dst) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  render_state& st = r->st;
  st.target_vp = srx_viewport{0, 0, 0, 0, 0, 0};
  st.target_samples = 0;
  st.color_targets.clear();
  st.ds_target = ref<surface_obj>();
  if (n >= 8) return SRX_FAILED;
  auto* ds = as<surface_obj>(dst, K_SURFACE);
  if ((n == 0 || !cts) && !ds) return SRX_FAILED;
  srx_viewport tvp{0, 0, 3.4028234663852886e38f, 3.4028234663852886e38f, 0, 0};
  size_t samples = 0;
  for (size_t i = 0; i < n; ++i) {
    auto* c = as<surface_obj>(cts[i], K_SURFACE);
    if (!c) continue;
    tvp.w = std::min(static_cast<float>(c->w), tvp.w);
    tvp.h = std::min(static_cast<float>(c->h), tvp.h);
    if (samples == 0) samples = c->samples;
    else if (static_cast<size_t>(c->samples) != samples) return SRX_FAILED;
  }
  if (ds) {
    if (ds->pf != PF_RG32F) return SRX_FAILED;
    if (n == 0) {
      samples = ds->samples;
      tvp.w = static_cast<float>(ds->w);
      tvp.h = static_cast<float>(ds->h);
    }
    if (static_cast<float>(ds->w) < tvp.w || static_cast<float>(ds->h) < tvp.h) return SRX_FAILED;
    if (static_cast<size_t>(ds->samples) != samples) return SRX_FAILED;
  }
  st.target_vp = tvp;
  st.target_samples = samples;
  for (size_t i = 0; i < n; ++i) st.color_targets.emplace_back(as<surface_obj>(cts[i], K_SURFACE));
  st.ds_target = ref<surface_obj>(ds);
  return SRX_OK;
}
int32_t srx_set_viewport(srx_renderer rr, const srx_viewport* vp) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !vp) return SRX_INVALID_PARAMETER;
  if (vp->x < 0 || vp->y < 0 || vp->w >= SRX_MAX_RENDER_TARGET_SIZE || vp->h >= SRX_MAX_RENDER_TARGET_SIZE) return SRX_FAILED;
  r->st.vp = *vp;
  return SRX_OK;
}
int32_t srx_get_viewport(srx_renderer rr, srx_viewport* vp) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !vp) return SRX_INVALID_PARAMETER;
  *vp = r->st.vp;
  return SRX_OK;
}

int32_t srx_begin(srx_renderer rr, srx_query qq) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* q = as<query_obj>(qq, K_QUERY);
  if (!r || !q) return SRX_INVALID_PARAMETER;
  if (q->running) return SRX_FAILED;
  for (auto& a : r->active)
    if (a->id == q->id) return SRX_FAILED;
  RCHK(collect_profiles(r));
  q->running = true;
  q->has_data = false;
  snapshot_stats(r, q->begin_snap);
  for (auto& v : q->prof) v = 0;
  r->active.emplace_back(q);
  return SRX_OK;
}
int32_t srx_end(srx_renderer rr, srx_query qq) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* q = as<query_obj>(qq, K_QUERY);
  if (!r || !q || !q->running) return SRX_FAILED;
  RCHK(collect_profiles(r));
  uint64_t now[ST_COUNT];
  snapshot_stats(r, now);
  uint64_t d[ST_COUNT];
  for (int i = 0; i < ST_COUNT; ++i) d[i] = now[i] - q->begin_snap[i];
  q->stats = srx_pipeline_statistics{};
  q->stats.ia_vertices = d[ST_IA_VERTICES];
  q->stats.ia_primitives = d[ST_IA_PRIMITIVES];
  q->stats.vs_invocations = d[ST_VS_INVOCATIONS];
  q->stats.cinvocations = d[ST_CINVOCATIONS];
  q->stats.cprimitives = d[ST_CPRIMITIVES];
  q->stats.ps_invocations = d[ST_PS_INVOCATIONS];
  q->istats.backend_input_pixels = d[ST_BACKEND_PIXELS];
  q->running = false;
  q->has_data = true;
  for (size_t i = 0; i < r->active.size(); ++i)
    if (r->active[i].get() == q) {
      r->active.erase(r->active.begin() + static_cast<long>(i));
      break;
    }
  return SRX_OK;
}
int32_t srx_get_data(srx_renderer rr, srx_query qq, void* data, size_t size, int32_t do_not_wait) {
  (void)do_not_wait;
  auto* q = as<query_obj>(qq, K_QUERY);
  if (!as<renderer_obj>(rr, K_RENDERER) || !q || !data || !q->has_data) return SRX_ASYNC_ERROR;
  if (q->id == SRX_QUERY_PIPELINE_STATISTICS && size >= sizeof(q->stats)) std::memcpy(data, &q->stats, sizeof(q->stats));
  else if (q->id == SRX_QUERY_INTERNAL_STATISTICS && size >= sizeof(q->istats)) std::memcpy(data, &q->istats, sizeof(q->istats));
  else if (q->id == SRX_QUERY_PIPELINE_PROFILES && size >= 8 * sizeof(uint64_t))
    std::memcpy(data, q->prof, std::min(size, sizeof(q->prof)));
  else return SRX_ASYNC_ERROR;
  return SRX_ASYNC_READY;
}
int32_t srx_draw(srx_renderer rr, size_t start, size_t primcnt) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  (void)start;  // the index fetcher ignores the start of non-indexed draws (index_fetcher.cpp:103-107)
  return execute_draw(r, false, 0, static_cast<uint32_t>(primcnt), 0);
}
int32_t srx_draw_index(srx_renderer rr, size_t start, size_t primcnt, int32_t basevert) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  return execute_draw(r, true, static_cast<uint32_t>(start), static_cast<uint32_t>(primcnt), basevert);
}
int32_t srx_clear_color(srx_renderer rr, srx_surface s, const float* c) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* o = as<surface_obj>(s, K_SURFACE);
  if (!r || !o || !c) return SRX_INVALID_PARAMETER;
  if (r->recording) return record_fill(r, o, make_float4(c[0], c[1], c[2], c[3]));
  launch_clear(o->desc(), make_float4(c[0], c[1], c[2], c[3]), r->stream);
  HIPCHK(hipGetLastError());
  return SRX_OK;
}
int32_t srx_clear_depth_stencil(srx_renderer rr, srx_surface s, uint32_t flags, float d, uint32_t st) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* o = as<surface_obj>(s, K_SURFACE);
  if (!r || !o || o->pf != PF_RG32F) return SRX_INVALID_PARAMETER;
  if (flags != (SRX_CLEAR_DEPTH | SRX_CLEAR_STENCIL) && flags != SRX_CLEAR_DEPTH && flags != SRX_CLEAR_STENCIL)
    return SRX_E_UNSUPPORTED;
  if (r->recording) {
    if (flags != (SRX_CLEAR_DEPTH | SRX_CLEAR_STENCIL))
      return r->recording->refuse(SRX_E_UNSUPPORTED, "batched depth/stencil clears set both");
    float sf;
    std::memcpy(&sf, &st, sizeof sf);
    return record_fill(r, o, make_float4(d, sf, 0.0f, 0.0f));
  }
  launch_clear_ds(o->desc(), flags, d, st, r->stream);
  HIPCHK(hipGetLastError());
  return SRX_OK;
}
int32_t srx_batch_begin(srx_renderer rr) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || r->recording) return SRX_INVALID_PARAMETER;
  r->recording = new batch_obj();
  return SRX_OK;
}
int32_t srx_batch_end(srx_renderer rr, srx_batch* out) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !out || !r->recording) return SRX_INVALID_PARAMETER;
  batch_obj* b = r->recording;
  r->recording = nullptr;
  if (b->error != SRX_OK) {
    const int32_t e = b->error;
    release(b);
    return e;
  }
  retain(r);
  b->r = r;
  b->owns_renderer = true;
  int32_t rc = batch_prepare(b);
  if (rc == SRX_OK) rc = batch_configure(b, b->tile, b->bin_pt, b->vis_threads);
  if (rc != SRX_OK) {
    release(b);
    return rc;
  }
  *out = reinterpret_cast<srx_batch>(static_cast<object*>(b));
  return SRX_OK;
}
int32_t srx_batch_launch(srx_batch bb) {
  auto* b = as<batch_obj>(bb, K_BATCH);
  if (!b || !b->exec) return SRX_INVALID_PARAMETER;
  HIPCHK(hipGraphLaunch(b->exec, b->r->stream));
  return SRX_OK;
}
int32_t srx_batch_status(srx_batch bb, uint32_t* status) {
  auto* b = as<batch_obj>(bb, K_BATCH);
  if (!b || !status) return SRX_INVALID_PARAMETER;
  HIPCHK(hipMemcpyAsync(b->status_h, b->status_d, sizeof(uint32_t), hipMemcpyDeviceToHost, b->r->stream));
  HIPCHK(hipStreamSynchronize(b->r->stream));
  *status = b->status_h[0];
  return SRX_OK;
}
int32_t srx_batch_configure(srx_batch bb, int32_t tile, int32_t bin_items, int32_t vis_threads) {
  auto* b = as<batch_obj>(bb, K_BATCH);
  if (!b || bin_items % 256) return SRX_INVALID_PARAMETER;
  return batch_configure(b, tile, bin_items / 256, vis_threads);
}
int32_t srx_batch_autotune(srx_batch bb, uint32_t reps, srx_batch_tuning* results, size_t capacity, size_t* count) {
  auto* b = as<batch_obj>(bb, K_BATCH);
  if (!b || reps == 0) return SRX_INVALID_PARAMETER;
  hipStream_t s = b->r->stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  size_t n = 0;
  float best = 0.0f;
  int best_tile = b->tile, best_pt = b->bin_pt, best_th = b->vis_threads;
  int32_t rc = SRX_OK;
  for (int tile : {8, 16, 32}) {
    for (int pt : {8, 16, 32}) {
      for (int th : {64, 128, 256}) {
        rc = batch_configure(b, tile, pt, th);
        if (rc == SRX_E_UNSUPPORTED) {  // a tile size some job cannot take: not a candidate
          rc = SRX_OK;
          continue;
        }
        if (rc != SRX_OK) break;
        hipError_t e = hipGraphLaunch(b->exec, s);  // warm-up
        if (e == hipSuccess) e = hipEventRecord(e0, s);
        for (uint32_t i = 0; i < reps && e == hipSuccess; ++i) e = hipGraphLaunch(b->exec, s);
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
          rc = hip_fail(e, "srx_batch_autotune");
          break;
        }
        ms /= static_cast<float>(reps);
        if (results && n < capacity) results[n] = srx_batch_tuning{tile, pt * 256, th, ms};
        ++n;
        if (n == 1 || ms < best) {
          best = ms;
          best_tile = tile;
          best_pt = pt;
          best_th = th;
        }
      }
      if (rc != SRX_OK) break;
    }
    if (rc != SRX_OK) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (count) *count = n;
  if (rc != SRX_OK) return rc;
  return batch_configure(b, best_tile, best_pt, best_th);
}
int32_t srx_batch_get_info(srx_batch bb, srx_batch_info* info) {
  auto* b = as<batch_obj>(bb, K_BATCH);
  if (!b || !info) return SRX_INVALID_PARAMETER;
  info->jobs = static_cast<uint32_t>(b->draws.size());
  info->clears = static_cast<uint32_t>(b->fills.size());
  info->resolves = static_cast<uint32_t>(b->resolves.size());
  info->tile = static_cast<uint32_t>(b->tile);
  info->bin_items = static_cast<uint32_t>(b->bin_pt * 256);
  info->vis_threads = static_cast<uint32_t>(b->vis_threads);
  info->vis_launches = static_cast<uint32_t>(b->m_vis.size());
  info->primitives = b->primitives;
  info->records = b->capacity;
  info->device_bytes = b->device_bytes;
  return SRX_OK;
}
void srx_default_server_desc(srx_server_desc* d) {
  *d = srx_server_desc{200ull << 30, 64ull << 30, 2ull << 30, 64, 3, 8192, 64};
}
int32_t srx_create_server(srx_renderer rr, const srx_server_desc* d, srx_server* out) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !d || !out || d->streams == 0 || d->streams > 8 || d->max_batch_jobs == 0 || !batch_bin_items_valid(d->bin_items / 256) ||
      d->bin_items % 256 || !(d->vis_threads == 64 || d->vis_threads == 128 || d->vis_threads == 256))
    return SRX_INVALID_PARAMETER;
  auto* S = new server_obj();
  S->desc = *d;
  retain(r);
  S->r = r;
  hipError_t e = hipMalloc(&S->dev, std::max<uint64_t>(d->arena_bytes, 256));
  if (e == hipSuccess) e = hipHostMalloc(&S->host, std::max<uint64_t>(d->staging_bytes, 256), hipHostMallocDefault);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&S->upload, hipStreamNonBlocking);
  for (uint32_t i = 0; i < d->streams && e == hipSuccess; ++i) {
    hipStream_t st = nullptr;
    e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) S->streams.push_back(st);
  }
  if (e != hipSuccess) {
    const int32_t rc = hip_fail(e, "srx_create_server");
    release(S);
    return rc;
  }
  S->stream_busy.assign(S->streams.size(), false);
  S->dev_arena.init(S->dev, d->arena_bytes);
  S->host_arena.init(S->host, d->staging_bytes);
  S->worker = std::thread(serve_loop, S);
  *out = reinterpret_cast<srx_server>(static_cast<object*>(S));
  return SRX_OK;
}
int32_t srx_server_put_mesh(srx_server ss, uint64_t key, const void* vertices, uint32_t n_tris) {
  auto* S = as<server_obj>(ss, K_SERVER);
  if (!S || n_tris == 0 || n_tris >= 0x15555555u) return SRX_INVALID_PARAMETER;
  std::lock_guard<std::mutex> lk(S->mu);
  return put_mesh(S, key, vertices, n_tris);
}
int32_t srx_server_put_texture(srx_server ss, uint64_t key, const void* rgba8, uint32_t size) {
  auto* S = as<server_obj>(ss, K_SERVER);
  if (!S || key == 0 || size == 0 || size > 16384 || (size & (size - 1))) return SRX_INVALID_PARAMETER;
  std::lock_guard<std::mutex> lk(S->mu);
  return put_texture(S, key, rgba8, size);
}
int32_t srx_server_submit(srx_server ss, const srx_request* q, uint64_t* ticket) {
  auto* S = as<server_obj>(ss, K_SERVER);
  if (!S || !q || !ticket) return SRX_INVALID_PARAMETER;
  if (!(q->samples == 1 || q->samples == 2 || q->samples == 4) || q->width == 0 || q->height == 0 || q->n_tris == 0 ||
      q->n_tris >= 0x15555555u)
    return SRX_INVALID_PARAMETER;
  if (((q->width + 15) / 16) * ((q->height + 15) / 16) > BATCH_MAX_TILES)
    return fail(SRX_E_UNSUPPORTED, "served target larger than 8192 16x16 tiles");
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->fatal != SRX_OK) return fail(S->fatal, S->fatal_msg.c_str());
  // each asset is held as soon as it is resident, so the next upload cannot evict it
  RCHK(put_mesh(S, q->mesh_key, q->mesh_data, q->n_tris));
  ++S->meshes[q->mesh_key].users;
  if (q->texture_key) {
    const int32_t rc = put_texture(S, q->texture_key, q->texture_data, q->texture_size);
    if (rc != SRX_OK) {
      --S->meshes[q->mesh_key].users;
      return rc;
    }
    ++S->textures[q->texture_key].users;
  }
  srv_job t;
  t.ticket = S->next_ticket++;
  t.req = *q;
  t.req.mesh_data = nullptr;
  t.req.texture_data = nullptr;
  t.est_records = estimate_records(S, *q, &t.exact);
  auto pos = S->queue.end();  // ordered by arrival time, ties in submission order
  while (pos != S->queue.begin() && std::prev(pos)->req.arrival_ns > q->arrival_ns) --pos;
  S->queue.insert(pos, t);
  *ticket = t.ticket;
  S->cv.notify_all();
  return SRX_OK;
}
int32_t srx_server_poll(srx_server ss, srx_completion* out, size_t capacity, size_t* count, int32_t wait_ms) {
  auto* S = as<server_obj>(ss, K_SERVER);
  if (!S || !count || (capacity && !out)) return SRX_INVALID_PARAMETER;
  std::unique_lock<std::mutex> lk(S->mu);
  if (S->unreported.empty() && wait_ms > 0 && S->fatal == SRX_OK)
    S->cv.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !S->unreported.empty() || S->stop; });
  size_t n = 0;
  while (n < capacity && !S->unreported.empty()) {
    const uint64_t tk = S->unreported.front();
    S->unreported.erase(S->unreported.begin());
    const srv_job& t = S->finished[tk];
    srx_completion& c = out[n++];
    c.ticket = tk;
    c.status = t.status;
    c.batch_jobs = t.batch_jobs;
    c.arrival_ns = t.req.arrival_ns;
    c.admitted_ns = t.admitted_ns;
    c.done_ns = t.done_ns;
    c.color = t.stage_bytes ? S->host + t.stage_off : nullptr;
    c.color_bytes = t.stage_bytes ? t.color_bytes : 0;
    c.depth = t.depth_bytes && t.stage_bytes ? S->host + t.stage_off + range_arena::round(t.color_bytes) : nullptr;
    c.depth_bytes = c.depth ? t.depth_bytes : 0;
  }
  *count = n;
  if (S->fatal != SRX_OK) return fail(S->fatal, S->fatal_msg.c_str());
  return SRX_OK;
}
int32_t srx_server_release(srx_server ss, uint64_t ticket) {
  auto* S = as<server_obj>(ss, K_SERVER);
  if (!S) return SRX_INVALID_PARAMETER;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->finished.find(ticket);
  if (it == S->finished.end()) return SRX_INVALID_PARAMETER;
  if (it->second.stage_bytes) S->host_arena.release(it->second.stage_off, it->second.stage_bytes);
  S->finished.erase(it);
  S->cv.notify_all();
  return SRX_OK;
}
int64_t srx_server_now(srx_server ss) {
  auto* S = as<server_obj>(ss, K_SERVER);
  return S ? S->now() : -1;
}
int32_t srx_server_get_stats(srx_server ss, srx_server_stats* st) {
  auto* S = as<server_obj>(ss, K_SERVER);
  if (!S || !st) return SRX_INVALID_PARAMETER;
  std::lock_guard<std::mutex> lk(S->mu);
  *st = S->stats;
  st->pool_bytes = S->pool_used;
  st->resident_meshes = S->meshes.size();
  st->resident_textures = S->textures.size();
  st->arena_peak_bytes = S->dev_arena.peak;
  st->queued = S->queue.size();
  st->running = S->running.size();
  return SRX_OK;
}
int32_t srx_flush(srx_renderer rr) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  HIPCHK(hipStreamSynchronize(r->stream));
  return SRX_OK;
}

}  // extern "C"
