// Internals of the HIP backend's host runtime shared by srx_runtime.cpp (renderer, resources, draws),
// srx_batch.cpp (recorded batches) and srx_server.cpp (request server).
#pragma once

#include "../../include/srx.h"
#include "srx_kernels.h"
#include "srx_sasl.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace srx_rt {
using namespace srx;


inline thread_local int32_t g_last_error = 0;
inline thread_local std::string g_last_error_str;

inline int32_t fail(int32_t code, const char* msg) {
  g_last_error = 0;
  g_last_error_str = msg;
  return code;
}
inline int32_t hip_fail(hipError_t e, const char* where) {
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
enum obj_kind { K_RENDERER = 1, K_BUFFER, K_TEXTURE, K_SURFACE, K_SAMPLER, K_LAYOUT, K_RASTER, K_DS, K_QUERY, K_BATCH, K_SERVER,
                K_SHADER, K_SWAP_CHAIN };

struct object {
  std::atomic<int> refs{1};
  int kind;
  explicit object(int k) : kind(k) {}
  virtual ~object() = default;
};
inline void retain(object* o) {
  if (o) o->refs.fetch_add(1);
}
inline void release(object* o) {
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

inline int pf_bytes(int pf) {
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

struct renderer_obj;
// A full clear issued by srx_clear_color / srx_clear_depth_stencil that has not run yet: it runs
// right before the next command that reads or writes the surface (a draw queues its targets'
// clears behind its record emission, where the GPU would otherwise wait for the record count).
struct pending_clear {
  int kind = 0;  // 0 none, 1 colour, 2 depth/stencil
  float4 color{};
  uint32_t flags = 0;  // depth/stencil: SRX_CLEAR_DEPTH | SRX_CLEAR_STENCIL
  float depth = 0.0f;
  uint32_t stencil = 0;
  renderer_obj* owner = nullptr;  // the issuing renderer (its stream runs the clear without a requester)
};

struct surface_obj : object {
  int w, h, samples, pf;
  uint8_t* d = nullptr;
  size_t bytes = 0;
  pending_clear pend;
  surface_obj(int w_, int h_, int s_, int pf_) : object(K_SURFACE), w(w_), h(h_), samples(s_), pf(pf_) {}
  ~surface_obj() override {
    if (d) (void)hipFree(d);
  }
  surface_desc desc() const { return surface_desc{d, w, h, samples, pf}; }
};

inline int32_t make_surface(int w, int h, int s, int pf, surface_obj** out) {
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

// A compiled SASL pixel shader (shader_object, salvia/include/salvia/shader/shader_object.h).
struct shader_obj : object {
  sasl_module mod;
  shader_obj() : object(K_SHADER) {}
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

// HIP events bracketing the stages of one profiled draw, on the renderer's stream:
// [0,1) vertex shader, [1,2) setup + clipping (viewport transform fused), [2,3) compaction scan,
// [3,4) record emission + tile sort, [4,5) visibility or in-order raster, [5,6) deferred shading.
struct draw_events {
  hipEvent_t ev[7];
};

// An async_object (async_object.h:17-233).  begin/end enqueue copies of the device counters into
// the query's pinned slots and end records `done`, so neither waits; get_data polls or waits for
// `done` (async_object::get, :62-81) and then folds host counters, device counters and the stage
// events of the profiled draws into the result once.
struct query_obj : object {
  uint32_t id;
  bool running = false, has_data = false, resolved = false;
  hipEvent_t done = nullptr;
  unsigned long long* dev_snap = nullptr;  // pinned: [0, ST_COUNT) at begin, [ST_COUNT, 2*ST_COUNT) at end
  uint64_t host_begin[ST_COUNT] = {}, host_end[ST_COUNT] = {};
  std::vector<draw_events> draws;  // pipeline_profiles: the draws between begin and end
  srx_pipeline_statistics stats{};
  srx_internal_statistics istats{};
  uint64_t prof[PROF_COUNT] = {};
  query_obj() : object(K_QUERY) {}
  ~query_obj() override {
    for (auto& d : draws)
      for (auto e : d.ev)
        if (e) (void)hipEventDestroy(e);
    if (done) (void)hipEventDestroy(done);
    if (dev_snap) (void)hipHostFree(dev_snap);
  }
};

inline std::string lower(const char* s) {
  std::string r(s ? s : "");
  for (auto& c : r)
    if ('A' <= c && c <= 'Z') c = static_cast<char>(c - 'A' + 'a');
  return r;
}
// semantic_value name normalisation (shader/constants.h:52-83)
inline std::string canon_sem(const std::string& l) {
  if (l == "position" || l == "sv_position") return "position";
  if (l == "color" || l == "sv_target") return "color";
  return l;
}

struct vs_binding {
  const char* sem;
  uint32_t index;
  int reg;
};
// Input registers of each built-in vertex shader (bind_semantic of the cpp shaders; the SASL
// shaders' input structs in declaration order).
inline int vs_bindings(int id, vs_binding* out) {
  switch (id) {
  case SRX_VS_PASSTHROUGH_COLOR:
  case SRX_VS_WVP_COLOR:
  case SRX_VS_WVP_COLOR_LIGHT:
    out[0] = {"position", 0, 0};
    out[1] = {"color", 0, 1};
    return 2;
  case SRX_VS_WVP_NORMAL_UV:
  case SRX_VS_SSM_DRAW:
  case SRX_VS_CUP:
    out[0] = {"position", 0, 0};
    out[1] = {"normal", 0, 1};
    out[2] = {"texcoord", 0, 2};
    return 3;
  case SRX_VS_COLORIZED:
  case SRX_VS_MESH:
    out[0] = {"position", 0, 0};
    out[1] = {"normal", 0, 1};
    return 2;
  case SRX_VS_BOX:
    out[0] = {"position", 0, 0};
    out[1] = {"texcoord", 0, 1};
    return 2;
  }
  return -1;
}
inline int vs_output_count(int id) {
  switch (id) {
  case SRX_VS_PASSTHROUGH_COLOR:
  case SRX_VS_WVP_COLOR: return 1;
  case SRX_VS_WVP_NORMAL_UV:
  case SRX_VS_WVP_COLOR_LIGHT:
  case SRX_VS_BOX: return 2;
  case SRX_VS_MESH: return 3;
  case SRX_VS_COLORIZED:
  case SRX_VS_CUP: return 4;
  case SRX_VS_SSM_DRAW: return 5;
  }
  return 0;
}
// Bytes of the constant blob each shader reads (srx.h lists the layouts).
inline size_t vs_const_bytes(int id) {
  switch (id) {
  case SRX_VS_WVP_NORMAL_UV:
  case SRX_VS_WVP_COLOR:
  case SRX_VS_BOX: return 64;
  case SRX_VS_CUP: return 96;
  case SRX_VS_COLORIZED: return 112;
  case SRX_VS_WVP_COLOR_LIGHT: return 128;
  case SRX_VS_MESH:
  case SRX_VS_SSM_DRAW: return 160;
  }
  return 0;
}
inline size_t ps_const_bytes(int id) {
  switch (id) {
  case SRX_PS_LAMBERT_TEX:
  case SRX_PS_CONSTANT:
  case SRX_PS_SHADOW_PCF: return 16;
  case SRX_PS_CUP: return 48;
  case SRX_PS_SSM: return 52;
  }
  return 0;
}
inline uint32_t fmt_bytes(uint32_t fmt) {
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
inline void make_ewa(float* t) {
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
  float vs_consts[VS_CONSTS] = {}, ps_consts[16] = {};
  ref<sampler_obj> samplers[SRX_MAX_SAMPLER_SLOTS];
  ref<sampler_obj> vs_samplers[SRX_MAX_SAMPLER_SLOTS];  // set_vs_sampler: no built-in vertex shader samples
  srx_raster_desc raster;
  srx_depth_stencil_desc ds;
  ref<raster_obj> raster_state;  // the objects behind raster / ds, for the getters
  ref<ds_obj> ds_state;
  int32_t stencil_ref = 0;
  std::vector<ref<surface_obj>> color_targets;
  ref<surface_obj> ds_target;
  srx_viewport target_vp{0, 0, 0, 0, 0, 0};
  size_t target_samples = 0;
  srx_viewport vp{0, 0, 0, 0, 0, 1};
  // set_pixel_shader_code (renderer_impl.cpp:310-316): when set it shades instead of ps_id
  // (rasterizer.cpp:1402-1407); its variables and samplers are the renderer's constant buffer
  // (shader_cbuffer, set by name and kept across shaders, render_core.cpp:80-93)
  ref<shader_obj> ps_code;
  std::vector<std::pair<std::string, std::vector<uint8_t>>> ps_vars;
  std::vector<std::pair<std::string, ref<sampler_obj>>> ps_named_samplers;
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
  int device = 0;                  // current at creation: the stream's device
  hipStream_t stream = nullptr;
  std::vector<ref<query_obj>> active;
  // per-draw device scratch, grown on demand
  dev_array<float4> vbuf, tri_rec, tri_attr, clip_verts;
  dev_array<ps_samplers> samplers_d;
  ps_samplers* samplers_h = nullptr;  // pinned staging of the per-draw sampler table
  // the table last uploaded to samplers_d by an allocating draw (the no-allocation region is shared)
  ps_samplers samplers_last{};
  const ps_samplers* samplers_last_d = nullptr;
  bool samplers_valid = false;
  // the SASL program of the draw being built (uniforms applied), its device copy, pinned staging
  // and the program last uploaded by an allocating draw
  sasl_program sasl_cur{};
  sasl_program* sasl_d = nullptr;
  sasl_program* sasl_h = nullptr;
  sasl_program sasl_last{};
  bool sasl_valid = false;
  dev_array<uint32_t> cnt, keys0, keys1, vals0, vals1, tstart, tend, hist, scan_tmp, clip_hdr, counters, vis;
  dev_array<uint32_t> order;  // GPU tile scheduling order
  dev_array<uint32_t> tile_count;  // per-tile record counts and cursors of the counting sort
  dev_array<uint32_t> wide;   // triangles counted by wave
  dev_array<uint32_t> bsum;   // records per 256-primitive block
  dev_array<uint4> pairs;     // setup's first four records of each primitive
  dev_array<uint4> vis_entries, vis_merges;  // visibility work table
  dev_array<uint32_t> vis_counts;
  dev_array<unsigned long long> vis_partial;
  dev_array<uint8_t> clip_flag;
  dev_array<int32_t> clip_slot;
  dev_array<unsigned long long> partial;
  unsigned long long* stats_d = nullptr;
  uint32_t* pinned = nullptr;  // host staging for counter read-back (CT_COUNT words)
  std::vector<ref<surface_obj>> pending;  // surfaces holding a pending clear issued here
  // the counters are read back on `side` (default priority) while the main stream emits records
  hipStream_t side = nullptr;
  hipEvent_t ev_counted = nullptr, ev_read = nullptr;
  // binned records of the last draws; their maximum sizes the next draw's record buffer, so that
  // frames whose draws differ in size do not overflow it in turn
  uint32_t recent_records[8] = {};
  uint32_t recent_next = 0;
  uint64_t host_stats[ST_COUNT] = {};
  // map/unmap (resource_manager.cpp:7-113): one mapped resource at a time, staged in pinned memory
  int32_t map_mode = 0;
  ref<buffer_obj> mapped_buf;
  ref<surface_obj> mapped_surf;
  uint8_t* map_host = nullptr;
  size_t map_cap = 0;
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

inline std::once_flag g_ewa_once;
inline int32_t g_ewa_status = SRX_OK;

inline int gpu_tile_size() {
  static int g = [] {
    const char* e = std::getenv("SRX_GPU_TILE");
    int v = e ? std::atoi(e) : 16;
    return (v == 32) ? 32 : 16;
  }();
  return g;
}

// Starts the read-back of the draw's device counters as they stood at ev_counted (recorded on the
// main stream), on the side stream, so that work queued after ev_counted overlaps it.
inline int32_t start_counter_read(renderer_obj* r) {
  HIPCHK(hipStreamWaitEvent(r->side, r->ev_counted, 0));
  HIPCHK(hipMemcpyAsync(r->pinned, r->counters.p, CT_CLASS * sizeof(uint32_t), hipMemcpyDeviceToHost, r->side));
  HIPCHK(hipEventRecord(r->ev_read, r->side));
  return SRX_OK;
}
inline int32_t finish_counter_read(renderer_obj* r, uint32_t* out) {
  HIPCHK(hipEventSynchronize(r->ev_read));
  for (int i = 0; i < CT_COUNT; ++i) out[i] = i < CT_CLASS ? r->pinned[i] : 0u;
  return SRX_OK;
}

#define RCHK(expr)                     \
  do {                                 \
    int32_t _rc = (expr);              \
    if (_rc != SRX_OK) return _rc;     \
  } while (0)

template <class T>
T* as(void* h, int kind) {
  auto* o = static_cast<object*>(h);
  return (o && o->kind == kind) ? static_cast<T*>(o) : nullptr;
}

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


// ---------------------------------------------------------------- across translation units
// A caller-free device region that a draw carves its scratch from instead of allocating.
struct fixed_scratch {
  uint8_t* base;
  size_t bytes;
};
// srx_runtime.cpp
// One draw on the renderer's stream; fixed == nullptr: scratch from the renderer's grow-only arrays.
int32_t execute_draw(renderer_obj* r, bool indexed, uint32_t start, uint32_t primcnt, int32_t basevert,
                     const fixed_scratch* fixed);
constexpr size_t MAX_DRAW_PRIMITIVES = 0x55554000u;  // clipped-triangle ids (3 per primitive) are u32
int32_t build_draw(renderer_obj* r, bool indexed, uint32_t start, uint32_t primcnt, int32_t basevert, draw_params& P,
                   ps_samplers& SM, bool& skip);
// srx_batch.cpp: commands recorded between srx_batch_begin and srx_batch_end
int32_t record_draw(renderer_obj* r, const draw_params& P, const ps_samplers& SM);
// Queues o's pending clear on stream s (its issuing renderer's stream when s is null); drop: the
// surface is about to be overwritten whole, so the clear is discarded instead.
int32_t run_pending_clear(surface_obj* o, hipStream_t s, bool drop = false);
// Deferred SASL shading with the program compiled to native code (srx_sasl_jit.cpp); *launched is
// false when it did not launch (run-time compilation off with SRX_SASL_JIT=0, a failed compilation,
// or a module not yet compiled when allow_compile is false): the interpreter kernel shades then.
int32_t sasl_jit_launch(const sasl_program& prog, const draw_params& P, const draw_buffers& B, int device, hipStream_t s,
                        bool allow_compile, bool* launched);
std::string sasl_jit_source(const sasl_program& p, int NA, int S);
// Queues every pending clear renderer r issued, on r's stream.
int32_t run_pending_clears(renderer_obj* r);
int32_t record_fill(renderer_obj* r, surface_obj* o, float4 v);
int32_t record_resolve(renderer_obj* r, surface_obj* a, surface_obj* d);
int32_t batch_refuse(renderer_obj* r, int32_t code, const char* msg);  // the first recording error sticks
void drop_recording(renderer_obj* r);

}  // namespace srx_rt
